/*
 * oracle.c -- CPU oracle for the UVaTRMM local-product path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built by oracle/Makefile into
 * oracle/liboracle.so; loaded by tests/, the smoke check and the benchmark's
 * CPU-baseline leg, never by the product.
 *
 * Pinning (what makes each function trustworthy):
 *   orc_map_regular / orc_map_balanced
 *       golden vectors produced by the reference's own map_* functions
 *       (tests/golden/mapping_golden.json) and its abort cases.
 *   orc_fill_closed_A / orc_fill_closed_B
 *       the reference program's printed Norm values (tests/golden/
 *       reference_norms.json) reproduced through orc_trmm + orc_inf_norm,
 *       and element-level closed-form checks.
 *   orc_seeded_value / orc_fill_seeded_*
 *       project-defined generator; pinned by published SplitMix64 outputs
 *       (known-answer values of the mixing function) and value-range checks.
 *   orc_trmm / orc_trmm_rows_dd / orc_stage_loop
 *       MKL cblas_dtrmm/cblas_dgemm fixtures (the reference's arithmetic
 *       engine), OpenBLAS (numpy) cross-checks, brute force in exact
 *       rational arithmetic on tiny inputs, and identities (A = I gives B,
 *       A = all-ones lower gives prefix sums of B, linearity).
 *   orc_inf_norm
 *       definition; checked against numpy's norm(C, inf).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* A1 mapping                                                          */
/* ------------------------------------------------------------------ */

/* map_regular, UVaTRMM/src/mm_mapping.c:27-67 */
int orc_map_regular(uint64_t n, uint32_t p, uint32_t i, uint64_t *begin, uint64_t *size)
{
    unsigned long equitative_size = n / p;
    unsigned long remainder = n % p;
    if (equitative_size < 1) equitative_size = 1;
    if (p != 1 && equitative_size * p > n) return ORC_EXIT_INVALID_PROCS_NUM; /* :40-54 */
    *begin = (unsigned long)i * equitative_size;                            /* :56 */
    *size = equitative_size;
    if (i == p - 1) *size += remainder;                                      /* :59-61 */
    return 0;
}

/* get_triangle_num_elements, mm_mapping.c:120-126.  The reference stores
 * 0.5*(r*(r+1)) in a 32-bit unsigned int; beyond UINT32_MAX the double to
 * unsigned conversion is undefined in C, so such n are reported undefined. */
static int orc_triangle_elems(unsigned long r, unsigned long *out)
{
    double t = 0.5 * (double)(r * (r + 1));
    if (t > 4294967295.0) return ORC_EXIT_UNDEFINED;
    unsigned int num_elements = (unsigned int)t;
    *out = num_elements;
    return 0;
}

/* get_rows_before, mm_mapping.c:177-181 (ideal is passed as a double) */
static unsigned long orc_rows_before(double ideal_block_size, int block_index)
{
    double numerator = -1 + sqrt(1 + 8 * (block_index * ideal_block_size));
    unsigned long rows_before = (numerator > 0) ? (unsigned long)(floor(numerator) / 2) : 0;
    return rows_before;
}

/* map_balanced, mm_mapping.c:319-366 with USE_OLD_BALANCED=0
 * (include/mm_triangular.h:54). */
int orc_map_balanced(uint64_t n, uint32_t p, uint32_t i, uint64_t *begin, uint64_t *size)
{
    unsigned long elems;
    /* get_matrix_num_elements_under_diagonal(n, n): square -> triangle only (:132-147) */
    if (orc_triangle_elems(n, &elems)) return ORC_EXIT_UNDEFINED;
    /* get_ideal_block_size (:153-168) */
    double ideal_block_size = (double)elems / (double)p;
    if (ideal_block_size < 1) return ORC_EXIT_INVALID_BLOCKS_NUM;
    unsigned int ideal = (unsigned int)ceil(ideal_block_size); /* :330 */
    unsigned int rows_begin = (unsigned int)orc_rows_before((double)ideal, (int)i);
    unsigned int rows_end = (unsigned int)orc_rows_before((double)ideal, (int)i + 1);
    unsigned long b = rows_begin;
    unsigned long s = (unsigned int)(rows_end - rows_begin);
    if (i == 0) { /* :346-353 */
        b = 0;
        s = (p > 1) ? orc_rows_before((double)ideal, 1) : n;
    }
    if (n > 1 && i == p - 1) { /* :354-356 */
        if (b > n) return ORC_EXIT_UNDEFINED; /* unsigned wrap in the reference */
        s = n - b;
    }
    if (s == 0) return ORC_EXIT_INVALID_RANGE; /* check_range, :184-191 */
    *begin = b;
    *size = s;
    return 0;
}

/* ------------------------------------------------------------------ */
/* A4 inputs                                                           */
/* ------------------------------------------------------------------ */

/* mat_consecutive_values with MININUMS=1, mm_triangular_common.c:1782-1805,
 * called through matrix_rangeSetA (:1830-1840) for a band {b,h} x {0,n}:
 * value index (i_local*n + j + 0) + b*n = row*n + j, VALUES_LIMIT = 2341
 * (include/mm_triangular.h:66-68); zero for j > row (triangular). */
void orc_fill_closed_A(int64_t n, double *A, int64_t lda)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) {
            if (j > i) {
                A[i * lda + j] = 0.0;
            } else {
                int v = (int)((i * n + j) % 2341) + 1;
                double d = v;
                A[i * lda + j] = 1 / d;
            }
        }
    }
}

/* matrix_rangeSetB (:1846-1857) for rows {0,n} x columns {c0,w}: value index
 * i*n + j + c0 = row*n + global column. */
void orc_fill_closed_B(int64_t n, int64_t m, double *B, int64_t ldb)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < m; j++) {
            int v = (int)((i * n + j) % 2341) + 1;
            double d = v;
            B[i * ldb + j] = 1 / d;
        }
    }
}

/* SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014; constants of the
 * published public-domain splitmix64.c by S. Vigna). */
static uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_splitmix64(uint64_t state, uint64_t k)
{
    return orc_mix64(state + k * 0x9E3779B97F4A7C15ULL);
}

double orc_seeded_value(uint64_t seed, uint32_t stream, uint64_t idx, int spread, int diag)
{
    const uint64_t gamma = 0x9E3779B97F4A7C15ULL;
    uint64_t key = orc_mix64(seed ^ ((uint64_t)stream * gamma));
    uint64_t z = orc_mix64(key + (idx + 1) * gamma);
    double u = (double)(z >> 11) * 0x1.0p-53; /* [0,1), exact */
    double v = 2.0 * u - 1.0;                 /* [-1,1), exact */
    if (diag) v = v + 2.0;
    if (spread) {
        int e = (int)(orc_mix64(z) % 17) - 8;
        uint64_t bits = (uint64_t)(1023 + e) << 52;
        double s;
        memcpy(&s, &bits, sizeof s);
        v = v * s; /* exact power-of-two scaling */
    }
    return v;
}

void orc_fill_seeded_A(int64_t n, uint64_t seed, int spread, double *A, int64_t lda)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++)
            A[i * lda + j] =
                (j > i) ? 0.0 : orc_seeded_value(seed, 1u, (uint64_t)(i * n + j), spread, i == j);
}

void orc_fill_seeded_B(int64_t n, int64_t m, uint64_t seed, int spread, double *B, int64_t ldb)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < m; j++)
            B[i * ldb + j] = orc_seeded_value(seed, 2u, (uint64_t)(i * m + j), spread, 0);
}

/* ------------------------------------------------------------------ */
/* A5-A8 product                                                       */
/* ------------------------------------------------------------------ */

/* C = tril(A) * B.  Rows in blocks of 8 share each B row read; columns in
 * blocks of 512.  Per element the sum is sequential in ascending k with a
 * separate multiply and add (built with -ffp-contract=off). */
void orc_trmm(int64_t n, int64_t m, const double *A, int64_t lda, const double *B, int64_t ldb,
              double *C, int64_t ldc)
{
    const int64_t RB = 8, CB = 512;
    int64_t nrb = (n + RB - 1) / RB;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t rb = nrb - 1; rb >= 0; rb--) {
        int64_t i0 = rb * RB, i1 = i0 + RB < n ? i0 + RB : n;
        for (int64_t j0 = 0; j0 < m; j0 += CB) {
            int64_t j1 = j0 + CB < m ? j0 + CB : m;
            for (int64_t i = i0; i < i1; i++)
                for (int64_t j = j0; j < j1; j++) C[i * ldc + j] = 0.0;
            for (int64_t k = 0; k < i1; k++) {
                const double *b = B + k * ldb;
                for (int64_t i = (k > i0 ? k : i0); i < i1; i++) {
                    double a = A[i * lda + k];
                    double *c = C + i * ldc;
                    for (int64_t j = j0; j < j1; j++) c[j] += a * b[j];
                }
            }
        }
    }
}

/* Error-free transformations (Knuth TwoSum; TwoProd with a correctly
 * rounded fma). */
static inline void two_sum(double a, double b, double *s, double *e)
{
    double x = a + b, bv = x - a, av = x - bv;
    *s = x;
    *e = (a - av) + (b - bv);
}

void orc_trmm_rows_dd(int64_t n, int64_t m, const double *A, int64_t lda, const double *B,
                      int64_t ldb, const int64_t *rows, int64_t nrows, double *out, int64_t ldo)
{
    (void)n;
    const int64_t CB = 256;
    int64_t ncb = (m + CB - 1) / CB;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int64_t r = 0; r < nrows; r++) {
        for (int64_t cb = 0; cb < ncb; cb++) {
            double hi[256], lo[256];
            int64_t i = rows[r];
            int64_t j0 = cb * CB, j1 = j0 + CB < m ? j0 + CB : m;
            for (int64_t j = j0; j < j1; j++) hi[j - j0] = lo[j - j0] = 0.0;
            for (int64_t k = 0; k <= i; k++) {
                double a = A[i * lda + k];
                const double *b = B + k * ldb;
                for (int64_t j = j0; j < j1; j++) {
                    double p = a * b[j];
                    double pe = fma(a, b[j], -p);
                    double s, e;
                    two_sum(hi[j - j0], p, &s, &e);
                    hi[j - j0] = s;
                    lo[j - j0] += e + pe;
                }
            }
            for (int64_t j = j0; j < j1; j++) out[r * ldo + j] = hi[j - j0] + lo[j - j0];
        }
    }
}

/* Rows of C = tril(A) * B in double-double for the seeded inputs of an n x n
 * A and n x m B, generating every element on the fly (no n x n host copy):
 * the same sums as orc_trmm_rows_dd over orc_fill_seeded_A/B. */
void orc_seeded_rows_dd(int64_t n, int64_t m, uint64_t seed, int spread, const int64_t *rows,
                        int64_t nrows, double *out, int64_t ldo)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t r = 0; r < nrows; r++) {
        const int64_t i = rows[r];
        double *hi = (double *)calloc((size_t)m, sizeof(double));
        double *lo = (double *)calloc((size_t)m, sizeof(double));
        if (!hi || !lo) {
            free(hi);
            free(lo);
            for (int64_t j = 0; j < m; j++) out[r * ldo + j] = NAN;
            continue;
        }
        for (int64_t k = 0; k <= i; k++) {
            const double a = orc_seeded_value(seed, 1u, (uint64_t)(i * n + k), spread, i == k);
            for (int64_t j = 0; j < m; j++) {
                const double b = orc_seeded_value(seed, 2u, (uint64_t)(k * m + j), spread, 0);
                double p = a * b;
                double pe = fma(a, b, -p);
                double sum, e;
                two_sum(hi[j], p, &sum, &e);
                hi[j] = sum;
                lo[j] += e + pe;
            }
        }
        for (int64_t j = 0; j < m; j++) out[r * ldo + j] = hi[j] + lo[j];
        free(hi);
        free(lo);
    }
}

/* ---- host BLAS (MKL) loaded at run time, for the stage-loop replay ---- */
typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double *, int, const double *,
                         int, double, double *, int);
typedef void (*dtrmm_fn)(int, int, int, int, int, int, int, double, const double *, int, double *,
                         int);
static dgemm_fn g_dgemm;
static dtrmm_fn g_dtrmm;
static char g_blas_version[256];
static int g_blas_state; /* 0 unknown, 1 loaded, -1 missing */

static void orc_blas_load(void)
{
    if (g_blas_state) return;
    static const char *names[] = {"libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1",
                                  "/opt/conda/lib/libmkl_rt.so"};
    void *h = NULL;
    for (unsigned i = 0; i < sizeof names / sizeof names[0] && !h; i++)
        h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        g_blas_state = -1;
        return;
    }
    int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
    void (*ver)(char *, int) = (void (*)(char *, int))dlsym(h, "MKL_Get_Version_String");
    g_dgemm = (dgemm_fn)dlsym(h, "cblas_dgemm");
    g_dtrmm = (dtrmm_fn)dlsym(h, "cblas_dtrmm");
    if (!g_dgemm || !g_dtrmm) {
        g_blas_state = -1;
        return;
    }
    if (set_layer) set_layer(1); /* MKL_THREADING_SEQUENTIAL, like -lmkl_sequential */
    if (ver) ver(g_blas_version, (int)sizeof g_blas_version);
    g_blas_state = 1;
}

/* Makes the CBLAS library at `path` the one the stage loop calls (a
 * CBLAS-compatible replacement for MKL under test); NULL returns to the MKL
 * search.  1 when both routines were found. */
int orc_blas_use(const char *path)
{
    g_blas_state = 0;
    g_dgemm = NULL;
    g_dtrmm = NULL;
    if (!path) return 1;
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return 0;
    g_dgemm = (dgemm_fn)dlsym(h, "cblas_dgemm");
    g_dtrmm = (dtrmm_fn)dlsym(h, "cblas_dtrmm");
    if (!g_dgemm || !g_dtrmm) {
        g_blas_state = -1;
        return 0;
    }
    snprintf(g_blas_version, sizeof g_blas_version, "%s", path);
    g_blas_state = 1;
    return 1;
}

int orc_blas_available(char *version, int cap)
{
    orc_blas_load();
    if (version && cap > 0) {
        snprintf(version, (size_t)cap, "%s", g_blas_state == 1 ? g_blas_version : "");
    }
    return g_blas_state == 1;
}

/* CBLAS enum values (public CBLAS interface). */
enum { ROWMAJOR = 101, NOTRANS = 111, LOWER = 122, NONUNIT = 131, LEFT = 141 };

/* one band of the stage loop, UVaTRMM/src/mm_triangular.c:330-403 */
static void orc_stage(int64_t n, int64_t w, uint64_t b, uint64_t h, const double *A, int64_t lda,
                      const double *B, int64_t ldb, double *C, int64_t ldc, double *scratch,
                      int blas)
{
    /* K1, :347-360: C[b:b+h,:] = A[b:b+h, 0:b+1] * B[0:b+1,:] */
    if (blas) {
        g_dgemm(ROWMAJOR, NOTRANS, NOTRANS, (int)h, (int)w, (int)(b + 1), 1.0, A + b * lda,
                (int)lda, B, (int)ldb, 0.0, C + b * ldc, (int)ldc);
    } else {
        for (uint64_t r = b; r < b + h; r++) {
            double *c = C + r * ldc;
            for (int64_t j = 0; j < w; j++) c[j] = 0.0;
            for (uint64_t k = 0; k <= b; k++) {
                double a = A[r * lda + k];
                const double *bb = B + k * ldb;
                for (int64_t j = 0; j < w; j++) c[j] += a * bb[j];
            }
        }
    }
    /* K2 geometry, :366-373: tri = min(h-1, n), left = b+1, above = 1 */
    uint64_t tri = h - 1;
    if (tri > (uint64_t)n) tri = (uint64_t)n;
    if (tri == 0) return;
    uint64_t left = b + h - tri;
    /* K4 save (:334-344) becomes a scratch copy of B rows [left, left+tri). */
    for (uint64_t r = 0; r < tri; r++)
        memcpy(scratch + r * w, B + (left + r) * ldb, (size_t)w * sizeof(double));
    const double *At = A + (b + (h - tri)) * lda + left;
    if (blas) {
        g_dtrmm(ROWMAJOR, LEFT, LOWER, NOTRANS, NONUNIT, (int)tri, (int)w, 1.0, At, (int)lda,
                scratch, (int)w);
    } else {
        for (int64_t r = (int64_t)tri - 1; r >= 0; r--) { /* bottom-up keeps it in place */
            double *t = scratch + r * w;
            double d = At[r * lda + r];
            for (int64_t j = 0; j < w; j++) t[j] *= d;
            for (int64_t k = 0; k < r; k++) {
                double a = At[r * lda + k];
                const double *s = scratch + k * w;
                for (int64_t j = 0; j < w; j++) t[j] += a * s[j];
            }
        }
    }
    /* K3, :388-393: C[left + r] += B'[r] */
    for (uint64_t r = 0; r < tri; r++) {
        double *c = C + (left + r) * ldc;
        const double *t = scratch + r * w;
        for (int64_t j = 0; j < w; j++) c[j] += t[j];
    }
}

int orc_stage_loop(int64_t n, int64_t m, uint32_t p, int balanced, const double *A, int64_t lda,
                   const double *B, int64_t ldb, double *C, int64_t ldc, int blas)
{
    if (blas) {
        orc_blas_load();
        if (g_blas_state != 1) blas = 0;
    }
    uint64_t maxh = 0;
    for (uint32_t s = 0; s < p; s++) {
        uint64_t b, h;
        int rc = balanced ? orc_map_balanced((uint64_t)n, p, s, &b, &h)
                          : orc_map_regular((uint64_t)n, p, s, &b, &h);
        if (rc) return rc;
        if (h > maxh) maxh = h;
    }
    double *scratch = (double *)malloc((size_t)(maxh ? maxh : 1) * (size_t)m * sizeof(double));
    if (!scratch) return -2;
    for (uint32_t s = 0; s < p; s++) {
        uint64_t b = 0, h = 0; /* every band mapped without error above */
        if (balanced)
            orc_map_balanced((uint64_t)n, p, s, &b, &h);
        else
            orc_map_regular((uint64_t)n, p, s, &b, &h);
        orc_stage(n, m, b, h, A, lda, B, ldb, C, ldc, scratch, blas);
    }
    free(scratch);
    return 0;
}

double orc_ranks_stage_loop(int64_t n, int64_t w, uint32_t p, int balanced, int ranks,
                            const double *A, int64_t lda, const double *B, int64_t ldb,
                            double *C, int64_t ldc, int blas)
{
    if (blas) {
        orc_blas_load();
        if (g_blas_state != 1) blas = 0;
    }
    double t0 = 0, t1 = 0;
    int bad = 0;
#pragma omp parallel num_threads(ranks) reduction(| : bad)
    {
#pragma omp barrier
#pragma omp master
        t0 = omp_get_wtime();
#pragma omp barrier
        int r = omp_get_thread_num();
        bad |= orc_stage_loop(n, w, p, balanced, A, lda, B + (int64_t)r * w, ldb,
                              C + (int64_t)r * w, ldc, blas) != 0;
#pragma omp barrier
#pragma omp master
        t1 = omp_get_wtime();
    }
    return bad ? -1.0 : t1 - t0;
}

/* ------------------------------------------------------------------ */
/* A10 infinity norm, matrix_rangeInfinityNorm,                        */
/* mm_triangular_common.c:2936-2991 (RECTANGULAR branch, :2954-2958)  */
/* ------------------------------------------------------------------ */
double orc_inf_norm(int64_t n, int64_t m, const double *C, int64_t ldc, double *rowsums)
{
    double norm = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int64_t j = 0; j < m; j++) s += fabs(C[i * ldc + j]);
        if (rowsums) rowsums[i] = s;
        norm = s > norm ? s : norm;
    }
    return norm;
}
