/* uvt_cblas.c -- libuvt_cblas.so: the CBLAS host interface of include/uvt_cblas.h
 * on top of libuvt's device BLAS forms (uvt_dgemm, uvt_dtrmm).
 *
 * Per call: check the parameters (CBLAS numbering), reduce a column-major
 * call to the row-major one it equals (a column-major matrix is the
 * transpose of a row-major one), copy the operands into a grow-only device
 * staging buffer, run, and copy the result back.  An operand whose rows lie
 * back to back in host memory keeps that pitch on the device (one copy);
 * others get an even pitch (16-byte loads in the tile kernel) and go through
 * libuvt's pinned strided copy.  One device handle per process, created on
 * the first call; a mutex serializes calls. */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "uvt.h"
#include "uvt_cblas.h"

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static uvt_handle g_h;
static void *g_stage;
static size_t g_stage_bytes;

static void die(const char *fn, const char *what, uvt_status st)
{
    fprintf(stderr, "libuvt_cblas: %s: %s failed: %s\n", fn, what, uvt_status_string(st));
    fflush(stderr);
    abort();
}

static void bad_param(const char *fn, int pos)
{
    fprintf(stderr, "libuvt_cblas: parameter %d was incorrect on entry to %s.\n", pos, fn);
}

static int64_t even(int64_t x) { return x + (x & 1); }
/* device row pitch of a host operand of `cols` columns and leading dimension ld */
static int64_t pitch(int64_t cols, int64_t ld) { return ld == cols ? cols : even(cols); }
static size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }
static int64_t imax1(int64_t x) { return x > 1 ? x : 1; }

/* the handle and a staging buffer of at least `bytes` (lock held) */
static char *stage(const char *fn, size_t bytes)
{
    uvt_status st;
    if (!g_h) {
        const char *dev = getenv("UVT_CBLAS_DEVICE");
        st = uvt_create(dev ? atoi(dev) : 0, &g_h);
        if (st != UVT_OK) die(fn, "opening the GPU", st);
    }
    if (g_stage_bytes < bytes) {
        if (g_stage) uvt_free(g_h, g_stage);
        g_stage = NULL;
        g_stage_bytes = 0;
        st = uvt_malloc(g_h, bytes, &g_stage);
        if (st != UVT_OK) die(fn, "device staging allocation", st);
        g_stage_bytes = bytes;
    }
    return (char *)g_stage;
}

static void to_dev(const char *fn, double *d, int64_t dpitch, const double *h, int64_t ld,
                   int64_t rows, int64_t cols)
{
    uvt_status st = uvt_memcpy_2d(g_h, d, (size_t)dpitch * 8, h, (size_t)ld * 8,
                                  (size_t)cols * 8, (size_t)rows, 0);
    if (st != UVT_OK) die(fn, "copy to the GPU", st);
}

static void to_host(const char *fn, double *h, int64_t ld, const double *d, int64_t dpitch,
                    int64_t rows, int64_t cols)
{
    uvt_status st = uvt_memcpy_2d(g_h, h, (size_t)ld * 8, d, (size_t)dpitch * 8,
                                  (size_t)cols * 8, (size_t)rows, 1);
    if (st != UVT_OK) die(fn, "copy from the GPU", st);
}

static int trans_ok(int t) { return t == CblasNoTrans || t == CblasTrans || t == CblasConjTrans; }

void cblas_dgemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE transa, CBLAS_TRANSPOSE transb, int m, int n,
                 int k, double alpha, const double *a, int lda, const double *b, int ldb,
                 double beta, double *c, int ldc)
{
    static const char *fn = "cblas_dgemm";
    const int row = layout == CblasRowMajor;
    if (!row && layout != CblasColMajor) return bad_param(fn, 1);
    if (!trans_ok(transa)) return bad_param(fn, 2);
    if (!trans_ok(transb)) return bad_param(fn, 3);
    if (m < 0) return bad_param(fn, 4);
    if (n < 0) return bad_param(fn, 5);
    if (k < 0) return bad_param(fn, 6);
    const int ta = transa != CblasNoTrans, tb = transb != CblasNoTrans;
    if (lda < imax1(row ? (ta ? m : k) : (ta ? k : m))) return bad_param(fn, 9);
    if (ldb < imax1(row ? (tb ? k : n) : (tb ? n : k))) return bad_param(fn, 11);
    if (ldc < imax1(row ? n : m)) return bad_param(fn, 14);
    if (m == 0 || n == 0) return;

    /* column-major C = op(A) op(B) is row-major C^T = op(B)^T op(A)^T */
    int64_t M = m, N = n, K = k, la = lda, lb = ldb;
    const double *A = a, *B = b;
    int tA = ta, tB = tb;
    if (!row) {
        M = n;
        N = m;
        A = b;
        B = a;
        la = ldb;
        lb = lda;
        tA = tb;
        tB = ta;
    }
    const int read_ab = K > 0 && alpha != 0.0;
    const int64_t ra = tA ? K : M, ca = tA ? M : K, rb = tB ? N : K, cb = tB ? K : N;
    const int64_t pa = pitch(imax1(ca), la), pb = pitch(imax1(cb), lb), pc = pitch(N, ldc);
    const size_t abytes = read_ab ? up256((size_t)ra * pa * 8) : 0;
    const size_t bbytes = read_ab ? up256((size_t)rb * pb * 8) : 0;
    const size_t cbytes = up256((size_t)M * pc * 8);

    pthread_mutex_lock(&g_mu);
    char *d = stage(fn, abytes + bbytes + cbytes);
    double *dA = (double *)d, *dB = (double *)(d + abytes), *dC = (double *)(d + abytes + bbytes);
    if (read_ab) {
        to_dev(fn, dA, pa, A, la, ra, ca);
        to_dev(fn, dB, pb, B, lb, rb, cb);
    }
    if (beta != 0.0) to_dev(fn, dC, pc, c, ldc, M, N);
    uvt_status st = uvt_dgemm(g_h, tA ? UVT_TRANS : UVT_NO_TRANS, tB ? UVT_TRANS : UVT_NO_TRANS, M,
                              N, read_ab ? K : 0, alpha, dA, pa, dB, pb, beta, dC, pc, NULL);
    if (st != UVT_OK) die(fn, "uvt_dgemm", st);
    to_host(fn, c, ldc, dC, pc, M, N);
    pthread_mutex_unlock(&g_mu);
}

void cblas_dtrmm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE transa,
                 CBLAS_DIAG diag, int m, int n, double alpha, const double *a, int lda, double *b,
                 int ldb)
{
    static const char *fn = "cblas_dtrmm";
    const int row = layout == CblasRowMajor;
    if (!row && layout != CblasColMajor) return bad_param(fn, 1);
    if (side != CblasLeft && side != CblasRight) return bad_param(fn, 2);
    if (uplo != CblasUpper && uplo != CblasLower) return bad_param(fn, 3);
    if (!trans_ok(transa)) return bad_param(fn, 4);
    if (diag != CblasNonUnit && diag != CblasUnit) return bad_param(fn, 5);
    if (m < 0) return bad_param(fn, 6);
    if (n < 0) return bad_param(fn, 7);
    if (lda < imax1(side == CblasLeft ? m : n)) return bad_param(fn, 10);
    if (ldb < imax1(row ? n : m)) return bad_param(fn, 12);
    if (m == 0 || n == 0) return;

    /* column-major: B^T := op(A)^T B^T with A^T stored row-major, i.e. the
     * other side and the other triangle, m and n swapped */
    int left = side == CblasLeft, lower = uplo == CblasLower;
    int64_t M = m, N = n;
    if (!row) {
        left = !left;
        lower = !lower;
        M = n;
        N = m;
    }
    const int64_t nt = left ? M : N;
    const int64_t pa = pitch(nt, lda), pb = pitch(N, ldb);
    const size_t abytes = alpha != 0.0 ? up256((size_t)nt * pa * 8) : 0;
    const size_t bbytes = up256((size_t)M * pb * 8);

    pthread_mutex_lock(&g_mu);
    char *d = stage(fn, abytes + bbytes);
    double *dA = (double *)d, *dB = (double *)(d + abytes);
    if (alpha != 0.0) {
        to_dev(fn, dA, pa, a, lda, nt, nt);
        to_dev(fn, dB, pb, b, ldb, M, N);
    }
    uvt_status st = uvt_dtrmm(g_h, left ? UVT_LEFT : UVT_RIGHT, lower ? UVT_LOWER : UVT_UPPER,
                              transa != CblasNoTrans ? UVT_TRANS : UVT_NO_TRANS,
                              diag == CblasUnit ? UVT_UNIT : UVT_NON_UNIT, M, N, alpha,
                              alpha != 0.0 ? dA : NULL, pa, dB, pb, NULL);
    if (st != UVT_OK) die(fn, "uvt_dtrmm", st);
    to_host(fn, b, ldb, dB, pb, M, N);
    pthread_mutex_unlock(&g_mu);
}
