/*
 * oracle.h -- CPU oracle for the UVaTRMM local-product path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libuvt, uvatrmm.x,
 * libuvt_cblas) links, loads or calls this code.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark use it, and
 * only as the checker / the CPU comparator, never as the thing measured
 * on the GPU.
 *
 * Every function restates the reference algorithm it follows (citations
 * are relative to the UVaTRMM source tree) or is a plain mathematical
 * definition (C = tril(A) * B, row sums, max).  How each function is pinned
 * (golden vectors, known answers, independent libraries, identities) is
 * listed next to it.
 */
#ifndef UVT_ORACLE_H
#define UVT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference process exit codes: error_mpi_abort exits with code+2
 * (UVaTRMM/src/mm_triangular_common.c:2854-2863; enum in include/error.h:20-39). */
#define ORC_EXIT_INVALID_PROCS_NUM 2  /* INVALID_PROCS_NUM(0)+2 */
#define ORC_EXIT_INVALID_ARGS 3       /* INVALID_ARGS(1)+2 (show_usage) */
#define ORC_EXIT_INVALID_BLOCKS_NUM 7 /* INVALID_BLOCKS_NUM(5)+2 */
#define ORC_EXIT_INVALID_RANGE 11     /* INVALID_RANGE(9)+2 (check_range) */
/* Not a reference code: inputs on which the reference's arithmetic is
 * undefined (32-bit overflow of the balanced-mapping element count). */
#define ORC_EXIT_UNDEFINED (-1)

/* ---- A1: row mapping (pinned by golden vectors from the reference's own
 *      map_regular/map_balanced and by its abort cases) ---- */
int orc_map_regular(uint64_t n, uint32_t p, uint32_t i, uint64_t *begin, uint64_t *size);
int orc_map_balanced(uint64_t n, uint32_t p, uint32_t i, uint64_t *begin, uint64_t *size);

/* ---- A4: inputs ---- */
/* Closed form of mat_consecutive_values with MININUMS=1 (the only input the
 * reference program can produce). A: n x n, zero above the diagonal.
 * B: n x m, value index uses n as the global column count. */
void orc_fill_closed_A(int64_t n, double *A, int64_t lda);
void orc_fill_closed_B(int64_t n, int64_t m, double *B, int64_t ldb);
/* Project-defined seeded inputs (no reference counterpart): SplitMix64
 * counter generator, uniform in [-1,1); spread!=0 multiplies each entry by
 * 2^e with e uniform in {-8..8}; A's diagonal gets +2.0 before scaling. */
double orc_seeded_value(uint64_t seed, uint32_t stream, uint64_t idx, int spread, int diag);
/* The k-th output of the published SplitMix64 stream started at `state`
 * (exposed so its known-answer values can be checked). */
uint64_t orc_splitmix64(uint64_t state, uint64_t k);
void orc_fill_seeded_A(int64_t n, uint64_t seed, int spread, double *A, int64_t lda);
void orc_fill_seeded_B(int64_t n, int64_t m, uint64_t seed, int spread, double *B, int64_t ldb);

/* ---- A5-A8: C = tril(A) * B (the stage loop's result) ---- */
/* Plain FP64, every row summed in ascending k; OpenMP over rows. */
void orc_trmm(int64_t n, int64_t m, const double *A, int64_t lda, const double *B, int64_t ldb,
              double *C, int64_t ldc);
/* Same product for selected rows, summed in double-double (error-free
 * TwoProd via fma + TwoSum), i.e. accurate to ~2^-100 relative per term. */
void orc_trmm_rows_dd(int64_t n, int64_t m, const double *A, int64_t lda, const double *B,
                      int64_t ldb, const int64_t *rows, int64_t nrows, double *out, int64_t ldo);
/* The reference's per-rank stage loop replayed band by band
 * (UVaTRMM/src/mm_triangular.c:319-410): GEMM box with K=b+1, TRMM of the
 * (h-1) diagonal triangle on a copy of B, C += that, for each band of the
 * given mapping.  blas!=0 uses the host BLAS (cblas_dgemm/cblas_dtrmm from
 * libmkl_rt, loaded at run time); blas==0 or no MKL uses the C loops. */
int orc_stage_loop(int64_t n, int64_t m, uint32_t p, int balanced, const double *A, int64_t lda,
                   const double *B, int64_t ldb, double *C, int64_t ldc, int blas);
/* Rows of C for the seeded inputs (n x n A, n x m B), elements generated on
 * the fly, summed in double-double: orc_trmm_rows_dd without host copies of
 * A and B (for orders whose A does not fit in host memory). */
void orc_seeded_rows_dd(int64_t n, int64_t m, uint64_t seed, int spread, const int64_t *rows,
                        int64_t nrows, double *out, int64_t ldo);
/* 1 if MKL's CBLAS could be loaded; fills a version string. */
int orc_blas_available(char *version, int cap);
/* Makes the CBLAS library at `path` the one orc_stage_loop calls (a
 * CBLAS-compatible replacement under test); NULL returns to the MKL search.
 * 1 when cblas_dgemm and cblas_dtrmm were found. */
int orc_blas_use(const char *path);
/* CPU baseline: `ranks` threads, each running the reference's stage loop
 * over its own column band of w columns (like one MPI rank with sequential
 * MKL).  Returns seconds of wall time for the timed loop; C gets the result. */
double orc_ranks_stage_loop(int64_t n, int64_t w, uint32_t p, int balanced, int ranks,
                            const double *A, int64_t lda, const double *B, int64_t ldb,
                            double *C, int64_t ldc, int blas);

/* ---- A10: infinity norm (max row sum of |C|) ---- */
double orc_inf_norm(int64_t n, int64_t m, const double *C, int64_t ldc, double *rowsums);

#ifdef __cplusplus
}
#endif
#endif
