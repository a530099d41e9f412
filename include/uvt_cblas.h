/*
 * uvt_cblas.h -- CBLAS-compatible host interface of libuvt_cblas.so, the
 * drop-in for the two MKL routines the UVaTRMM stage loop calls
 * (UVaTRMM/src/mm_triangular.c:347-360 cblas_dgemm, :374-387 cblas_dtrmm),
 * with MKL's LP64 interface (32-bit int sizes, UVaTRMM/Makefile:19) and the
 * public CBLAS enum values.  A caller linked with -luvt_cblas in place of
 * -lmkl_intel_lp64 -lmkl_sequential -lmkl_core runs unchanged.
 *
 * Semantics are those of CBLAS: host pointers owned by the caller, both
 * layouts, every transpose / side / triangle / diagonal variant, alpha and
 * beta, dtrmm in place on B, nothing returned.  Each call copies its
 * operands to the GPU, runs the FP64 MFMA kernels of libuvt, synchronizes
 * and copies the result back; calls are serialized across threads.
 *
 * Errors: an illegal parameter prints its position (CBLAS numbering) to
 * stderr and returns with no output written, as MKL does; a GPU failure (no
 * device, out of memory, a HIP error) prints the reason and aborts, since a
 * silently unwritten result would corrupt the caller's computation.
 * Environment: UVT_CBLAS_DEVICE selects the GPU (default 0).
 */
#ifndef UVT_CBLAS_H
#define UVT_CBLAS_H

#ifdef __cplusplus
extern "C" {
#endif
/* the library builds with hidden symbols; these declarations are its exports */
#pragma GCC visibility push(default)

/* Callers that already include a cblas.h define UVT_CBLAS_NO_ENUMS. */
#ifndef UVT_CBLAS_NO_ENUMS
typedef enum { CblasRowMajor = 101, CblasColMajor = 102 } CBLAS_LAYOUT;
typedef CBLAS_LAYOUT CBLAS_ORDER;
typedef enum { CblasNoTrans = 111, CblasTrans = 112, CblasConjTrans = 113 } CBLAS_TRANSPOSE;
typedef enum { CblasUpper = 121, CblasLower = 122 } CBLAS_UPLO;
typedef enum { CblasNonUnit = 131, CblasUnit = 132 } CBLAS_DIAG;
typedef enum { CblasLeft = 141, CblasRight = 142 } CBLAS_SIDE;
#endif

/* C = alpha op(A) op(B) + beta C. */
void cblas_dgemm(CBLAS_LAYOUT layout, CBLAS_TRANSPOSE transa, CBLAS_TRANSPOSE transb, int m, int n,
                 int k, double alpha, const double *a, int lda, const double *b, int ldb,
                 double beta, double *c, int ldc);
/* B := alpha op(A) B (left) or alpha B op(A) (right), A triangular. */
void cblas_dtrmm(CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, CBLAS_TRANSPOSE transa,
                 CBLAS_DIAG diag, int m, int n, double alpha, const double *a, int lda, double *b,
                 int ldb);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
