/*
 * scalapack_trmm.c -- the reference's second CPU comparator, ScaLAPACK's
 * distributed TRMM (UVaTRMM ScaLAPACK/scalapack_sample_trmm.c:135-139:
 * pdtrmm_('L','L','N','N', n, n, 1.0, A, 1, 1, descA, B, 1, 1, descB) on a
 * 2-D block-cyclic grid, nb = 256 by default), restated as a timing harness.
 *
 * TEST INFRASTRUCTURE ONLY: a CPU baseline run by the benchmark's CPU leg,
 * never by the product.  It differs from the reference sample in what the
 * comparison needs: B may have m != n columns, A and B hold the seeded
 * inputs of the GPU runs (orc_seeded_value, A's upper triangle zero), the
 * product is repeated `reps` times on fresh copies of B and timed from a
 * barrier to a barrier, and rank 0 prints one line
 *   n m nb nprow npcol ranks median_s min_s max_s frobenius_norm_of_C
 * (the norm lets the caller compare the result with the GPU's).  B is the
 * first m columns of the seeded n x mgen matrix (mgen >= m, default m: the
 * generator indexes element (i, j) as i * mgen + j), so a column sample of a
 * wider workload gets exactly that workload's values.
 *
 *   mpiexec -n P ./scalapack_trmm.x n m nb nprow npcol reps seed spread [mgen]
 */
#include <math.h>
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* MKL ScaLAPACK / BLACS (Fortran symbols, LP64 integers) */
void blacs_pinfo_(int *iam, int *nprocs);
void blacs_get_(int *ctxt, int *what, int *val);
void blacs_gridinit_(int *ctxt, char *order, int *nprow, int *npcol);
void blacs_gridinfo_(int *ctxt, int *nprow, int *npcol, int *myrow, int *mycol);
void blacs_gridexit_(int *ctxt);
int numroc_(int *n, int *nb, int *iproc, int *isrcproc, int *nprocs);
void descinit_(int *desc, int *m, int *n, int *mb, int *nb, int *irsrc, int *icsrc, int *ctxt,
               int *lld, int *info);
void pdtrmm_(char *side, char *uplo, char *trans, char *diag, int *m, int *n, double *alpha,
             double *a, int *ia, int *ja, int *desca, double *b, int *ib, int *jb, int *descb);

/* global index of local index l on a 1-D block-cyclic axis */
static int64_t global_of(int64_t l, int nb, int p, int np) { return (l / nb * np + p) * (int64_t)nb + l % nb; }

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    if (argc != 9 && argc != 10) {
        if (!rank) fprintf(stderr, "usage: %s n m nb nprow npcol reps seed spread [mgen]\n", argv[0]);
        MPI_Finalize();
        return 2;
    }
    int n = atoi(argv[1]), m = atoi(argv[2]), nb = atoi(argv[3]);
    int nprow = atoi(argv[4]), npcol = atoi(argv[5]), reps = atoi(argv[6]);
    const uint64_t seed = strtoull(argv[7], NULL, 10);
    const int spread = atoi(argv[8]);
    const int64_t mgen = argc == 10 ? atoll(argv[9]) : m;
    if (nprow * npcol != size || n < 1 || m < 1 || nb < 1 || reps < 1 || mgen < m) {
        if (!rank) fprintf(stderr, "grid %dx%d does not match %d ranks, or bad sizes\n", nprow, npcol, size);
        MPI_Finalize();
        return 2;
    }
    int iam, nprocs, ctxt, zero = 0, myrow, mycol, info = 0;
    char order = 'R';
    blacs_pinfo_(&iam, &nprocs);
    blacs_get_(&zero, &zero, &ctxt);
    blacs_gridinit_(&ctxt, &order, &nprow, &npcol);
    blacs_gridinfo_(&ctxt, &nprow, &npcol, &myrow, &mycol);
    /* local pieces: A n x n, B n x m, column-major with leading dimension lr */
    const int lr = numroc_(&n, &nb, &myrow, &zero, &nprow);
    const int lca = numroc_(&n, &nb, &mycol, &zero, &npcol);
    const int lcb = numroc_(&m, &nb, &mycol, &zero, &npcol);
    int ld = lr > 1 ? lr : 1;
    double *A = (double *)calloc((size_t)ld * (lca > 0 ? lca : 1), sizeof(double));
    double *B0 = (double *)calloc((size_t)ld * (lcb > 0 ? lcb : 1), sizeof(double));
    double *B = (double *)calloc((size_t)ld * (lcb > 0 ? lcb : 1), sizeof(double));
    if (!A || !B0 || !B) {
        fprintf(stderr, "rank %d: out of memory\n", rank);
        MPI_Abort(MPI_COMM_WORLD, 3);
    }
    for (int64_t c = 0; c < lca; c++) {
        const int64_t J = global_of(c, nb, mycol, npcol);
        for (int64_t r = 0; r < lr; r++) {
            const int64_t I = global_of(r, nb, myrow, nprow);
            A[c * ld + r] = J <= I ? orc_seeded_value(seed, 1u, (uint64_t)(I * n + J), spread, I == J) : 0.0;
        }
    }
    for (int64_t c = 0; c < lcb; c++) {
        const int64_t J = global_of(c, nb, mycol, npcol);
        for (int64_t r = 0; r < lr; r++) {
            const int64_t I = global_of(r, nb, myrow, nprow);
            B0[c * ld + r] = orc_seeded_value(seed, 2u, (uint64_t)(I * mgen + J), spread, 0);
        }
    }
    int descA[9], descB[9];
    descinit_(descA, &n, &n, &nb, &nb, &zero, &zero, &ctxt, &ld, &info);
    if (!info) descinit_(descB, &n, &m, &nb, &nb, &zero, &zero, &ctxt, &ld, &info);
    if (info) {
        fprintf(stderr, "rank %d: descinit info %d\n", rank, info);
        MPI_Abort(MPI_COMM_WORLD, 4);
    }
    double *times = (double *)malloc((size_t)reps * sizeof(double));
    char side = 'L', uplo = 'L', trans = 'N', diag = 'N';
    double alpha = 1.0;
    int one = 1;
    for (int r = 0; r < reps; r++) {
        memcpy(B, B0, (size_t)ld * (lcb > 0 ? lcb : 1) * sizeof(double));
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        pdtrmm_(&side, &uplo, &trans, &diag, &n, &m, &alpha, A, &one, &one, descA, B, &one, &one, descB);
        MPI_Barrier(MPI_COMM_WORLD);
        times[r] = MPI_Wtime() - t0;
    }
    double local = 0.0, total = 0.0;
    for (int64_t c = 0; c < lcb; c++)
        for (int64_t r = 0; r < lr; r++) local += B[c * ld + r] * B[c * ld + r];
    MPI_Reduce(&local, &total, 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    if (!rank) {
        double lo = times[0], hi = times[0];
        for (int r = 1; r < reps; r++) {
            lo = times[r] < lo ? times[r] : lo;
            hi = times[r] > hi ? times[r] : hi;
        }
        qsort(times, (size_t)reps, sizeof(double), cmp_double);
        const double med = reps % 2 ? times[reps / 2] : 0.5 * (times[reps / 2 - 1] + times[reps / 2]);
        printf("%d %d %d %d %d %d %.6f %.6f %.6f %.17g\n", n, m, nb, nprow, npcol, size, med, lo, hi,
               sqrt(total));
    }
    free(times);
    free(A);
    free(B0);
    free(B);
    blacs_gridexit_(&ctxt);
    MPI_Finalize();
    return 0;
}
