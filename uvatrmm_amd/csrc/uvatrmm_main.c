/*
 * uvatrmm_main.c -- the uvatrmm.x command line on one MI355X.
 *
 *   uvatrmm.x <N> <reg|balanced> <boxes|trapezoid> [options]
 *
 * Same positional arguments, usage text, exit codes and stdout lines as the
 * reference program (UVaTRMM/src/mm_triangular.c:7-436 with
 * UVaTRMM/src/mm_triangular_common.c:1206-1225, 1446-1513, 2936-2991).  The P
 * MPI ranks of the reference become P virtual row bands on one GPU
 * (--bands P, default 1); "Total mpis" reports P.
 *
 * Options (none of them exist in the reference; defaults reproduce it):
 *   --bands P        number of row bands (the reference's mpirun -np P)
 *   --m M            columns of B and C (reference: M = N)
 *   --prec fp64|split   arithmetic (default fp64); --slices S for split
 *   --init ref|seeded   reference closed-form inputs (default) or seeded
 *   --seed S, --spread  seeded inputs, exponent spread 2^[-8,8]
 *   --reps R         timed repetitions after one warm-up run (default 1)
 *   --device D       GPU ordinal (default 0)
 *   --json           also print one JSON line with the measurements
 *   --dump DIR       write C in the reference's binary result format, one
 *                    file per band's column slice
 *                    (DIR/output_uvatrmm.<N>.<P>.<rank>.dat)
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "uvt.h"

/* error_mpi_abort exits with the error enum + 2 (mm_triangular_common.c:2854-2863;
 * include/error.h:20-39). */
#define EXIT_ARGS 3    /* INVALID_ARGS */
#define EXIT_MEMORY 6  /* MEMORY_ERROR */
#define EXIT_OTHER 18  /* OTHER_ERROR */

static void show_usage(const char *prog, const char *msg)
{
    /* text of show_usage, mm_triangular_common.c:1206-1225 */
    fprintf(stderr, "\nError: %s\n", msg);
    fprintf(stderr, "\nUsage: %s <num_rows> <map_func> <comm_shape>\n\n", prog);
    fprintf(stderr, "Map functions:\n");
    fprintf(stderr, "\treg\tRegular with remainder in the last proc.\n");
    fprintf(stderr, "\tbalanced\tBalanced according to the total number of non-zero elements\n");
    fprintf(stderr, "\n");
    fprintf(stderr, "Communication shapes:\n");
    fprintf(stderr, "\tboxes\tMinimal rectangular shapes, bounding boxes including the triangle\n");
    fprintf(stderr,
            "\ttrapezoid\tExact comm. One indexed type with the trapezoid (rectangle + triangle\n");
    fprintf(stderr, "\n");
    fprintf(stderr, "Options of this build: [--bands P] [--m M] [--prec fp64|split] [--slices S]\n"
                    "  [--init ref|seeded] [--seed S] [--spread] [--reps R] [--device D] [--json]\n"
                    "  [--dump DIR]\n");
    exit(EXIT_ARGS);
}

static int fail(uvt_status s, const char *what)
{
    fprintf(stderr, "uvatrmm.x: %s: %s\n", what, uvt_status_string(s));
    return s == UVT_ERR_NOMEM ? EXIT_MEMORY : EXIT_OTHER;
}

#define CHECK(call, what)                       \
    do {                                        \
        uvt_status st_ = (call);                \
        if (st_ != UVT_OK) return fail(st_, what); \
    } while (0)

/* Result dump in the reference's layout (print_output_results,
 * mm_triangular_common.c:2103-2132), one file per virtual rank r of P:
 * DIR/output_uvatrmm.<N>.<P>.<r>.dat holding int rows, int columns,
 * int range[4] = {0, rows, column begin, column count} and then rows x count
 * doubles, the rank's column band being map_regular over the columns
 * (mm_triangular.c:90). */
static int dump_result(const char *dir, long n, long m, unsigned p, const double *C)
{
    for (unsigned r = 0; r < p; r++) {
        uvt_range cols;
        if (uvt_map_band((uint64_t)m, p, r, UVT_MAP_REG, &cols, NULL) != UVT_OK) return -1;
        char path[4096];
        snprintf(path, sizeof path, "%s/output_uvatrmm.%ld.%u.%u.dat", dir, n, p, r);
        FILE *f = fopen(path, "wb");
        if (!f) return -1;
        int hdr[6] = {(int)n, (int)m, 0, (int)n, (int)cols.begin, (int)cols.size};
        int ok = fwrite(hdr, sizeof(int), 6, f) == 6;
        for (long i = 0; ok && i < n && cols.size; i++)
            ok = fwrite(C + (size_t)i * m + cols.begin, sizeof(double), cols.size, f) == cols.size;
        if (fclose(f) != 0 || !ok) return -1;
    }
    return 0;
}

/* Host wall clock in seconds (monotonic). */
static double wall_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv)
{
    if (argc < 4) show_usage(argv[0], "Number of arguments");
    long n = atol(argv[1]);
    uvt_map map;
    uvt_shape shape;
    if (!strcmp(argv[2], "reg")) map = UVT_MAP_REG;
    else if (!strcmp(argv[2], "balanced")) map = UVT_MAP_BALANCED;
    else show_usage(argv[0], "Unknown name of mapping function");
    if (!strcmp(argv[3], "boxes")) shape = UVT_SHAPE_BOXES;
    else if (!strcmp(argv[3], "trapezoid")) shape = UVT_SHAPE_TRAPEZOID;
    else show_usage(argv[0], "Unknown name of communication shape");

    long bands = 1, m = -1, reps = 1, seed = 1, slices = 13;
    int device = 0, json = 0, seeded = 0, spread = 0, split = 0;
    const char *dump = NULL;
    for (int i = 4; i < argc; i++) {
        const char *a = argv[i];
        const char *v = (i + 1 < argc) ? argv[i + 1] : NULL;
        if (!strcmp(a, "--bands") && v) bands = atol(v), i++;
        else if (!strcmp(a, "--m") && v) m = atol(v), i++;
        else if (!strcmp(a, "--reps") && v) reps = atol(v), i++;
        else if (!strcmp(a, "--seed") && v) seed = atol(v), i++;
        else if (!strcmp(a, "--slices") && v) slices = atol(v), i++;
        else if (!strcmp(a, "--device") && v) device = atoi(v), i++;
        else if (!strcmp(a, "--dump") && v) dump = v, i++;
        else if (!strcmp(a, "--init") && v) {
            if (!strcmp(v, "ref")) seeded = 0;
            else if (!strcmp(v, "seeded")) seeded = 1;
            else show_usage(argv[0], "Unknown --init");
            i++;
        } else if (!strcmp(a, "--prec") && v) {
            if (!strcmp(v, "fp64")) split = 0;
            else if (!strcmp(v, "split")) split = 1;
            else show_usage(argv[0], "Unknown --prec");
            i++;
        } else if (!strcmp(a, "--spread")) spread = 1;
        else if (!strcmp(a, "--json")) json = 1;
        else show_usage(argv[0], "Unknown option");
    }
    if (m < 0) m = n;
    if (bands < 1 || reps < 1 || m < 1 || n < 1)
        show_usage(argv[0], "Sizes, bands and repetitions must be positive");
    /* mm_triangular.c:79 */
    if (n < bands)
        show_usage(argv[0], "Number of rows should be equal or greater than number of processors");

    int ref_exit = 0;
    uvt_plan plan = NULL;
    uvt_status st = uvt_plan_create(n, m, (uint32_t)bands, map, shape, &plan, &ref_exit);
    if (st == UVT_ERR_MAPPING) return ref_exit;
    CHECK(st, "plan");

    uvt_handle h = NULL;
    CHECK(uvt_create(device, &h), "device");
    double *A = NULL, *B = NULL, *C = NULL;
    CHECK(uvt_malloc(h, (size_t)n * n * sizeof(double), (void **)&A), "allocate A");
    CHECK(uvt_malloc(h, (size_t)n * m * sizeof(double), (void **)&B), "allocate B");
    CHECK(uvt_malloc(h, (size_t)n * m * sizeof(double), (void **)&C), "allocate C");
    if (seeded) CHECK(uvt_fill_seeded(h, n, m, (uint64_t)seed, spread, A, n, B, m, NULL), "init");
    else CHECK(uvt_fill_closed_form(h, n, m, A, n, B, m, NULL), "init");

    void *work = NULL;
    size_t work_bytes = 0;
    if (split) {
        CHECK(uvt_trmm_split_workspace(n, m, (int32_t)slices, &work_bytes), "split workspace");
        CHECK(uvt_malloc(h, work_bytes, &work), "allocate split workspace");
    }
    void *stream = NULL, *e0 = NULL, *e1 = NULL;
    CHECK(uvt_stream_create(h, &stream), "stream");
    CHECK(uvt_event_create(h, &e0), "event");
    CHECK(uvt_event_create(h, &e1), "event");

    /* Timers with the reference's meanings (mm_triangular.c:316-329, 407-423):
     * Total t spans the whole product on the host clock, from after the
     * stream is idle (the reference's first barrier) until its work has
     * completed (the second barrier), launches and synchronisation included;
     * Mult t is the time the device spent on the product, timed with events
     * on its stream around the launches of every band (all bands are one
     * tile walk here).  The last repetition is printed. */
    double best = 1e300, total = 0.0, last_total = 0.0, last_mult = 0.0;
    for (long r = -1; r < reps; r++) { /* r = -1: untimed warm-up */
        CHECK(uvt_stream_sync(h, stream), "sync");
        const double t0 = wall_seconds();
        CHECK(uvt_event_record(h, e0, stream), "event");
        if (split)
            CHECK(uvt_trmm_split(h, n, m, (int32_t)slices, A, n, B, m, C, m, work, work_bytes,
                                 stream),
                  "trmm");
        else
            CHECK(uvt_trmm_f64(h, plan, A, n, B, m, C, m, stream), "trmm");
        CHECK(uvt_event_record(h, e1, stream), "event");
        CHECK(uvt_stream_sync(h, stream), "sync");
        const double t1 = wall_seconds();
        float ms = 0.f;
        CHECK(uvt_event_elapsed(h, e0, e1, &ms), "timer");
        if (r >= 0) {
            last_mult = ms * 1e-3;
            last_total = t1 - t0;
            total += last_mult;
            if (last_mult < best) best = last_mult;
        }
    }
    /* mm_triangular.c:426 */
    printf("\nTotal t: %lf\nTotal mpis: %d\nMult t: %lf\n", last_total, (int)bands, last_mult);

    double norm = 0.0;
    CHECK(uvt_inf_norm(h, n, m, C, m, &norm, stream), "norm");
    /* mm_triangular_common.c:2986-2988 */
    printf("\nNorm: %.3lf\n", norm);
    FILE *nf = fopen("norm_ff.txt", "w+");
    if (nf) {
        fprintf(nf, "%.3lf", norm);
        fclose(nf);
    }
    if (json) {
        double fl = uvt_plan_flops(plan);
        printf("{\"n\": %ld, \"m\": %ld, \"bands\": %ld, \"map\": \"%s\", \"shape\": \"%s\", "
               "\"prec\": \"%s\", \"reps\": %ld, \"best_s\": %.6e, \"mean_s\": %.6e, "
               "\"tflops_best\": %.4f, \"norm\": %.17g}\n",
               n, m, bands, argv[2], argv[3], split ? "split" : "fp64", reps, best,
               total / reps, fl / best * 1e-12, norm);
    }
    if (dump) {
        double *hc = (double *)malloc((size_t)n * m * sizeof(double));
        if (!hc) return EXIT_MEMORY;
        CHECK(uvt_memcpy(h, hc, C, (size_t)n * m * sizeof(double), 1), "copy C");
        if (dump_result(dump, n, m, (unsigned)bands, hc)) {
            fprintf(stderr, "uvatrmm.x: cannot write result dump in %s\n", dump);
            return EXIT_OTHER;
        }
        free(hc);
    }
    uvt_event_destroy(h, e0);
    uvt_event_destroy(h, e1);
    uvt_stream_destroy(h, stream);
    uvt_free(h, work);
    uvt_free(h, A);
    uvt_free(h, B);
    uvt_free(h, C);
    uvt_destroy(h);
    uvt_plan_destroy(plan);
    return 0;
}
