/*
 * uvt.h -- C ABI of the MI355X TRMM engine (libuvt.so).
 *
 * The engine computes the local product of the UVaTRMM program,
 * C = tril(A) * B for a lower-triangular, non-unit, row-major FP64 A (n x n)
 * and row-major B, C (n x m), on one AMD MI355X (gfx950).
 *
 * What each group of entry points replaces in the reference (paths relative
 * to the UVaTRMM source tree):
 *   uvt_map_band / uvt_plan_bands   map_regular / map_balanced,
 *                                   UVaTRMM/src/mm_mapping.c:27-67, 319-366
 *   uvt_plan_*                      the stage loop's per-band shapes,
 *                                   UVaTRMM/src/mm_triangular.c:319-410
 *                                   (and the boxes/trapezoid shape choice,
 *                                   UVaTRMM/src/mm_triangular_common.c:1477-1513)
 *   uvt_trmm_f64                    the per-rank cblas_dgemm + cblas_dtrmm +
 *                                   C += B sequence, mm_triangular.c:347-393
 *   uvt_inf_norm / uvt_row_abs_sums matrix_rangeInfinityNorm,
 *                                   mm_triangular_common.c:2936-2991
 *   uvt_fill_closed_form            mat_consecutive_values (MININUMS=1),
 *                                   mm_triangular_common.c:1782-1857
 *   uvt_trmm_split, uvt_batch_*,    no counterpart in the reference (split
 *   uvt_serve_*                     precision, batched and served TRMMs)
 *
 * Conventions: every pointer named d* is device memory owned by the caller;
 * sizes and leading dimensions are element counts (int64); every function
 * returns a uvt_status and never aborts the process; work is asynchronous on
 * the given stream (a hipStream_t passed as void*, NULL = the null stream)
 * unless a comment says "synchronous".  A handle is not thread-safe: use one
 * per host thread.
 */
#ifndef UVT_H
#define UVT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
/* opaque handles, declared before the export region: their members are the
 * library's own */
struct uvt_plan_s;
struct uvt_handle_s;
struct uvt_batch_s;
struct uvt_server_s;
/* the library builds with hidden symbols; these declarations are its exports */
#pragma GCC visibility push(default)

typedef enum {
    UVT_OK = 0,
    UVT_ERR_ARG = 1,         /* null pointer, negative size, bad enum, ld too small */
    UVT_ERR_SIZE = 2,        /* a size the engine does not support */
    UVT_ERR_MAPPING = 3,     /* the reference would abort on this mapping (see ref_exit) */
    UVT_ERR_HIP = 4,         /* a HIP runtime call failed */
    UVT_ERR_NOMEM = 5,       /* host or device allocation failed */
    UVT_ERR_BUDGET = 6,      /* request exceeds the serving memory budget */
    UVT_ERR_UNSUPPORTED = 7, /* no GPU, or a feature this build lacks */
    UVT_ERR_STATE = 8,       /* call out of order (e.g. server closed) */
    UVT_ERR_AGAIN = 9        /* nothing ready yet (polling) */
} uvt_status;

/* Range{begin,size} of include/mkl_range.h:18-21 (rows of a band). */
typedef struct {
    uint64_t begin;
    uint64_t size;
} uvt_range;

typedef enum { UVT_MAP_REG = 0, UVT_MAP_BALANCED = 1 } uvt_map;
/* boxes: per band a rectangular GEMM box over k in [0, b] followed by the
 * (h-1)-row diagonal triangle accumulated into C (two launches).
 * trapezoid: one triangle-aware tile walk over the lower trapezoid. */
typedef enum { UVT_SHAPE_BOXES = 0, UVT_SHAPE_TRAPEZOID = 1 } uvt_shape;
typedef enum { UVT_PREC_F64 = 0, UVT_PREC_SPLIT = 1 } uvt_prec;

const char *uvt_status_string(uvt_status s);
const char *uvt_version(void);

/* ------------------------------------------------------------------ */
/* Band mapping (host only; no GPU needed).                            */
/* ------------------------------------------------------------------ */
/* Band i of p over n rows, bit-identical to the reference.  When the
 * reference would abort, returns UVT_ERR_MAPPING and sets *ref_exit to the
 * reference's process exit code (7: too many bands, 11: an empty band);
 * n whose balanced element count overflows the reference's 32-bit counter
 * (n >= 92682) give UVT_ERR_SIZE.  ref_exit may be NULL. */
uvt_status uvt_map_band(uint64_t n, uint32_t p, uint32_t i, uvt_map map, uvt_range *out,
                        int *ref_exit);
/* All p bands (out has room for p). */
uvt_status uvt_plan_bands(uint64_t n, uint32_t p, uvt_map map, uvt_range *out, int *ref_exit);

/* ------------------------------------------------------------------ */
/* Tile plans (host objects; uploaded to the device on first use).     */
/* ------------------------------------------------------------------ */
typedef struct uvt_plan_s *uvt_plan;

/* One output tile of the FP64 kernels: rows [r0, r0+rows), columns
 * [c0, c0+tile_n) clipped to m, reduction over k in [k_lo, k_hi).  cfg is
 * its shape: a uvt_tile_cfg, or UVT_TILE_256x128 in the plans of large
 * problems whose 256 x 128 tiles are all full (n a multiple of 256, m of
 * 128, at least 512 tiles). */
typedef struct {
    int32_t r0, rows, c0, k_lo, k_hi, flags, band, cfg;
} uvt_tile;
#define UVT_TILE_256x128 4 /* 256 x 128 tiles (not selectable per batch problem) */
#define UVT_TILE_LOWER 1 /* A[i][k] is treated as 0 for k > i */
#define UVT_TILE_ACCUM 2 /* C += product (else C = product) */

/* Plans C = tril(A) * B for an n x n A and n x m B over p bands of the given
 * mapping and shape.  Errors as uvt_plan_bands. */
uvt_status uvt_plan_create(int64_t n, int64_t m, uint32_t p, uvt_map map, uvt_shape shape,
                           uvt_plan *out, int *ref_exit);
uvt_status uvt_plan_destroy(uvt_plan plan);
/* Number of launches (phases): 2 for boxes, 1 for trapezoid. */
int32_t uvt_plan_phases(uvt_plan plan);
int64_t uvt_plan_num_tiles(uvt_plan plan, int32_t phase);
uvt_status uvt_plan_get_tiles(uvt_plan plan, int32_t phase, uvt_tile *out, int64_t cap);
uvt_status uvt_plan_get_bands(uvt_plan plan, uvt_range *out, uint32_t cap);
/* Tile edge lengths used by the plan (rows, columns). */
uvt_status uvt_plan_tile_shape(uvt_plan plan, int32_t *tile_m, int32_t *tile_n);
/* Algorithmic flops m*n^2 (LAWN 41 count for a left triangular product). */
double uvt_plan_flops(uvt_plan plan);
/* MFMA flops actually issued, including masked diagonal tiles and padding. */
double uvt_plan_issued_flops(uvt_plan plan);

/* ------------------------------------------------------------------ */
/* Device handle.                                                      */
/* ------------------------------------------------------------------ */
typedef struct uvt_handle_s *uvt_handle;
uvt_status uvt_create(int device, uvt_handle *out);
uvt_status uvt_destroy(uvt_handle h);
/* Number of compute units of the handle's device. */
int32_t uvt_device_cus(uvt_handle h);

/* FP64 C = tril(A) * B following the plan (out of place; A and B are not
 * written; C is fully overwritten on rows [0, n) x columns [0, m)).  Entries
 * of A above the diagonal are never read.  The kernels multiply whole
 * 16-deep k panels of B, so a NaN or infinity in B[k][j] also reaches the
 * rows above k in k's 128-row tile through A's structural zeros (0 * Inf =
 * NaN), where a BLAS dtrmm would leave them finite; finite inputs are not
 * affected.  On its first use on a device the plan allocates its launch
 * records there.  A trapezoid plan of at least 2e9 flops whose whole tiles
 * would not finish together (narrow B: few, deep tiles) splits its long
 * reductions and adds the pieces after the first of a tile to C afterwards
 * in k order (results repeat bit for bit), with up to 64 MiB of partial
 * tiles on the device: with 128x128 tiles whose whole-tile span is 1.3x or
 * more an even share of the work, stream-K (up to one workgroup per slot,
 * each the same number of k panels, in groups that run a row tile's column
 * tiles side by side); with smaller tiles whose deepest tile is 1.5x or
 * more a slot's share, cuts of those tiles into pieces.  The device state
 * belongs to the plan: calls with one plan on different streams must be
 * serialized. */
uvt_status uvt_trmm_f64(uvt_handle h, uvt_plan plan, const double *dA, int64_t lda,
                        const double *dB, int64_t ldb, double *dC, int64_t ldc, void *stream);
/* How the plan launches on the handle's device (placing its device copies
 * there if needed): workgroups of its tile launches (for stream-K, up to one
 * per workgroup slot), and partial tiles (0 when every tile writes C directly).
 * Either pointer may be NULL. */
uvt_status uvt_plan_launch_info(uvt_handle h, uvt_plan plan, int64_t *workgroups, int64_t *partial_tiles);

/* Split-precision C ~= tril(A) * B: the box below each 256-row diagonal
 * block as exact int8 products of residues modulo `slices` coprime moduli
 * (2..14; int32 accumulation), recombined exactly (Chinese remainder) and
 * rounded to FP64; the diagonal blocks in FP64.  More moduli, smaller
 * error: 13 gives ~1e-13 relative at n = 32768.  dwork: 256-byte aligned
 * device workspace of uvt_trmm_split_workspace bytes (n <= 131072). */
uvt_status uvt_trmm_split_workspace(int64_t n, int64_t m, int32_t slices, size_t *bytes);
uvt_status uvt_trmm_split(uvt_handle h, int64_t n, int64_t m, int32_t slices, const double *dA,
                          int64_t lda, const double *dB, int64_t ldb, double *dC, int64_t ldc,
                          void *dwork, size_t work_bytes, void *stream);
/* The error statement of the last uvt_trmm_split run on this workspace with
 * these n, m, slices (synchronous on `stream`), both relative to ||C||_F and
 * computed on the device from the scales and norms of that call:
 *   bound:    rigorous, >= ||C - tril(A) B||_F / ||tril(A) B||_F (from the
 *             worst-case integer roundings, the FP64 diagonal blocks and the
 *             final roundings; inf if it exceeds ||C||_F);
 *   estimate: the expected value of that error when the roundings to integers
 *             are independent and uniform (random inputs measure ~0.9x of it;
 *             the bound is typically 30-150x larger).
 * Both are NaN when A or B holds a NaN or infinity in the part the integer
 * products read; those entries of C are then NaN. */
uvt_status uvt_trmm_split_error(uvt_handle h, int64_t n, int64_t m, int32_t slices,
                                const void *dwork, size_t work_bytes, double *bound,
                                double *estimate, void *stream);

/* ------------------------------------------------------------------ */
/* BLAS forms of the two per-band calls, device-resident and row-major */
/* (the CBLAS shim libuvt_cblas maps the host interface onto them).    */
/* ------------------------------------------------------------------ */
typedef enum { UVT_NO_TRANS = 0, UVT_TRANS = 1 } uvt_trans;
typedef enum { UVT_LEFT = 0, UVT_RIGHT = 1 } uvt_side;
typedef enum { UVT_LOWER = 0, UVT_UPPER = 1 } uvt_uplo;
typedef enum { UVT_NON_UNIT = 0, UVT_UNIT = 1 } uvt_diag;
/* C = alpha op(A) op(B) + beta C; C (m x n) is not read when beta == 0, A and
 * B are not read when k == 0 or alpha == 0.  Replaces the cblas_dgemm call of
 * UVaTRMM/src/mm_triangular.c:347-360 (RowMajor, NoTrans, NoTrans, alpha 1,
 * beta 0) and covers the other transposes.  Uses the handle's workspace:
 * calls on one handle that use it (uvt_dgemm, uvt_dtrmm, uvt_inf_norm) must
 * be on one stream or serialized by the caller. */
uvt_status uvt_dgemm(uvt_handle h, uvt_trans ta, uvt_trans tb, int64_t m, int64_t n, int64_t k,
                     double alpha, const double *dA, int64_t lda, const double *dB, int64_t ldb,
                     double beta, double *dC, int64_t ldc, void *stream);
/* B (m x n) := alpha op(A) B (left, A m x m) or alpha B op(A) (right, A n x n)
 * in place, A triangular (uplo; only that triangle is read; unit: its
 * diagonal is taken as 1).  Replaces the cblas_dtrmm call of
 * mm_triangular.c:374-387 (RowMajor, Left, Lower, NoTrans, NonUnit, alpha 1)
 * and covers every other side / uplo / trans / diag.  Uses the handle's
 * workspace (see uvt_dgemm). */
uvt_status uvt_dtrmm(uvt_handle h, uvt_side side, uvt_uplo uplo, uvt_trans ta, uvt_diag diag,
                     int64_t m, int64_t n, double alpha, const double *dA, int64_t lda,
                     double *dB, int64_t ldb, void *stream);

/* uvt_dtrmm that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results (bit for bit) and errors as uvt_dtrmm, but its workspace is a fixed
 * per-device scratch array of UVT_NOALLOC_SCRATCH_MB MiB inside the library
 * (held by every device the library runs on until the process ends) instead
 * of the handle's workspace: calls needing more (for a left-side lower,
 * non-transposed, non-unit A: 8 * m * (n rounded up to even) bytes; other
 * forms also pack A and/or B) return UVT_ERR_SIZE.  Every call on a device
 * shares that device's array, so calls that overlap on one device (other
 * streams, other handles, other threads) race: serialize them.  Nothing in
 * the array needs to start at zero. */
#define UVT_NOALLOC_SCRATCH_MB 128
uvt_status uvt_dtrmm_noalloc(uvt_handle h, uvt_side side, uvt_uplo uplo, uvt_trans ta,
                             uvt_diag diag, int64_t m, int64_t n, double alpha, const double *dA,
                             int64_t lda, double *dB, int64_t ldb, void *stream);

/* ------------------------------------------------------------------ */
/* Batched TRMMs (no counterpart in the reference).                    */
/* ------------------------------------------------------------------ */
typedef struct {
    const double *A; /* device, n x n lower, row-major */
    int64_t lda;
    const double *B; /* device, n x m */
    int64_t ldb;
    double *C;       /* device, n x m */
    int64_t ldc;
    int64_t n, m;
    int32_t prec;    /* UVT_PREC_F64 or UVT_PREC_SPLIT */
    int32_t route;   /* UVT_ROUTE_*: where a split-class problem's box part runs */
} uvt_problem;

/* auto: the int8 path where it is measured faster in a batch (n >= 8192 and
 * m >= 256), FP64 elsewhere; int8: the int8 path wherever it supports the
 * size (n > 256, n <= 131072, enough moduli for the depth); fp64: FP64.  FP64
 * problems take auto or fp64. */
enum { UVT_ROUTE_AUTO = 0, UVT_ROUTE_INT8 = 1, UVT_ROUTE_FP64 = 2 };

typedef struct uvt_batch_s *uvt_batch;

/* Plans a batch: the FP64 problems become one longest-first tile queue
 * consumed by persistent kernels (tile shape per problem from the tuning
 * table: a file written by uvt_autotune_f64, or NULL for the built-in one);
 * the split problems run through batched split stages with `slices` moduli.
 * Everything is captured once into a hipGraph on the handle's device.  Each
 * problem's result does not depend on the other problems of the batch. */
uvt_status uvt_batch_create(uvt_handle h, const uvt_problem *probs, int32_t count, int32_t slices,
                            const char *tune_table, uvt_batch *out);
/* One hipGraphLaunch of the whole batch on `stream`. */
uvt_status uvt_batch_run(uvt_batch b, void *stream);
/* New pointers for the same shapes and precisions (descriptors live in
 * device memory, so the captured graph picks them up); synchronous. */
uvt_status uvt_batch_update(uvt_batch b, const uvt_problem *probs, int32_t count);
/* Algorithmic flops (sum of m n^2), the part of them whose box runs on the
 * int8 path (see uvt_problem.route), FP64 tiles queued, graph nodes.  Any
 * pointer may be NULL. */
uvt_status uvt_batch_info(uvt_batch b, double *flops, double *int8_flops, int64_t *fp64_tiles,
                          int32_t *graph_nodes);
/* The error statement of every problem after the last run (synchronous):
 * for problems whose box ran on the int8 path, the rigorous bound and the
 * expected value of the relative Frobenius error (see uvt_trmm_split_error);
 * -1 for problems that ran wholly in FP64.  Arrays of count entries (either
 * may be NULL). */
uvt_status uvt_batch_errors(uvt_batch b, double *bound, double *estimate, int32_t count);
uvt_status uvt_batch_destroy(uvt_batch b);

/* FP64 tile shapes selectable per problem. */
typedef enum { UVT_TILE_128x128 = 0, UVT_TILE_64x64 = 1, UVT_TILE_128x32 = 2 } uvt_tile_cfg;
/* Times every tile shape on (n, m) buckets (n 64..4096, m 16..1024) and
 * writes the winners as a JSON table to out_path (synchronous; seconds). */
uvt_status uvt_autotune_f64(uvt_handle h, const char *out_path);
/* The tile shape the table (NULL: built-in) gives an n x n by n x m problem. */
uvt_status uvt_tune_choice(const char *tune_table, int64_t n, int64_t m, int32_t *cfg);

/* ------------------------------------------------------------------ */
/* Serving (no counterpart in the reference).                          */
/* ------------------------------------------------------------------ */
/* A server owns a scheduler thread, a device memory arena of hbm_budget
 * bytes and max_slots streams, each with a fixed buffer for the partial
 * tiles of a stream-K FP64 product (64 MiB on a GPU of 256 CUs; an FP64
 * request runs stream-K where a plan of its size would).  Submitted
 * requests wait in FIFO order and
 * are admitted as soon as a slot is free and their operands (A, B, C and,
 * for split requests, the workspace) fit in the unused part of the arena;
 * admitted requests run concurrently on their slot streams, so requests that
 * arrive while others are on the GPU join them (continuous batching).  A
 * request's result does not depend on what else is in flight.  The server
 * uses only the handle's device (not its workspace), so the handle stays
 * usable from the caller's thread.
 * Large requests can be capped: with max_big > 0, at most max_big requests
 * of m n^2 >= big_flops run at once, and smaller requests queued behind a
 * held-back large one are admitted past it (large ones keep their order).
 * Sharing the GPU among fewer large requests shortens each of them, which
 * is what sets the tail latency of a mixed stream.  With big_wait_ms > 0 the
 * large requests go shortest first, which lowers the latency of most of
 * them, and the longest then wait up to about big_wait_ms plus the work
 * ahead of them. */
typedef struct {
    size_t hbm_budget; /* bytes of the operand arena (the slots' partial tiles come on top) */
    int32_t max_slots; /* requests in flight on the device (1..64) */
    int32_t slices;    /* moduli for split-class requests */
    const char *tune_table; /* FP64 tile shapes (NULL: built-in) */
    double big_flops;  /* m n^2 from which a request counts as large */
    int32_t max_big;   /* large requests in flight at once (0: no cap) */
    int32_t big_wait_ms; /* order of large requests held by the cap: 0 first come first
                            served; > 0 the one of fewest flops first, except that one
                            waiting this many ms or longer goes first (oldest first) */
} uvt_serve_config;

typedef struct {
    int64_t n, m;
    int32_t prec;     /* tolerance class: UVT_PREC_F64 (<= 1e-13) or UVT_PREC_SPLIT */
    int32_t seeded;   /* 1: A and B generated on the device from seed (uvt_fill_seeded);
                         0: read from the pointers A and B */
    const double *A;  /* n x n row-major (seeded = 0): host, or device with UVT_REQ_AB_DEVICE */
    const double *B;  /* n x m row-major (seeded = 0): host, or device with UVT_REQ_AB_DEVICE */
    double *C;        /* n x m row-major destination: host (copied back before the
                         completion is posted), device with UVT_REQ_C_DEVICE (written in
                         place), or NULL to leave the result in the arena and drop it */
    uint64_t seed;
    int32_t spread;
    int32_t flags;    /* UVT_REQ_* */
    uint64_t tag;     /* returned in the completion */
} uvt_request;

enum {
    UVT_REQ_AB_DEVICE = 1, /* A and B are device pointers, read in place (no arena copy) */
    UVT_REQ_C_DEVICE = 2   /* C is a device pointer, written in place */
};

typedef struct {
    uint64_t ticket, tag;
    uvt_status status;
    int32_t path;                     /* 1: the box ran on the int8 path, 0: FP64 */
    double t_submit, t_admit, t_done; /* seconds since the server started */
    double err_bound, err_estimate;   /* int8 path: as uvt_trmm_split_error; FP64: -1 */
} uvt_completion;

typedef struct {
    int64_t submitted, completed, rejected;
    double p50_s, p99_s, mean_s, max_s; /* latency: submit -> result ready */
    double wall_s;                      /* first submit -> last completion */
    double flops;                       /* sum of m n^2 over completed requests */
    double peak_bytes;                  /* peak arena use */
    int32_t peak_inflight;
    int32_t reserved;
} uvt_serve_stats;

typedef struct uvt_server_s *uvt_server;
uvt_status uvt_serve_create(uvt_handle h, const uvt_serve_config *cfg, uvt_server *out);
/* Non-blocking.  UVT_ERR_BUDGET if the request alone exceeds the budget. */
uvt_status uvt_serve_submit(uvt_server s, const uvt_request *r, uint64_t *ticket);
/* Non-blocking: up to cap finished requests (UVT_OK even if none). */
uvt_status uvt_serve_poll(uvt_server s, uvt_completion *out, int32_t cap, int32_t *count);
/* Blocks until every submitted request has completed. */
uvt_status uvt_serve_drain(uvt_server s);
uvt_status uvt_serve_stats_get(uvt_server s, uvt_serve_stats *st);
/* Device bytes a request with host operands holds while in flight (A, B, C
 * and the split workspace); device-resident operands are not charged. */
uvt_status uvt_serve_request_bytes(const uvt_serve_config *cfg, int64_t n, int64_t m, int32_t prec,
                                   size_t *bytes);
uvt_status uvt_serve_destroy(uvt_server s);

/* Row sums of |C| into d_rowsum[n] (device). */
uvt_status uvt_row_abs_sums(uvt_handle h, int64_t n, int64_t m, const double *dC, int64_t ldc,
                            double *d_rowsum, void *stream);
/* max_i sum_j |C[i][j]|; synchronous, result in *norm (host). */
uvt_status uvt_inf_norm(uvt_handle h, int64_t n, int64_t m, const double *dC, int64_t ldc,
                        double *norm, void *stream);

/* Inputs.  Closed form of the reference (A lower with zeros above the
 * diagonal, B n x m with value index row*n + column). */
uvt_status uvt_fill_closed_form(uvt_handle h, int64_t n, int64_t m, double *dA, int64_t lda,
                                double *dB, int64_t ldb, void *stream);
/* Seeded inputs: SplitMix64 counter generator, U[-1,1), A's diagonal +2.0,
 * spread!=0 scales each entry by 2^e, e uniform in {-8..8}.  Bit-identical
 * to the host generator of the same name. Either pointer may be NULL. */
uvt_status uvt_fill_seeded(uvt_handle h, int64_t n, int64_t m, uint64_t seed, int32_t spread,
                           double *dA, int64_t lda, double *dB, int64_t ldb, void *stream);

/* ------------------------------------------------------------------ */
/* Small runtime helpers for callers without a HIP toolchain (ctypes,  */
/* the CLI).                                                           */
/* ------------------------------------------------------------------ */
uvt_status uvt_malloc(uvt_handle h, size_t bytes, void **dptr);
uvt_status uvt_free(uvt_handle h, void *dptr);
/* kind: 0 host->device, 1 device->host, 2 device->device; synchronous.
 * Host copies of 1 MiB or more, and strided host copies, go through two
 * 16 MiB pinned buffers that the handle allocates on first use and keeps
 * (rows packed on the host, a chunk in flight while the next is packed):
 * host memory is read or written as plain memory, never registered. */
uvt_status uvt_memcpy(uvt_handle h, void *dst, const void *src, size_t bytes, int32_t kind);
/* Strided copy of `height` rows of `width` bytes (pitches in bytes). */
uvt_status uvt_memcpy_2d(uvt_handle h, void *dst, size_t dpitch, const void *src, size_t spitch,
                         size_t width, size_t height, int32_t kind);
uvt_status uvt_memset(uvt_handle h, void *dptr, int32_t value, size_t bytes);
uvt_status uvt_stream_create(uvt_handle h, void **stream);
uvt_status uvt_stream_destroy(uvt_handle h, void *stream);
uvt_status uvt_stream_sync(uvt_handle h, void *stream);
uvt_status uvt_event_create(uvt_handle h, void **event);
uvt_status uvt_event_destroy(uvt_handle h, void *event);
uvt_status uvt_event_record(uvt_handle h, void *event, void *stream);
/* Milliseconds between two recorded events (synchronises on `end`). */
uvt_status uvt_event_elapsed(uvt_handle h, void *start, void *end, float *ms);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
