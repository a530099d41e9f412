// uvt_runtime.hpp -- handle layout and HIP error plumbing shared by the HIP
// translation units of libuvt.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include <vector>

#include "uvt_internal.h"

struct uvt_handle_s {
    int device;
    int cus;
    size_t vram;
    void* work;        // grows on demand (norm row sums, BLAS-form packing)
    size_t work_bytes;
    void* scratch;     // this device's fixed scratch array (uvt_dtrmm_noalloc)
    // pinned bounce buffers of strided host copies (uvt_memcpy_2d), on first use
    void* pin[2];
    hipEvent_t pin_done[2];
    bool pin_used[2];
};

constexpr int kMaxDevices = 64;

// Stores the handle's device's address of the fixed scratch array of
// uvt_dtrmm_noalloc in h->scratch (looked up once per device); called by
// uvt_create with the device current.
UVT_INTERNAL uvt_status uvt_noalloc_setup(uvt_handle h);

// Returns UVT_ERR_HIP from the enclosing function when a HIP call fails,
// clearing HIP's sticky last error so later calls start clean.
#define UVT_HIP(call)                                 \
    do {                                              \
        hipError_t uvt_e_ = (call);                   \
        if (uvt_e_ != hipSuccess) {                   \
            (void)hipGetLastError();                  \
            return UVT_ERR_HIP;                       \
        }                                             \
    } while (0)

UVT_INTERNAL uvt_status uvt_handle_workspace(uvt_handle h, size_t bytes, void** p);
// FP64 C = tril(A) * B as one trapezoid walk with the given tile shape,
// tile coordinates from the workgroup index (no plan, nothing to upload).
UVT_INTERNAL uvt_status uvt_launch_trapezoid_f64(int32_t cfg, int64_t n, int64_t m, const double* A, int64_t lda,
                                    const double* B, int64_t ldb, double* C, int64_t ldc,
                                    hipStream_t s);
// The same walk split stream-K (uvt_sk.hpp) over at most `workgroups`
// workgroups (in column groups where there are at least 8 of them), then
// its fixup launch; parts holds uvt_sk_parts_bytes(cfg, workgroups) bytes.
// Each result is the same on every run with the same arguments and GPU.
UVT_INTERNAL uvt_status uvt_launch_trapezoid_sk_f64(int32_t cfg, int64_t n, int64_t m, const double* A,
                                                    int64_t lda, const double* B, int64_t ldb, double* C,
                                                    int64_t ldc, int64_t workgroups, double* parts,
                                                    hipStream_t s);
// Workgroups for the stream-K walk of an n x n by n x m product in cfg tiles
// on a GPU of `cus` compute units (one per workgroup slot), or 0 for whole
// tiles: stream-K runs 128x128 tiles of at least 2e9 flops whose whole-tile
// span is at least 1.3x an even share of the work (narrow B: few, deep tiles).
UVT_INTERNAL int64_t uvt_sk_workgroups(int32_t cfg, int64_t n, int64_t m, int cus);
UVT_INTERNAL size_t uvt_sk_parts_bytes(int32_t cfg, int64_t workgroups);
// The most partial-tile bytes any uvt_sk_workgroups result needs on `cus` CUs.
UVT_INTERNAL size_t uvt_sk_max_parts_bytes(int cus);
// FP64 C[0:M, 0:N] = A[0:M, 0:K] * B[0:K, 0:N] (no triangle), tile coordinates
// from the workgroup index.
UVT_INTERNAL uvt_status uvt_launch_gemm_f64(int32_t cfg, int64_t M, int64_t N, int64_t K, const double* A,
                               int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                               hipStream_t s);

// Tile-shape table written by uvt_autotune_f64: entries matched in order,
// the first with n <= n_max and m <= m_max gives the shape.
struct uvt_tune_entry {
    int64_t n_max, m_max;
    int cfg;
};
UVT_INTERNAL bool uvt_tune_load(const char* path, std::vector<uvt_tune_entry>& out);
// Shape for (n, m); the built-in default when no entry matches.
UVT_INTERNAL int uvt_tune_pick(const std::vector<uvt_tune_entry>& table, int64_t n, int64_t m);

// Where a split-class problem (error <= 1e-12 acceptable) runs on the int8
// path rather than in FP64, which is always at least as accurate.  Measured on
// MI355X with 12 moduli (FP64 time / split time):
//   alone (serving): the int8 path is faster from m >= 1024 once n >= 4096
//     (1.14x - 1.98x), and from m >= 256 at n >= 6144 only by 0-6 %, so the
//     line is drawn at m >= 512 there; narrow B (m < 256) and n < 4096 never
//     pay for the residue and reconstruction passes;
//   in a batch (throughput, persistent FP64 queues): FP64 wins below n = 8192.
inline bool uvt_split_faster_alone(int64_t n, int64_t m)
{
    return (n >= 4096 && m >= 1024) || (n >= 6144 && m >= 512);
}
inline bool uvt_split_faster_in_batch(int64_t n, int64_t m) { return n >= 8192 && m >= 256; }
