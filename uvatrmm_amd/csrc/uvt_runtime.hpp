// uvt_runtime.hpp -- handle layout and HIP error plumbing shared by the HIP
// translation units of libuvt.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "uvt_internal.h"

struct uvt_handle_s {
    int device;
    int cus;
    size_t vram;
    void* work;        // grows on demand (norm row sums, split slices)
    size_t work_bytes;
};

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

uvt_status uvt_handle_workspace(uvt_handle h, size_t bytes, void** p);
uvt_status uvt_launch_tiles_f64(int32_t cfg, const uvt_tile* dtiles, int64_t ntiles,
                                const double* A, int64_t lda, const double* B, int64_t ldb,
                                double* C, int64_t ldc, int64_t m, hipStream_t s);
