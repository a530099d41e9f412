// uvt_api.hip -- C ABI of libuvt: device handle, FP64 TRMM launches, norm,
// input generators and small runtime helpers.  Declarations: include/uvt.h.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "uvt_f64_tile.hpp"
#include "uvt_internal.h"
#include "uvt_runtime.hpp"

using namespace uvt;

// ---------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------

namespace {  // kernels: internal to this file

template <int TM, int TN, bool VEC>
__global__ __launch_bounds__((64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N),
                             (TileWaves<TM, TN>::WAVES_PER_EU))
void k_trmm_tiles_f64(const uvt_tile* __restrict__ tiles, const double* __restrict__ A, int64_t lda,
                      const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
                      int64_t ldc, int64_t m)
{
    using W = TileWaves<TM, TN>;
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<TM, TN, W::M, W::N>::SMEM_DOUBLES];
    const uvt_tile t = tiles[blockIdx.x];
    f64_tile<TM, TN, VEC, W::M, W::N>(t.r0, t.rows, t.c0, t.k_lo, t.k_hi, t.flags, gbl(A), lda,
                                      gbl(B), ldb, gbl(C), ldc, m, smem);
}

// The trapezoid walk without a tile list: workgroup b takes row tile
// nrt-1-b/nct (longest reduction first) and column tile b%nct.
template <int TM, int TN, bool VEC>
__global__ __launch_bounds__((64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N),
                             (TileWaves<TM, TN>::WAVES_PER_EU))
void k_trmm_trapezoid_f64(int64_t n, int64_t m, int nrt, int nct, const double* __restrict__ A,
                          int64_t lda, const double* __restrict__ B, int64_t ldb,
                          double* __restrict__ C, int64_t ldc)
{
    using W = TileWaves<TM, TN>;
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<TM, TN, W::M, W::N>::SMEM_DOUBLES];
    const int rt = nrt - 1 - (int)(blockIdx.x / nct), ct = (int)(blockIdx.x % nct);
    const int r0 = rt * TM;
    const int rows = (int)(n - r0 < TM ? n - r0 : TM);
    f64_tile<TM, TN, VEC, W::M, W::N>(r0, rows, ct * TN, 0, r0 + rows, kTileLower, gbl(A), lda,
                                      gbl(B), ldb, gbl(C), ldc, m, smem);
}

// One wave per row: sum_j |C[i][j]|, lanes strided over columns, then a
// fixed butterfly, so the order (and the result) is the same on every run.
__global__ __launch_bounds__(256) void k_row_abs_sums(int64_t n, int64_t m,
                                                      const double* __restrict__ C, int64_t ldc,
                                                      double* __restrict__ out)
{
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double* c = C + row * ldc;
    double s = 0.0;
    for (int64_t j = lane; j < m; j += 64) s += fabs(c[j]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[row] = s;
}

__global__ __launch_bounds__(1024) void k_max_reduce(int64_t n, const double* __restrict__ v,
                                                     double* __restrict__ out)
{
    __shared__ double part[16];
    double x = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) x = fmax(x, v[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x < 64) {
        x = threadIdx.x < 16 ? part[threadIdx.x] : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
        if (threadIdx.x == 0) out[0] = x;
    }
}

// mat_consecutive_values with MININUMS=1 (UVaTRMM/src/mm_triangular_common.c:1782-1805):
// value index row*n + column modulo VALUES_LIMIT=2341, plus one, reciprocal.
__global__ void k_fill_closed(int64_t n, int64_t rows, int64_t cols, int lower,
                              double* __restrict__ X, int64_t ldx)
{
    for (int64_t i = blockIdx.y; i < rows; i += gridDim.y)
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cols;
         j += (int64_t)gridDim.x * blockDim.x) {
        double v;
        if (lower && j > i) {
            v = 0.0;
        } else {
            const int64_t idx = i * n + j;
            v = 1.0 / (double)(idx % 2341 + 1);
        }
        X[i * ldx + j] = v;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Seeded generator, bit-identical to the host one: SplitMix64 (Steele, Lea,
// Flood 2014) applied to key + (idx+1)*gamma; 53 random bits -> U[-1,1);
// optional +2 on the diagonal; optional exact 2^e scaling, e in {-8..8}.
__global__ void k_fill_seeded(int64_t rows, int64_t cols, uint64_t key, int stream_is_a,
                              int spread, double* __restrict__ X, int64_t ldx)
{
    const uint64_t gamma = 0x9E3779B97F4A7C15ULL;
    for (int64_t i = blockIdx.y; i < rows; i += gridDim.y)
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < cols;
         j += (int64_t)gridDim.x * blockDim.x) {
        double v;
        if (stream_is_a && j > i) {
            v = 0.0;
        } else {
            const uint64_t idx = (uint64_t)(i * cols + j);
            const uint64_t z = mix64(key + (idx + 1) * gamma);
            const double u = (double)(z >> 11) * 0x1.0p-53;
            v = 2.0 * u - 1.0;
            if (stream_is_a && i == j) v = v + 2.0;
            if (spread) {
                const int e = (int)(mix64(z) % 17) - 8;
                v = v * __longlong_as_double((long long)((uint64_t)(1023 + e) << 52));
            }
        }
        X[i * ldx + j] = v;
    }
}

}  // namespace

// ---------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------

extern "C" uvt_status uvt_create(int device, uvt_handle* out)
{
    if (!out) return UVT_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return UVT_ERR_UNSUPPORTED;
    if (device < 0 || device >= count) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    UVT_HIP(hipGetDeviceProperties(&prop, device));
    uvt_handle h = (uvt_handle)calloc(1, sizeof(struct uvt_handle_s));
    if (!h) return UVT_ERR_NOMEM;
    h->device = device;
    h->cus = prop.multiProcessorCount;
    h->vram = prop.totalGlobalMem;
    const uvt_status st = uvt_noalloc_setup(h);
    if (st != UVT_OK) {
        free(h);
        return st;
    }
    *out = h;
    return UVT_OK;
}

extern "C" uvt_status uvt_destroy(uvt_handle h)
{
    if (!h) return UVT_ERR_ARG;
    (void)hipSetDevice(h->device);
    if (h->work) (void)hipFree(h->work);
    free(h);
    return UVT_OK;
}

extern "C" int32_t uvt_device_cus(uvt_handle h) { return h ? h->cus : 0; }

uvt_status uvt_handle_workspace(uvt_handle h, size_t bytes, void** p)
{
    if (h->work_bytes < bytes) {
        if (h->work) {
            UVT_HIP(hipDeviceSynchronize());
            UVT_HIP(hipFree(h->work));
            h->work = nullptr;
            h->work_bytes = 0;
        }
        if (hipMalloc(&h->work, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return UVT_ERR_NOMEM;
        }
        h->work_bytes = bytes;
    }
    *p = h->work;
    return UVT_OK;
}

// ---------------------------------------------------------------------
// Plans on the device
// ---------------------------------------------------------------------

extern "C" void uvt_plan_release_device(uvt_plan plan)
{
    if (!plan) return;
    for (int ph = 0; ph < 2; ++ph) {
        if (plan->dev_tiles[ph]) {
            (void)hipSetDevice(plan->dev_id);
            (void)hipFree(plan->dev_tiles[ph]);
            plan->dev_tiles[ph] = nullptr;
        }
    }
    plan->dev_id = -1;
}

static uvt_status plan_upload(uvt_handle h, uvt_plan plan)
{
    if (plan->dev_id == h->device) return UVT_OK;
    uvt_plan_release_device(plan);
    for (int ph = 0; ph < plan->phases; ++ph) {
        size_t bytes = (size_t)plan->ntiles[ph] * sizeof(uvt_tile);
        if (!bytes) continue;
        if (hipMalloc(&plan->dev_tiles[ph], bytes) != hipSuccess) {
            (void)hipGetLastError();
            uvt_plan_release_device(plan);
            return UVT_ERR_NOMEM;
        }
        UVT_HIP(hipMemcpy(plan->dev_tiles[ph], plan->tiles[ph], bytes, hipMemcpyHostToDevice));
    }
    plan->dev_id = h->device;
    return UVT_OK;
}

// ---------------------------------------------------------------------
// FP64 TRMM
// ---------------------------------------------------------------------

template <int TM, int TN>
static void launch_tiles(const uvt_tile* dtiles, int64_t ntiles, bool vec, const double* A,
                         int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                         int64_t m, hipStream_t s)
{
    dim3 grid((unsigned)ntiles), block(64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N);
    if (vec)
        hipLaunchKernelGGL((k_trmm_tiles_f64<TM, TN, true>), grid, block, 0, s, dtiles, A, lda, B,
                           ldb, C, ldc, m);
    else
        hipLaunchKernelGGL((k_trmm_tiles_f64<TM, TN, false>), grid, block, 0, s, dtiles, A, lda, B,
                           ldb, C, ldc, m);
}

uvt_status uvt_launch_tiles_f64(int32_t cfg, const uvt_tile* dtiles, int64_t ntiles,
                                const double* A, int64_t lda, const double* B, int64_t ldb,
                                double* C, int64_t ldc, int64_t m, hipStream_t s)
{
    if (ntiles <= 0) return UVT_OK;
    if (ntiles > 0x7FFFFFFF) return UVT_ERR_SIZE;
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 2 == 0 &&
                     ldb % 2 == 0;
    if (cfg == UVT_CFG_64x64)
        launch_tiles<64, 64>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, s);
    else if (cfg == UVT_CFG_128x32)
        launch_tiles<128, 32>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, s);
    else
        launch_tiles<128, 128>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, s);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

template <int TM, int TN>
static void launch_trapezoid(bool vec, int64_t n, int64_t m, const double* A, int64_t lda,
                             const double* B, int64_t ldb, double* C, int64_t ldc, hipStream_t s)
{
    const int nrt = (int)((n + TM - 1) / TM), nct = (int)((m + TN - 1) / TN);
    const dim3 grid((unsigned)((int64_t)nrt * nct)), block(64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N);
    if (vec)
        hipLaunchKernelGGL((k_trmm_trapezoid_f64<TM, TN, true>), grid, block, 0, s, n, m, nrt, nct, A,
                           lda, B, ldb, C, ldc);
    else
        hipLaunchKernelGGL((k_trmm_trapezoid_f64<TM, TN, false>), grid, block, 0, s, n, m, nrt, nct,
                           A, lda, B, ldb, C, ldc);
}

uvt_status uvt_launch_trapezoid_f64(int32_t cfg, int64_t n, int64_t m, const double* A, int64_t lda,
                                    const double* B, int64_t ldb, double* C, int64_t ldc,
                                    hipStream_t s)
{
    if (n <= 0 || m <= 0 || lda < n || ldb < m || ldc < m) return UVT_ERR_ARG;
    if (n > 0x3FFFFFFF || m > 0x3FFFFFFF) return UVT_ERR_SIZE;
    const int64_t tm = uvt_cfg_tile_m(cfg), tn = uvt_cfg_tile_n(cfg);
    if (((n + tm - 1) / tm) * ((m + tn - 1) / tn) > 0x7FFFFFFF) return UVT_ERR_SIZE;
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 2 == 0 &&
                     ldb % 2 == 0;
    if (cfg == UVT_CFG_64x64) launch_trapezoid<64, 64>(vec, n, m, A, lda, B, ldb, C, ldc, s);
    else if (cfg == UVT_CFG_128x32) launch_trapezoid<128, 32>(vec, n, m, A, lda, B, ldb, C, ldc, s);
    else launch_trapezoid<128, 128>(vec, n, m, A, lda, B, ldb, C, ldc, s);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

extern "C" uvt_status uvt_trmm_f64(uvt_handle h, uvt_plan plan, const double* dA, int64_t lda,
                                   const double* dB, int64_t ldb, double* dC, int64_t ldc,
                                   void* stream)
{
    if (!h || !plan || !dA || !dB || !dC) return UVT_ERR_ARG;
    if (lda < plan->n || ldb < plan->m || ldc < plan->m) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    uvt_status st = plan_upload(h, plan);
    if (st != UVT_OK) return st;
    hipStream_t s = (hipStream_t)stream;
    for (int ph = 0; ph < plan->phases; ++ph) {
        st = uvt_launch_tiles_f64(plan->cfg, (const uvt_tile*)plan->dev_tiles[ph], plan->ntiles[ph],
                                  dA, lda, dB, ldb, dC, ldc, plan->m, s);
        if (st != UVT_OK) return st;
    }
    return UVT_OK;
}

// ---------------------------------------------------------------------
// Norm
// ---------------------------------------------------------------------

extern "C" uvt_status uvt_row_abs_sums(uvt_handle h, int64_t n, int64_t m, const double* dC,
                                       int64_t ldc, double* d_rowsum, void* stream)
{
    if (!h || !dC || !d_rowsum || n <= 0 || m <= 0 || ldc < m) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    const int64_t blocks = (n + 3) / 4;
    if (blocks > 0x7FFFFFFF) return UVT_ERR_SIZE;
    hipLaunchKernelGGL(k_row_abs_sums, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       n, m, dC, ldc, d_rowsum);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

extern "C" uvt_status uvt_inf_norm(uvt_handle h, int64_t n, int64_t m, const double* dC,
                                   int64_t ldc, double* norm, void* stream)
{
    if (!h || !dC || !norm || n <= 0 || m <= 0 || ldc < m) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    void* w = nullptr;
    uvt_status st = uvt_handle_workspace(h, (size_t)(n + 1) * sizeof(double), &w);
    if (st != UVT_OK) return st;
    double* rows = (double*)w;
    st = uvt_row_abs_sums(h, n, m, dC, ldc, rows, stream);
    if (st != UVT_OK) return st;
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(1024), 0, (hipStream_t)stream, n, rows,
                       rows + n);
    UVT_HIP(hipGetLastError());
    UVT_HIP(hipMemcpyAsync(norm, rows + n, sizeof(double), hipMemcpyDeviceToHost,
                           (hipStream_t)stream));
    UVT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return UVT_OK;
}

// ---------------------------------------------------------------------
// Inputs
// ---------------------------------------------------------------------

static dim3 fill_grid(int64_t rows, int64_t cols)
{
    int64_t bx = (cols + 255) / 256;
    if (bx > 64) bx = 64;
    return dim3((unsigned)(bx < 1 ? 1 : bx), (unsigned)(rows < 32768 ? rows : 32768));
}

extern "C" uvt_status uvt_fill_closed_form(uvt_handle h, int64_t n, int64_t m, double* dA,
                                           int64_t lda, double* dB, int64_t ldb, void* stream)
{
    if (!h || n <= 0 || m <= 0) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    if (dA) {
        if (lda < n) return UVT_ERR_ARG;
        hipLaunchKernelGGL(k_fill_closed, fill_grid(n, n), dim3(256), 0, s, n, n, n, 1, dA, lda);
    }
    if (dB) {
        if (ldb < m) return UVT_ERR_ARG;
        hipLaunchKernelGGL(k_fill_closed, fill_grid(n, m), dim3(256), 0, s, n, n, m, 0, dB, ldb);
    }
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

static uint64_t host_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

extern "C" uvt_status uvt_fill_seeded(uvt_handle h, int64_t n, int64_t m, uint64_t seed,
                                      int32_t spread, double* dA, int64_t lda, double* dB,
                                      int64_t ldb, void* stream)
{
    if (!h || n <= 0 || m <= 0) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    const uint64_t gamma = 0x9E3779B97F4A7C15ULL;
    if (dA) {
        if (lda < n) return UVT_ERR_ARG;
        hipLaunchKernelGGL(k_fill_seeded, fill_grid(n, n), dim3(256), 0, s, n, n,
                           host_mix64(seed ^ (1ULL * gamma)), 1, (int)spread, dA, lda);
    }
    if (dB) {
        if (ldb < m) return UVT_ERR_ARG;
        hipLaunchKernelGGL(k_fill_seeded, fill_grid(n, m), dim3(256), 0, s, n, m,
                           host_mix64(seed ^ (2ULL * gamma)), 0, (int)spread, dB, ldb);
    }
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

// ---------------------------------------------------------------------
// Runtime helpers
// ---------------------------------------------------------------------

extern "C" uvt_status uvt_malloc(uvt_handle h, size_t bytes, void** dptr)
{
    if (!h || !dptr) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        *dptr = nullptr;
        return UVT_ERR_NOMEM;
    }
    return UVT_OK;
}

extern "C" uvt_status uvt_free(uvt_handle h, void* dptr)
{
    if (!h) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    if (dptr) UVT_HIP(hipFree(dptr));
    return UVT_OK;
}

extern "C" uvt_status uvt_memcpy(uvt_handle h, void* dst, const void* src, size_t bytes,
                                 int32_t kind)
{
    if (!h || (!dst && bytes) || (!src && bytes)) return UVT_ERR_ARG;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                      : kind == 1 ? hipMemcpyDeviceToHost
                                  : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    if (bytes) UVT_HIP(hipMemcpy(dst, src, bytes, k));
    return UVT_OK;
}

extern "C" uvt_status uvt_memcpy_2d(uvt_handle h, void* dst, size_t dpitch, const void* src,
                                    size_t spitch, size_t width, size_t height, int32_t kind)
{
    if (!h || kind < 0 || kind > 2) return UVT_ERR_ARG;
    if (!width || !height) return UVT_OK;
    if (!dst || !src || dpitch < width || spitch < width) return UVT_ERR_ARG;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                            : kind == 1 ? hipMemcpyDeviceToHost
                                        : hipMemcpyDeviceToDevice;
    UVT_HIP(hipSetDevice(h->device));
    UVT_HIP(hipMemcpy2D(dst, dpitch, src, spitch, width, height, k));
    return UVT_OK;
}

extern "C" uvt_status uvt_memset(uvt_handle h, void* dptr, int32_t value, size_t bytes)
{
    if (!h || (!dptr && bytes)) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    if (bytes) UVT_HIP(hipMemset(dptr, value, bytes));
    return UVT_OK;
}

extern "C" uvt_status uvt_stream_create(uvt_handle h, void** stream)
{
    if (!h || !stream) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    hipStream_t s;
    UVT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return UVT_OK;
}

extern "C" uvt_status uvt_stream_destroy(uvt_handle h, void* stream)
{
    if (!h) return UVT_ERR_ARG;
    if (stream) UVT_HIP(hipStreamDestroy((hipStream_t)stream));
    return UVT_OK;
}

extern "C" uvt_status uvt_stream_sync(uvt_handle h, void* stream)
{
    if (!h) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    UVT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return UVT_OK;
}

extern "C" uvt_status uvt_event_create(uvt_handle h, void** event)
{
    if (!h || !event) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    hipEvent_t e;
    UVT_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return UVT_OK;
}

extern "C" uvt_status uvt_event_destroy(uvt_handle h, void* event)
{
    if (!h) return UVT_ERR_ARG;
    if (event) UVT_HIP(hipEventDestroy((hipEvent_t)event));
    return UVT_OK;
}

extern "C" uvt_status uvt_event_record(uvt_handle h, void* event, void* stream)
{
    if (!h || !event) return UVT_ERR_ARG;
    UVT_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return UVT_OK;
}

extern "C" uvt_status uvt_event_elapsed(uvt_handle h, void* start, void* end, float* ms)
{
    if (!h || !start || !end || !ms) return UVT_ERR_ARG;
    UVT_HIP(hipEventSynchronize((hipEvent_t)end));
    UVT_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return UVT_OK;
}
