// uvt_api.hip -- C ABI of libuvt: device handle, FP64 TRMM launches, norm,
// input generators and small runtime helpers.  Declarations: include/uvt.h.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "uvt_f64_tile.hpp"
#include "uvt_internal.h"
#include "uvt_runtime.hpp"
#include "uvt_sk.hpp"

using namespace uvt;

// ---------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------

namespace {  // kernels: internal to this file

// One workgroup of a plan launch: a plan tile (or a piece of its k range)
// and where its product goes: C (part < 0), or partial tile `part`, which a
// later k_add_parts launch adds to C.
struct LaunchTile {
    uvt_tile t;
    int32_t part;
};
// C[r0 + r][c0 + c] += partial tiles first, first + 1, ..., first + count - 1,
// in that order, for rows r < rows of a TM x TN tile.
struct PartSum {
    int32_t r0, rows, c0, first, count;
};

// PARTS: the records are LaunchTiles (some write partial tiles); else plain
// uvt_tiles, every one writing C.
template <int TM, int TN, bool VEC, bool PARTS>
__global__ __launch_bounds__((64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N),
                             (TileWaves<TM, TN>::WAVES_PER_EU))
void k_trmm_tiles_f64(const void* __restrict__ tiles, const double* __restrict__ A, int64_t lda,
                      const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
                      int64_t ldc, int64_t m, double* __restrict__ parts)
{
    using W = TileWaves<TM, TN>;
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<TM, TN, W::M, W::N>::SMEM_DOUBLES];
    if constexpr (PARTS) {
        const LaunchTile lt = static_cast<const LaunchTile*>(tiles)[blockIdx.x];
        const uvt_tile& t = lt.t;
        UVT_GLOBAL double* part = lt.part >= 0 ? gbl(parts) + (int64_t)lt.part * (TM * TN) : nullptr;
        f64_tile<TM, TN, VEC, W::M, W::N>(t.r0, t.rows, t.c0, t.k_lo, t.k_hi, t.flags, gbl(A), lda,
                                          gbl(B), ldb, gbl(C), ldc, m, smem, -1, part);
    } else {
        const uvt_tile t = static_cast<const uvt_tile*>(tiles)[blockIdx.x];
        f64_tile<TM, TN, VEC, W::M, W::N>(t.r0, t.rows, t.c0, t.k_lo, t.k_hi, t.flags, gbl(A), lda,
                                          gbl(B), ldb, gbl(C), ldc, m, smem);
    }
}

// Adds the partial tiles of each record to C in the record's order (one
// workgroup per record; records cover disjoint parts of C), so the result
// does not depend on which workgroup finished first.
__global__ __launch_bounds__(256) void k_add_parts(const PartSum* __restrict__ sums, int tn, int tile,
                                                   const double* __restrict__ parts, double* __restrict__ C,
                                                   int64_t ldc, int64_t m)
{
    const PartSum ps = sums[blockIdx.x];
    const UVT_GLOBAL double* P = gbl(parts) + (int64_t)ps.first * tile;
    for (int e = threadIdx.x; e < ps.rows * tn; e += 256) {
        const int r = e / tn, c = e % tn;
        if ((int64_t)ps.c0 + c >= m) continue;
        UVT_GLOBAL double* p = gbl(C) + (int64_t)(ps.r0 + r) * ldc + (ps.c0 + c);
        double x = *p;
        for (int q = 0; q < ps.count; ++q) x += P[(int64_t)q * tile + e];
        *p = x;
    }
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

// The trapezoid walk split stream-K (uvt_sk.hpp): workgroup w of group
// b = w / ncg runs the pieces of panels [b P / G, (b+1) P / G) one after
// another, in column tile w % ncg of each piece's row tile when the launch
// runs column groups; its first piece writes partial tile w if it starts
// inside its tile, every other piece starts at k = 0 and writes C.
template <int TM, int TN, bool VEC>
__global__ __launch_bounds__((64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N),
                             (TileWaves<TM, TN>::WAVES_PER_EU))
void k_trmm_sk_f64(SkGeo g, const double* __restrict__ A, int64_t lda, const double* __restrict__ B,
                   int64_t ldb, double* __restrict__ C, int64_t ldc, double* __restrict__ parts)
{
    using W = TileWaves<TM, TN>;
    static_assert(F64Cfg<TM, TN, W::M, W::N>::BK == kSkPanel, "stream-K panels are the kernel's k panels");
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<TM, TN, W::M, W::N>::SMEM_DOUBLES];
    // The piece arithmetic (64-bit divisions) runs in vector registers; its
    // results are the same in every lane and go back to scalar registers, so
    // they do not take vector registers from the tile.
    auto uni = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
    auto uni64 = [&](int64_t v) {
        return (int64_t)(((uint64_t)(uint32_t)uni((int)(v >> 32)) << 32) | (uint32_t)uni((int)v));
    };
    // The 8 XCDs take contiguous runs of the launch, so the workgroups that
    // share an XCD's L2 work on neighbouring tiles, or (column groups) on the
    // same rows of A at the same k.
    const int64_t w = sk_xcd_order(blockIdx.x, sk_launch(g));
    const int64_t b = w / g.ncg;
    const int cx = (int)(w % g.ncg) * TN;  // the member's column tile (0 without groups)
    int64_t x = uni64(sk_start(g, b));
    const int64_t xe = uni64(sk_start(g, b + 1));
    UVT_GLOBAL double* part = gbl(parts) + w * (TM * TN);
    bool first = true;
    while (x < xe) {
        const SkPiece s = sk_piece(g, x, xe);
        if (!first) __syncthreads();  // the LDS tile is reused by the next piece
        // the thread index, opaque per piece: the tile's per-thread offsets are
        // then computed inside each piece instead of held (and spilled) across
        // the piece loop
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        f64_tile<TM, TN, VEC, W::M, W::N, kSkPanel, false>(uni(s.r0), uni(s.rows), uni(s.c0 + cx), uni(s.k_lo), uni(s.k_hi),
                                          kTileLower, gbl(A), lda, gbl(B), ldb, gbl(C), ldc, g.m, smem,
                                          tid, uni(s.p0 > 0) ? part : nullptr);
        first = false;
        x = uni64(s.next);
    }
}

// After k_trmm_sk_f64: for each tile split between workgroups, the
// workgroup b whose first piece is the tile's first partial tile adds
// partial tiles b, b+1, ... (k order) to C, which holds the k = 0 piece; with
// column groups, group b's member in each column tile does so in that column
// tile.  (gridDim.y workgroups per split tile, each over its share of the rows.)
template <int TM, int TN>
__global__ __launch_bounds__(256) void k_sk_fixup(SkGeo g, const double* __restrict__ parts,
                                                  double* __restrict__ C, int64_t ldc)
{
    const int64_t w = blockIdx.x;  // partial tile w: group b, column tile cx
    const int64_t b = w / g.ncg, cx = w % g.ncg;
    if (b == 0) return;
    const int64_t x = sk_start(g, b);
    const int j = sk_row(g, x);
    const int64_t d = sk_depth(g, j), off = x - sk_before(g, j), p0 = off % d;
    if (p0 == 0) return;                     // b starts its tile: it wrote C
    const int64_t tile0 = x - p0;
    if (sk_start(g, b - 1) > tile0) return;  // b - 1 holds this tile's first partial tile
    int64_t cnt = 1;
    while (b + cnt < g.G && sk_start(g, b + cnt) < tile0 + d) ++cnt;
    const int r0 = j * TM, rows = (int)(g.n - r0 < TM ? g.n - r0 : TM);
    const int64_t c0 = (off / d + cx) * TN;
    const UVT_GLOBAL double* P = gbl(parts) + w * (TM * TN);
    const int64_t step = g.ncg * (int64_t)(TM * TN);  // group b + q's partial tile in this column tile
    const int per = (rows + (int)gridDim.y - 1) / (int)gridDim.y;
    const int e_lo = (int)blockIdx.y * per * TN;
    const int e_hi = min(rows, ((int)blockIdx.y + 1) * per) * TN;
    for (int e = e_lo + (int)threadIdx.x; e < e_hi; e += 256) {
        const int r = e / TN, c = e % TN;
        if (c0 + c >= g.m) continue;
        UVT_GLOBAL double* p = gbl(C) + (int64_t)(r0 + r) * ldc + (c0 + c);
        double v = *p;
        for (int64_t q = 0; q < cnt; ++q) v += P[q * step + e];
        *p = v;
    }
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
    for (int i = 0; i < 2; ++i) {
        if (h->pin_done[i]) (void)hipEventSynchronize(h->pin_done[i]);
        if (h->pin[i]) (void)hipHostFree(h->pin[i]);
        if (h->pin_done[i]) (void)hipEventDestroy(h->pin_done[i]);
    }
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
    void** bufs[4] = {&plan->dev_launch[0], &plan->dev_launch[1], &plan->dev_sums, &plan->dev_parts};
    for (void** b : bufs) {
        if (*b) {
            (void)hipSetDevice(plan->dev_id);
            (void)hipFree(*b);
            *b = nullptr;
        }
    }
    plan->nlaunch[0] = plan->nlaunch[1] = plan->nsums = plan->nparts = plan->sk_grid = 0;
    plan->dev_id = -1;
}

namespace {

constexpr size_t kMaxPartBytes = (size_t)64 << 20;  // partial tiles of one plan
// Below this many flops a plan's launch time, not one tile's depth, bounds
// it, and the plan keeps whole tiles.
constexpr double kSplitMinFlops = 2e9;
// Stream-K walks run column groups (uvt_sk.hpp) where there are at least 8 of
// them.  Measured against one sequence over every tile (three alternating
// passes, one box): 8192 x 1024 1.20 -> 1.18 ms, 8165 x 939 1.30 -> 1.29 ms,
// 6772 x 1548 and 5000 x 2000 within 1 % either way.
constexpr bool kSkGroupCols = true;

int blocks_per_cu(int32_t cfg)
{
    if (cfg == UVT_CFG_256x128) return TileWaves<256, 128>::BLOCKS_PER_CU;
    if (cfg == UVT_CFG_64x64) return TileWaves<64, 64>::BLOCKS_PER_CU;
    if (cfg == UVT_CFG_128x32) return TileWaves<128, 32>::BLOCKS_PER_CU;
    return TileWaves<128, 128>::BLOCKS_PER_CU;
}

// The launches of a plan on a GPU of `cus` compute units: its phases, one
// launch each, with whole tiles, except for a trapezoid plan whose longest
// reduction is much longer than an even share of the work would give a
// workgroup slot (narrow B, few tiles: the launch then lasts as long as its
// deepest tile).  Such a plan cuts every reduction longer than about half
// that share, kc, into pieces that run at once: the first piece writes C,
// the others write partial tiles, which k_add_parts adds to C afterwards in
// k order.  Measured at 9000 x 250 (the deepest tile 2.7x the share): 0.73
// -> 0.45 ms; at 4096 x 4096 (0.97x) cutting cost 7 %, between 0.9x and
// 1.6x it changed -1 to +7 %, and cutting the boxes of the N=2048 boxes
// plan (1.5x) cost 11 %, so boxes plans and plans below 1.5x keep whole
// tiles (and the kernel without partial tiles).  (A trapezoid plan of
// 128x128 tiles whose whole tiles would not finish together runs stream-K
// instead, decided before this in plan_upload; so cuts serve the smaller
// tile shapes, where they were faster than stream-K.)
void build_launch(const uvt_plan_s& pl, int cus, std::vector<LaunchTile> launch[2],
                  std::vector<PartSum>& sums, int64_t& nparts)
{
    for (int ph = 0; ph < 2; ++ph) launch[ph].clear();
    sums.clear();
    nparts = 0;
    for (int ph = 0; ph < pl.phases; ++ph)
        for (int64_t i = 0; i < pl.ntiles[ph]; ++i) launch[ph].push_back(LaunchTile{pl.tiles[ph][i], -1});
    if (pl.phases != 1 || pl.flops < kSplitMinFlops || pl.ntiles[0] == 0) return;
    const int64_t slots = (int64_t)cus * blocks_per_cu(pl.cfg);
    int64_t work = 0, deepest = 0;  // k steps, summed over the tiles and the longest
    for (int64_t i = 0; i < pl.ntiles[0]; ++i) {
        const int64_t len = (int64_t)pl.tiles[0][i].k_hi - pl.tiles[0][i].k_lo;
        work += len;
        deepest = len > deepest ? len : deepest;
    }
    const int64_t share = work / slots;
    if (2 * deepest < 3 * share) return;
    int64_t kc = (share / 2 + 127) / 128 * 128;
    kc = kc < 512 ? 512 : kc;
    std::vector<LaunchTile> cut;
    std::vector<PartSum> cut_sums;
    int64_t cut_parts = 0;
    for (int64_t i = 0; i < pl.ntiles[0]; ++i) {
        const uvt_tile& t = pl.tiles[0][i];
        const int64_t len = (int64_t)t.k_hi - t.k_lo;
        const int64_t np = len > kc ? (len + kc - 1) / kc : 1;
        if (np == 1) {
            cut.push_back(LaunchTile{t, -1});
            continue;
        }
        // piece boundaries on whole 16-deep panels, as even as that allows
        cut_sums.push_back(PartSum{t.r0, t.rows, t.c0, (int32_t)cut_parts, (int32_t)(np - 1)});
        for (int64_t q = 0; q < np; ++q) {
            uvt_tile piece = t;
            piece.k_lo = q == 0 ? t.k_lo : (int32_t)(t.k_lo + (len * q / np) / 16 * 16);
            piece.k_hi = q == np - 1 ? t.k_hi : (int32_t)(t.k_lo + (len * (q + 1) / np) / 16 * 16);
            cut.push_back(LaunchTile{piece, q == 0 ? -1 : (int32_t)cut_parts++});
        }
    }
    const size_t tile_bytes = (size_t)uvt_cfg_tile_m(pl.cfg) * uvt_cfg_tile_n(pl.cfg) * sizeof(double);
    if ((size_t)cut_parts * tile_bytes > kMaxPartBytes) return;
    // deepest first (the planner's order among equals)
    std::stable_sort(cut.begin(), cut.end(), [](const LaunchTile& a, const LaunchTile& b) {
        return a.t.k_hi - a.t.k_lo > b.t.k_hi - b.t.k_lo;
    });
    launch[0].swap(cut);
    sums.swap(cut_sums);
    nparts = cut_parts;
}

}  // namespace

static uvt_status plan_upload(uvt_handle h, uvt_plan plan)
{
    if (plan->dev_id == h->device) return UVT_OK;
    uvt_plan_release_device(plan);
    // A trapezoid plan whose whole tiles would not finish together runs
    // stream-K: nothing to upload but its partial tiles.
    const int64_t sk = plan->phases == 1 ? uvt_sk_workgroups(plan->cfg, plan->n, plan->m, h->cus) : 0;
    if (sk > 0) {
        if (hipMalloc(&plan->dev_parts, uvt_sk_parts_bytes(plan->cfg, sk)) != hipSuccess) {
            (void)hipGetLastError();
            plan->dev_parts = nullptr;
            return UVT_ERR_NOMEM;
        }
        plan->dev_id = h->device;
        plan->sk_grid = sk;
        const SkGeo g =
            sk_geometry(plan->n, plan->m, uvt_cfg_tile_m(plan->cfg), uvt_cfg_tile_n(plan->cfg), sk, kSkGroupCols);
        for (int64_t b = 1; b < g.G; ++b)
            if (sk_piece(g, sk_start(g, b), sk_start(g, b + 1)).p0 > 0) plan->nparts += g.ncg;
        return UVT_OK;
    }
    std::vector<LaunchTile> launch[2];
    std::vector<PartSum> sums;
    int64_t nparts = 0;
    build_launch(*plan, h->cus, launch, sums, nparts);
    auto upload = [&](void** dst, const void* src, size_t bytes) -> uvt_status {
        if (!bytes) return UVT_OK;
        if (hipMalloc(dst, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return UVT_ERR_NOMEM;
        }
        if (src && hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipGetLastError();
            return UVT_ERR_HIP;
        }
        return UVT_OK;
    };
    plan->dev_id = h->device;  // what is allocated below belongs to this device
    uvt_status st = UVT_OK;
    for (int ph = 0; ph < 2 && st == UVT_OK; ++ph) {
        if (nparts > 0) {
            st = upload(&plan->dev_launch[ph], launch[ph].data(), launch[ph].size() * sizeof(LaunchTile));
        } else {  // whole tiles: the plain records of the kernel without partial tiles
            std::vector<uvt_tile> plain;
            for (const LaunchTile& lt : launch[ph]) plain.push_back(lt.t);
            st = upload(&plan->dev_launch[ph], plain.data(), plain.size() * sizeof(uvt_tile));
        }
        plan->nlaunch[ph] = (int64_t)launch[ph].size();
    }
    if (st == UVT_OK) st = upload(&plan->dev_sums, sums.data(), sums.size() * sizeof(PartSum));
    plan->nsums = (int64_t)sums.size();
    const int32_t tm = uvt_cfg_tile_m(plan->cfg), tn = uvt_cfg_tile_n(plan->cfg);
    if (st == UVT_OK) st = upload(&plan->dev_parts, nullptr, (size_t)nparts * tm * tn * sizeof(double));
    plan->nparts = nparts;
    if (st != UVT_OK) uvt_plan_release_device(plan);
    return st;
}

// ---------------------------------------------------------------------
// FP64 TRMM
// ---------------------------------------------------------------------

// parts == nullptr: dtiles holds uvt_tiles; else LaunchTiles.
template <int TM, int TN>
static void launch_tiles(const void* dtiles, int64_t ntiles, bool vec, const double* A,
                         int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                         int64_t m, double* parts, hipStream_t s)
{
    dim3 grid((unsigned)ntiles), block(64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N);
    if (parts) {
        if (vec)
            hipLaunchKernelGGL((k_trmm_tiles_f64<TM, TN, true, true>), grid, block, 0, s, dtiles, A, lda,
                               B, ldb, C, ldc, m, parts);
        else
            hipLaunchKernelGGL((k_trmm_tiles_f64<TM, TN, false, true>), grid, block, 0, s, dtiles, A, lda,
                               B, ldb, C, ldc, m, parts);
    } else {
        if (vec)
            hipLaunchKernelGGL((k_trmm_tiles_f64<TM, TN, true, false>), grid, block, 0, s, dtiles, A, lda,
                               B, ldb, C, ldc, m, parts);
        else
            hipLaunchKernelGGL((k_trmm_tiles_f64<TM, TN, false, false>), grid, block, 0, s, dtiles, A, lda,
                               B, ldb, C, ldc, m, parts);
    }
}

static uvt_status launch_tiles_f64(int32_t cfg, const void* dtiles, int64_t ntiles,
                                   const double* A, int64_t lda, const double* B, int64_t ldb,
                                   double* C, int64_t ldc, int64_t m, double* parts, hipStream_t s)
{
    if (ntiles <= 0) return UVT_OK;
    if (ntiles > 0x7FFFFFFF) return UVT_ERR_SIZE;
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 2 == 0 &&
                     ldb % 2 == 0;
    if (cfg == UVT_CFG_256x128)
        launch_tiles<256, 128>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, parts, s);
    else if (cfg == UVT_CFG_64x64)
        launch_tiles<64, 64>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, parts, s);
    else if (cfg == UVT_CFG_128x32)
        launch_tiles<128, 32>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, parts, s);
    else
        launch_tiles<128, 128>(dtiles, ntiles, vec, A, lda, B, ldb, C, ldc, m, parts, s);
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

template <int TM, int TN>
static void launch_trapezoid_sk(bool vec, const SkGeo& g, const double* A, int64_t lda, const double* B,
                                int64_t ldb, double* C, int64_t ldc, double* parts, hipStream_t s)
{
    const dim3 grid((unsigned)sk_launch(g)), block(64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N);
    if (vec)
        hipLaunchKernelGGL((k_trmm_sk_f64<TM, TN, true>), grid, block, 0, s, g, A, lda, B, ldb, C, ldc,
                           parts);
    else
        hipLaunchKernelGGL((k_trmm_sk_f64<TM, TN, false>), grid, block, 0, s, g, A, lda, B, ldb, C, ldc,
                           parts);
    hipLaunchKernelGGL((k_sk_fixup<TM, TN>), dim3((unsigned)sk_launch(g), 4), dim3(256), 0, s, g, parts, C, ldc);
}

namespace {

// When the trapezoid walk runs stream-K (measured on MI355X, 26 shapes from
// 1500 x 1500 to 16384 x 4096 against whole tiles): with 128x128 tiles whose
// whole-tile span is 1.4-2.0x an even share (narrow B: 512-700 deep tiles on
// 512 slots), stream-K took 0.73-0.97x the time (3000 x 3500 at 1.42x: 1.0x);
// where whole tiles balance (span 1.0-1.1x) it took 1.05-1.23x, because its
// workgroups no longer go through k together and share less in L2; with 64x64
// tiles it took 1.02-1.25x at any span (there the partial-tile cut of the
// plan path is faster).  More workgroups than slots (2x, 4x) did not help.
constexpr double kSkMinFlops = 2e9;  // below: launch time, not balance, bounds a walk
constexpr double kSkGain = 1.3;      // whole-tile span over an even share that stream-K needs
                                     // (with column groups: 1-14 % slower up to 1.23, 2 % faster at 1.29)

// How long whole tiles take, in panels per workgroup slot: the launch hands
// tiles, deepest first, to the slot that frees first (list scheduling).
int64_t whole_tile_panels(const SkGeo& g, int64_t slots)
{
    std::vector<int64_t> heap(slots, 0);  // min-heap of the slots' finish times
    auto later = [](int64_t a, int64_t b) { return a > b; };
    int64_t span = 0;
    for (int j = g.nrt - 1; j >= 0; --j) {
        const int64_t d = sk_depth(g, j);
        for (int c = 0; c < g.nct; ++c) {
            std::pop_heap(heap.begin(), heap.end(), later);
            heap.back() += d;
            span = std::max(span, heap.back());
            std::push_heap(heap.begin(), heap.end(), later);
        }
    }
    return span;
}

}  // namespace

int64_t uvt_sk_workgroups(int32_t cfg, int64_t n, int64_t m, int cus)
{
    if (cfg != UVT_CFG_128x128 || (double)m * (double)n * (double)n < kSkMinFlops || n > 0x3FFFFFFF ||
        m > 0x3FFFFFFF)
        return 0;
    const int tm = uvt_cfg_tile_m(cfg), tn = uvt_cfg_tile_n(cfg);
    const int64_t slots = (int64_t)cus * blocks_per_cu(cfg);
    const SkGeo g = sk_geometry(n, m, tm, tn, slots);
    const int64_t tiles = (int64_t)g.nrt * g.nct;
    if (tiles > 16 * slots || g.G < slots) return 0;  // many tiles balance by themselves
    const int64_t even = (g.P + slots - 1) / slots;
    // List scheduling never takes longer than an even share plus the deepest
    // tile (Graham 1969), so with a deepest tile below (kSkGain - 1) x the
    // share the span cannot reach kSkGain x: no need to simulate it.
    if ((double)sk_depth(g, g.nrt - 1) < (kSkGain - 1.0) * (double)even) return 0;
    return (double)whole_tile_panels(g, slots) >= kSkGain * (double)even ? slots : 0;
}

size_t uvt_sk_parts_bytes(int32_t cfg, int64_t workgroups)
{
    return (size_t)workgroups * uvt_cfg_tile_m(cfg) * uvt_cfg_tile_n(cfg) * sizeof(double);
}

size_t uvt_sk_max_parts_bytes(int cus)
{
    return uvt_sk_parts_bytes(UVT_CFG_128x128, (int64_t)cus * blocks_per_cu(UVT_CFG_128x128));
}

uvt_status uvt_launch_trapezoid_sk_f64(int32_t cfg, int64_t n, int64_t m, const double* A, int64_t lda,
                                       const double* B, int64_t ldb, double* C, int64_t ldc,
                                       int64_t workgroups, double* parts, hipStream_t s)
{
    if (n <= 0 || m <= 0 || lda < n || ldb < m || ldc < m || workgroups <= 0 || !parts) return UVT_ERR_ARG;
    if (cfg == UVT_CFG_256x128) return UVT_ERR_ARG;  // stream-K runs 128x128 tiles (uvt_sk_workgroups)
    if (n > 0x3FFFFFFF || m > 0x3FFFFFFF || workgroups > 0x7FFFFFFF) return UVT_ERR_SIZE;
    const SkGeo g = sk_geometry(n, m, uvt_cfg_tile_m(cfg), uvt_cfg_tile_n(cfg), workgroups, kSkGroupCols);
    const bool vec = ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0) && lda % 2 == 0 &&
                     ldb % 2 == 0;
    if (cfg == UVT_CFG_64x64) launch_trapezoid_sk<64, 64>(vec, g, A, lda, B, ldb, C, ldc, parts, s);
    else if (cfg == UVT_CFG_128x32) launch_trapezoid_sk<128, 32>(vec, g, A, lda, B, ldb, C, ldc, parts, s);
    else launch_trapezoid_sk<128, 128>(vec, g, A, lda, B, ldb, C, ldc, parts, s);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
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
    if (cfg == UVT_CFG_256x128) launch_trapezoid<256, 128>(vec, n, m, A, lda, B, ldb, C, ldc, s);
    else if (cfg == UVT_CFG_64x64) launch_trapezoid<64, 64>(vec, n, m, A, lda, B, ldb, C, ldc, s);
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
    double* parts = (double*)plan->dev_parts;
    if (plan->sk_grid > 0)
        return uvt_launch_trapezoid_sk_f64(plan->cfg, plan->n, plan->m, dA, lda, dB, ldb, dC, ldc,
                                           plan->sk_grid, parts, s);
    if (plan->nsums > 0x7FFFFFFF) return UVT_ERR_SIZE;
    for (int ph = 0; ph < 2; ++ph) {
        st = launch_tiles_f64(plan->cfg, plan->dev_launch[ph], plan->nlaunch[ph], dA, lda, dB, ldb,
                              dC, ldc, plan->m, plan->nparts > 0 ? parts : nullptr, s);
        if (st != UVT_OK) return st;
        if (ph == 0 && plan->nsums > 0) {  // the partial tiles belong to phase 0
            const int32_t tm = uvt_cfg_tile_m(plan->cfg), tn = uvt_cfg_tile_n(plan->cfg);
            hipLaunchKernelGGL(k_add_parts, dim3((unsigned)plan->nsums), dim3(256), 0, s,
                               (const PartSum*)plan->dev_sums, (int)tn, (int)(tm * tn),
                               (const double*)parts, dC, ldc, plan->m);
            UVT_HIP(hipGetLastError());
        }
    }
    return UVT_OK;
}

extern "C" uvt_status uvt_plan_launch_info(uvt_handle h, uvt_plan plan, int64_t* workgroups,
                                           int64_t* partial_tiles)
{
    if (!h || !plan) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    const uvt_status st = plan_upload(h, plan);
    if (st != UVT_OK) return st;
    if (workgroups)
        *workgroups = plan->sk_grid > 0 ? sk_launch(sk_geometry(plan->n, plan->m, uvt_cfg_tile_m(plan->cfg),
                                                                uvt_cfg_tile_n(plan->cfg), plan->sk_grid,
                                                                kSkGroupCols))
                                        : plan->nlaunch[0] + plan->nlaunch[1];
    if (partial_tiles) *partial_tiles = plan->nparts;
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

// Bytes of each pinned bounce buffer of host copies (two per handle), and
// the size from which a contiguous host copy goes through them: staged
// through pinned memory a copy runs at about 37 GB/s on MI355X, a pageable
// one at about 27 GB/s.
constexpr size_t kPinBytes = (size_t)16 << 20;
constexpr size_t kPinMinBytes = (size_t)1 << 20;

static uvt_status pin_setup(uvt_handle h)
{
    if (h->pin[0]) return UVT_OK;
    for (int i = 0; i < 2; ++i) {
        if (hipHostMalloc(&h->pin[i], kPinBytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&h->pin_done[i], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (int j = 0; j < 2; ++j) {
                if (h->pin[j]) (void)hipHostFree(h->pin[j]);
                if (h->pin_done[j]) (void)hipEventDestroy(h->pin_done[j]);
                h->pin[j] = nullptr;
                h->pin_done[j] = nullptr;
            }
            return UVT_ERR_NOMEM;
        }
        h->pin_used[i] = false;
    }
    return UVT_OK;
}

// Contiguous host <-> device copy through the handle's pinned buffers, a
// chunk in flight while the host fills (or drains) the other.
static uvt_status copy_pinned(uvt_handle h, char* dst, const char* src, size_t bytes, bool to_device)
{
    const size_t chunks = (bytes + kPinBytes - 1) / kPinBytes;
    auto len = [&](size_t c) { return std::min(kPinBytes, bytes - c * kPinBytes); };
    if (to_device) {
        for (size_t c = 0; c < chunks; ++c) {
            const int b = (int)(c & 1);
            if (h->pin_used[b]) UVT_HIP(hipEventSynchronize(h->pin_done[b]));
            memcpy(h->pin[b], src + c * kPinBytes, len(c));
            UVT_HIP(hipMemcpyAsync(dst + c * kPinBytes, h->pin[b], len(c), hipMemcpyHostToDevice, 0));
            UVT_HIP(hipEventRecord(h->pin_done[b], 0));
            h->pin_used[b] = true;
        }
        UVT_HIP(hipStreamSynchronize(0));
        return UVT_OK;
    }
    auto issue = [&](size_t c) -> hipError_t {
        const int b = (int)(c & 1);
        hipError_t e = hipMemcpyAsync(h->pin[b], src + c * kPinBytes, len(c), hipMemcpyDeviceToHost, 0);
        if (e == hipSuccess) e = hipEventRecord(h->pin_done[b], 0);
        h->pin_used[b] = true;
        return e;
    };
    UVT_HIP(issue(0));
    for (size_t c = 0; c < chunks; ++c) {
        if (c + 1 < chunks) UVT_HIP(issue(c + 1));
        const int b = (int)(c & 1);
        UVT_HIP(hipEventSynchronize(h->pin_done[b]));
        memcpy(dst + c * kPinBytes, h->pin[b], len(c));
    }
    return UVT_OK;
}

// Strided host <-> device copy through the handle's pinned buffers: rows are
// packed (or unpacked) on the host a chunk at a time, and each chunk moves as
// one rectangular copy on the null stream while the host packs the next.  A
// pageable strided copy instead pays a fixed cost per row (about 9 us on
// MI355X: 56 ms for the three 2048 x 2048 operand copies of one call).
static uvt_status copy_2d_pinned(uvt_handle h, char* dst, size_t dpitch, const char* src, size_t spitch,
                                 size_t width, size_t height, bool to_device)
{
    const size_t per = kPinBytes / width;
    if (to_device) {
        for (size_t r = 0, c = 0; r < height; r += per, ++c) {
            const size_t rows = std::min(per, height - r);
            const int b = (int)(c & 1);
            char* buf = (char*)h->pin[b];
            if (h->pin_used[b]) UVT_HIP(hipEventSynchronize(h->pin_done[b]));
            for (size_t i = 0; i < rows; ++i) memcpy(buf + i * width, src + (r + i) * spitch, width);
            if (dpitch == width)
                UVT_HIP(hipMemcpyAsync(dst + r * dpitch, buf, rows * width, hipMemcpyHostToDevice, 0));
            else
                UVT_HIP(hipMemcpy2DAsync(dst + r * dpitch, dpitch, buf, width, width, rows,
                                         hipMemcpyHostToDevice, 0));
            UVT_HIP(hipEventRecord(h->pin_done[b], 0));
            h->pin_used[b] = true;
        }
        UVT_HIP(hipStreamSynchronize(0));
        return UVT_OK;
    }
    // device -> host: chunk c + 1 is in flight while chunk c is unpacked
    auto issue = [&](size_t c) -> hipError_t {
        const size_t r = c * per, rows = std::min(per, height - r);
        const int b = (int)(c & 1);
        hipError_t e = spitch == width
                           ? hipMemcpyAsync(h->pin[b], src + r * spitch, rows * width, hipMemcpyDeviceToHost, 0)
                           : hipMemcpy2DAsync(h->pin[b], width, src + r * spitch, spitch, width, rows,
                                              hipMemcpyDeviceToHost, 0);
        if (e == hipSuccess) e = hipEventRecord(h->pin_done[b], 0);
        h->pin_used[b] = true;
        return e;
    };
    const size_t chunks = (height + per - 1) / per;
    UVT_HIP(issue(0));
    for (size_t c = 0; c < chunks; ++c) {
        if (c + 1 < chunks) UVT_HIP(issue(c + 1));
        const int b = (int)(c & 1);
        UVT_HIP(hipEventSynchronize(h->pin_done[b]));
        const size_t r = c * per, rows = std::min(per, height - r);
        const char* buf = (const char*)h->pin[b];
        for (size_t i = 0; i < rows; ++i) memcpy(dst + (r + i) * dpitch, buf + i * width, width);
    }
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
    if (kind != 2 && bytes >= kPinMinBytes) {
        const uvt_status st = pin_setup(h);
        if (st != UVT_OK) return st;
        return copy_pinned(h, (char*)dst, (const char*)src, bytes, kind == 0);
    }
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
    if (dpitch == width && spitch == width)  // one contiguous block
        return uvt_memcpy(h, dst, src, width * height, kind);
    if (kind != 2 && width <= kPinBytes && height > 1) {
        const uvt_status st = pin_setup(h);
        if (st != UVT_OK) return st;
        return copy_2d_pinned(h, (char*)dst, dpitch, (const char*)src, spitch, width, height, kind == 0);
    }
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
