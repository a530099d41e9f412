// uvt_f64_tile.hpp -- FP64 MFMA tile of C = tril(A) * B for gfx950.
//
// One workgroup (WAVES_M x WAVES_N wave64 waves) computes a TM x TN tile of C over a k range
// [k_lo, k_hi), optionally masking A above the diagonal (A[i][k] = 0 for
// k > i) and optionally accumulating into C.  This one routine covers every
// product the reference issues per band (UVaTRMM/src/mm_triangular.c:347-393):
//   * the GEMM box  C[b:b+h] = A[b:b+h, 0:b+1] * B[0:b+1]   (no mask, store)
//   * the triangle  C[b+1:b+h] += tril(A_tri) * B_tri       (mask, accumulate)
//   * the trapezoid C[r0:r0+TM] = A[r0:, 0:r0+TM] * B        (mask, store)
//
// Data path: A and B k-panels (BK = 16 deep) stream from HBM with 16-byte
// loads into registers one panel ahead (two with AHEAD2), are written to a double-buffered LDS
// tile, and each wave multiplies its (TM/WAVES_M) x (TN/WAVES_N) sub-tile with
// v_mfma_f64_16x16x4_f64 (one double of A and B per lane, four doubles of
// C per lane).  A keeps its row-major layout in LDS (rows of BK + 2 doubles)
// and B its k-major one, so every LDS write is one 16-byte piece a thread
// loaded; the writes for the next panel are issued during the current
// panel's products (see compute()), and each lane reads its A values for a
// panel as 16-byte pairs of one row.  (A transposed to k-major on the way in took
// four conflicting 8-byte writes per thread and cost about 10 % of the MFMA
// rate, measured register-only.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace uvt {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

// Global-memory pointers.  A generic pointer (one read from a descriptor in
// memory, say) is accessed with flat instructions: 64-bit addresses, and
// their completion is counted together with LDS traffic, so a wait for an
// LDS read also waits for every load in flight.  Typed global, the same
// access compiles to a global load or store (scalar base + lane offset).
#define UVT_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const UVT_GLOBAL T* gbl(const T* p)
{
    return (const UVT_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ UVT_GLOBAL T* gbl(T* p)
{
    return (UVT_GLOBAL T*)p;
}

constexpr int kTileLower = 1;  // UVT_TILE_LOWER
constexpr int kTileAccum = 2;  // UVT_TILE_ACCUM

template <int TM, int TN, int WAVES_M = 2, int WAVES_N = 2, int BK_ = 16>
struct F64Cfg {
    static constexpr int NT = 64 * WAVES_M * WAVES_N;
    static constexpr int BK = BK_;  // k depth of one LDS panel (BK/4 MFMA steps)
    static constexpr int WM = TM / WAVES_M, WN = TN / WAVES_N;
    static constexpr int FM = WM / 16, FN = WN / 16;
    // LDS row strides (doubles).  A: row-major rows of BK + 2 doubles (144 B,
    // 16-byte aligned; the 16 rows a fragment read touches fall on distinct
    // bank pairs).  B: k rows of TN + 16 doubles; a fragment read touches four
    // k rows BK/4 apart (the k order of compute()), which share one half of
    // the 64 LDS banks (a pitch of TN + 4, alternating halves, measured the same).
    static constexpr int SA = BK + 2;
    static constexpr int LDB_S = TN + 16;
    static constexpr int A_ELEMS = TM * SA;
    static constexpr int B_ELEMS = BK * LDB_S;
    static constexpr int SMEM_DOUBLES = 2 * (A_ELEMS + B_ELEMS);
    // 2-double pieces of a panel, and per thread (a panel smaller than the
    // workgroup leaves the last threads idle for it: B_PART / A_PART)
    static constexpr int A_PIECES = TM * BK / 2, B_PIECES = TN * BK / 2;
    static constexpr int A_VEC = (A_PIECES + NT - 1) / NT;
    static constexpr int B_VEC = (B_PIECES + NT - 1) / NT;
    static constexpr bool A_PART = A_PIECES < NT, B_PART = B_PIECES < NT;
    static_assert(A_PIECES % NT == 0 || A_PART, "uneven A panel split");
    static_assert(B_PIECES % NT == 0 || B_PART, "uneven B panel split");
    static_assert(FM >= 1 && FN >= 1, "wave tile below 16x16");
};

// VEC: A, B 16-byte aligned and lda, ldb even, so 2-double loads are legal.
// EDGE_IDLE: waves of an edge tile that hold no row or column of C run no
// products (off where a caller's own registers leave no room for the test:
// the stream-K kernel then spills inside its masked k loop).
// AHEAD2: the global loads of panel kt + 2 are issued right after panel
// kt + 1 goes to LDS (during the products of panel kt) instead of at the top
// of panel kt + 1, so the load registers are never live across a panel's
// start.  It lowers register pressure where one kernel holds several tile
// shapes (the persistent batch kernel: 12 bytes of spills per lane instead
// of 52, 1.5-3 % faster); the single-shape kernels ran 0.2-1.7 % slower
// with it.  The sums and their order are the same either way.
template <int TM, int TN, bool VEC, int WAVES_M = 2, int WAVES_N = 2, int BK_ = 16, bool EDGE_IDLE = true,
          bool AHEAD2 = false>
__device__ __forceinline__ void f64_tile(int r0, int rows, int c0, int k_lo, int k_hi, int flags,
                                         const UVT_GLOBAL double* __restrict__ A, int64_t lda,
                                         const UVT_GLOBAL double* __restrict__ B, int64_t ldb,
                                         UVT_GLOBAL double* __restrict__ C, int64_t ldc, int64_t m,
                                         double* smem, int thread = -1,
                                         UVT_GLOBAL double* __restrict__ part = nullptr)
{
    using G = F64Cfg<TM, TN, WAVES_M, WAVES_N, BK_>;
    constexpr int BK = G::BK, NT = G::NT;
    // thread: the caller's thread index when it must stay opaque to the
    // compiler (a persistent kernel running several tile shapes, whose
    // per-thread offsets would otherwise be hoisted out of its tile loop for
    // every shape at once and spilled); -1: threadIdx.x.
    // part: if set, the tile's product goes there instead of to C, as a
    // TM x TN row-major block (row r, column c of the tile at r * TN + c;
    // only rows < rows and columns inside C are written), to be added to C
    // later in a fixed order.
    const int tid = thread >= 0 ? thread : (int)threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: branches on it stay scalar
    const int wm = wave / WAVES_N, wn = wave % WAVES_N;
    const bool lower = (flags & kTileLower) != 0;
    double* As = smem;                    // [2][TM][SA]
    double* Bs = smem + 2 * G::A_ELEMS;   // [2][BK][LDB_S]

    const int k_start = k_lo & ~1;  // even, so 2-double loads stay aligned
    const int nk = (k_hi > k_start) ? (k_hi - k_start + BK - 1) / BK : 0;
    const bool cols_full = (int64_t)c0 + TN <= m;

    dbl4 acc[G::FM][G::FN];
#pragma unroll
    for (int i = 0; i < G::FM; ++i)
#pragma unroll
        for (int j = 0; j < G::FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    dbl2 ra[G::A_VEC], rb[G::B_VEC];

    // full_tag: std::true_type when the whole k panel lies inside the tile's
    // k range, below the diagonal and inside C (no masks at all); otherwise
    // every element is checked.
    auto load = [&](int k0, auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        const bool a_full = FULL;
        const bool b_full = FULL;
#pragma unroll
        for (int q = 0; q < G::A_VEC; ++q) {
            const int v = tid + q * NT;
            if (G::A_PART && v >= G::A_PIECES) break;
            const int row = v / (BK / 2), kp = (v % (BK / 2)) * 2;
            const int gk = k0 + kp, gr = r0 + row;
            const UVT_GLOBAL double* p = A + (int64_t)gr * lda + gk;
            if (a_full) {
                if (VEC) ra[q] = *reinterpret_cast<const UVT_GLOBAL dbl2*>(p);
                else ra[q] = dbl2{p[0], p[1]};
            } else {
                const bool rv = row < rows;
                const bool v0 = rv && gk >= k_lo && gk < k_hi && (!lower || gk <= gr);
                const bool v1 = rv && gk + 1 >= k_lo && gk + 1 < k_hi && (!lower || gk + 1 <= gr);
                ra[q] = dbl2{v0 ? p[0] : 0.0, v1 ? p[1] : 0.0};
            }
        }
#pragma unroll
        for (int q = 0; q < G::B_VEC; ++q) {
            const int v = tid + q * NT;
            if (G::B_PART && v >= G::B_PIECES) break;
            const int kr = v / (TN / 2), col = (v % (TN / 2)) * 2;
            const int gk = k0 + kr;
            const UVT_GLOBAL double* p = B + (int64_t)gk * ldb + (c0 + col);
            if (b_full) {
                if (VEC) rb[q] = *reinterpret_cast<const UVT_GLOBAL dbl2*>(p);
                else rb[q] = dbl2{p[0], p[1]};
            } else {
                const bool kv = gk >= k_lo && gk < k_hi;
                const bool v0 = kv && (int64_t)c0 + col < m;
                const bool v1 = kv && (int64_t)c0 + col + 1 < m;
                rb[q] = dbl2{v0 ? p[0] : 0.0, v1 ? p[1] : 0.0};
            }
        }
    };

    auto store = [&](int buf) {
        double* as = As + buf * G::A_ELEMS;
        double* bs = Bs + buf * G::B_ELEMS;
#pragma unroll
        for (int q = 0; q < G::A_VEC; ++q) {
            const int v = tid + q * NT;
            if (G::A_PART && v >= G::A_PIECES) break;
            const int row = v / (BK / 2), kp = (v % (BK / 2)) * 2;
            *reinterpret_cast<dbl2*>(&as[row * G::SA + kp]) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < G::B_VEC; ++q) {
            const int v = tid + q * NT;
            if (G::B_PART && v >= G::B_PIECES) break;
            const int kr = v / (TN / 2), col = (v % (TN / 2)) * 2;
            *reinterpret_cast<dbl2*>(&bs[kr * G::LDB_S + col]) = rb[q];
        }
    };

    // The products of the panel in LDS buffer buf; with st, the registers
    // loaded for the next panel go to buffer buf ^ 1.  The source places the
    // writes after the second of the four MFMA steps (measured at N=32768,
    // M=8192: 69.5-70.4 TFLOP/s against 67.4-69.2 with them after the
    // first, third or fourth step); the compiler's scheduler then issues them
    // after the first 4 products of the panel (256x128 tiles) or 8 (128x128),
    // and moves the barrier that follows ahead of the last step's products.
    // Pinning the writes with __builtin_amdgcn_sched_barrier after 0, 2, 4,
    // 8, 16 or 24 products measured 3-12 % slower (4 included: the pins also
    // stop the other interleaving), as did the max-ilp and iterative-ilp
    // scheduler strategies (0.8-1.2 %); same bits in every case.
    //
    // k order inside a panel: at MFMA step s, lane group g = lane / 16 takes
    // k = g * BK/4 + s (A and B alike, so the four steps still cover the
    // panel's BK values of k once each).  A lane's A values for the whole
    // panel are then BK/4 consecutive doubles of one row, read as 16-byte
    // pairs: half the LDS reads of one 8-byte read per step (1.4 % on a
    // register-only probe of this loop, 73.2 -> 74.3 TFLOP/s).
    static_assert(BK % 8 == 0, "panel depth: whole 16-byte A pairs per lane group");
    auto compute = [&](int buf, bool st, auto&& next) {
        const int kg = (lane >> 4) * (BK / 4);
        const double* as = As + buf * G::A_ELEMS + (wm * G::WM + (lane & 15)) * G::SA + kg;
        const double* bs = Bs + buf * G::B_ELEMS + kg * G::LDB_S + wn * G::WN + (lane & 15);
#pragma unroll
        for (int sp = 0; sp < BK / 8; ++sp) {
            dbl2 a2[G::FM];
#pragma unroll
            for (int i = 0; i < G::FM; ++i)
                a2[i] = *reinterpret_cast<const dbl2*>(&as[i * 16 * G::SA + 2 * sp]);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int s = 2 * sp + h;
                double b[G::FN];
#pragma unroll
                for (int j = 0; j < G::FN; ++j) b[j] = bs[s * G::LDB_S + j * 16];
#pragma unroll
                for (int i = 0; i < G::FM; ++i)
#pragma unroll
                    for (int j = 0; j < G::FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[i][h], b[j], acc[i][j], 0, 0, 0);
                if (s == 1 && st) {
                    store(buf ^ 1);
                    next();  // AHEAD2: the loads of the panel after next
                }
            }
        }
    };

    // Panels kt in [kt0, nfull) need no masks: k0 >= k_lo (only panel 0 can
    // start below an odd k_lo), k0 + BK <= k_hi, left of the diagonal for a
    // lower tile, and the tile inside C.  They form the steady-state loop,
    // free of per-panel branches; the rest (next to the diagonal, at the
    // tile's k end) run masked.
    if (nk > 0) {
        const std::true_type full{};
        const std::false_type masked{};
        const int lim = lower ? (k_hi < r0 + 1 ? k_hi : r0 + 1) : k_hi;
        const int kt0 = k_start < k_lo ? 1 : 0;
        const int nfull = (rows == TM && cols_full && lim >= k_start) ? (lim - k_start) / BK : 0;
        if (kt0 == 0 && nfull > 0) load(k_start, full);
        else load(k_start, masked);
        store(0);
        if (AHEAD2 && nk > 1) {
            if (1 < nfull) load(k_start + BK, full);  // panel 1 is full from kt0 <= 1 on
            else load(k_start + BK, masked);
        }
        __syncthreads();
        int kt = 0;
        // Registers hold the next panel's values from here on when AHEAD2.
        if constexpr (AHEAD2) {
            for (; kt + 2 < nfull; ++kt) {  // the panel after next is full
                compute(kt & 1, true, [&] { load(k_start + (kt + 2) * BK, full); });
                __syncthreads();
            }
        } else {
            for (; kt + 1 < nfull; ++kt) {  // the next panel is full
                load(k_start + (kt + 1) * BK, full);
                compute(kt & 1, true, [] {});
                __syncthreads();
            }
        }
        // In a lower tile, panels that start past the last of this wave's
        // rows meet only masked (zero) A values: the wave runs them without
        // products (the second loop: it only stores the next panel), which
        // leaves the CU's matrix units to the other waves.  The panel count
        // is uniform over the wave.  Skipping adds of exact zeros changes no
        // result (the sign of a zero aside).
        // Likewise a wave whose rows all lie past the tile's last row, or
        // whose columns all lie past C's last column (a tile at the bottom or
        // right edge, which never takes the steady-state loop), multiplies
        // only zeros and writes nothing: it runs no products.
        const bool idle = EDGE_IDLE && (wm * G::WM >= rows || (int64_t)c0 + wn * G::WN >= m);
        int nk_w = nk;
        if (lower) {
            const int last = idle ? -1 : r0 + wm * G::WM + G::WM - 1;  // last row of this wave
            nk_w = last >= k_start ? (last - k_start) / BK + 1 : 0;
            nk_w = nk_w < kt ? kt : (nk_w > nk ? nk : nk_w);
        }
        for (; kt < nk_w; ++kt) {
            const bool more = kt + 1 < nk;
            if (!AHEAD2 && more) load(k_start + (kt + 1) * BK, masked);
            compute(kt & 1, more, [&] {
                if (AHEAD2 && kt + 2 < nk) load(k_start + (kt + 2) * BK, masked);
            });
            __syncthreads();
        }
        for (; kt < nk; ++kt) {
            if (kt + 1 < nk) {
                if (!AHEAD2) load(k_start + (kt + 1) * BK, masked);
                store((kt & 1) ^ 1);
                if (AHEAD2 && kt + 2 < nk) load(k_start + (kt + 2) * BK, masked);
            }
            __syncthreads();
        }
    }

    // D layout of v_mfma_f64_16x16x4_f64 (measured on gfx950): lane l,
    // element v holds row l/16 + 4*v, column l%16 of the 16x16 block.
    const bool accum = (flags & kTileAccum) != 0;
#pragma unroll
    for (int i = 0; i < G::FM; ++i) {
#pragma unroll
        for (int j = 0; j < G::FN; ++j) {
            const int col = wn * G::WN + j * 16 + (lane & 15);
            const bool cv = (int64_t)c0 + col < m;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wm * G::WM + i * 16 + (lane >> 4) + 4 * v;
                if (cv && row < rows) {
                    if (part) {
                        part[row * TN + col] = acc[i][j][v];
                    } else {
                        UVT_GLOBAL double* p = C + (int64_t)(r0 + row) * ldc + (c0 + col);
                        if (accum) *p += acc[i][j][v];
                        else *p = acc[i][j][v];
                    }
                }
            }
        }
    }
}

// Workgroup shapes of the FP64 tile kernels (waves along rows x columns and
// workgroups resident per CU).  A 128x128 tile runs as 8 waves of 32x64 (4
// waves per SIMD with two workgroups per CU, which hides global-load latency
// better than 4 waves of 64x64); 64x64 tiles (small problems) as 4 waves of
// 32x32; 128x32 tiles (narrow B) as 4 waves of 32x32 stacked along rows.
// WAVES_PER_EU (the second __launch_bounds__ argument on AMD) caps the
// registers per wave so that BLOCKS_PER_CU workgroups fit on one CU.
template <int TM, int TN> struct TileWaves;
template <> struct TileWaves<128, 128> {
    static constexpr int M = 4, N = 2, BLOCKS_PER_CU = 2, WAVES_PER_EU = 4;
};
// 256x128 tiles: 16 waves of 32x64; as 8 waves of 64x64 (a third fewer LDS
// reads, two waves per SIMD, 234 VGPRs) they measured 11-13 % slower.
template <> struct TileWaves<256, 128> {
    static constexpr int M = 8, N = 2, BLOCKS_PER_CU = 1, WAVES_PER_EU = 4;
};
template <> struct TileWaves<64, 64> {
    static constexpr int M = 2, N = 2, BLOCKS_PER_CU = 3, WAVES_PER_EU = 3;
};
template <> struct TileWaves<128, 32> {
    static constexpr int M = 4, N = 1, BLOCKS_PER_CU = 3, WAVES_PER_EU = 3;
};

}  // namespace uvt
