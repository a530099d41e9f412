// uvt_split.hip -- split-precision TRMM: C ~= tril(A) * B with the bulk of the
// products done exactly in int8 on the MFMA units (no counterpart in the
// reference; it replaces the FP64 cblas_dgemm box of the stage loop,
// UVaTRMM/src/mm_triangular.c:347-360, by an integer-modular emulation).
//
// Method (Ozaki scheme II: Ozaki, Uchino, Imamura, "Ozaki Scheme II: A
// GEMM-oriented emulation of floating-point matrix multiplication using an
// integer modular technique", 2024), restated for a lower-triangular A:
//   * Rows of C are cut in 256-row tiles.  For a row tile starting at r0 the
//     product over k in [r0, row] (the diagonal block, lower triangle) runs in
//     FP64 on the FP64 tile kernel; the box k in [0, r0) runs in integers.
//   * Each row i of the box part of A is scaled by 2^alpha_i and each column j
//     of B by 2^beta_j and rounded to integers A', B', with alpha, beta chosen
//     from the 2-norms so that |A'B'| <= ||A'_i|| ||B'_j|| < M/2, where
//     M = m_1 * ... * m_N for N pairwise coprime moduli m_l <= 256.
//   * For each modulus, the symmetric residues of A' and B' fit in int8; the
//     product of residues with int32 accumulation is exact (|sum| < 2^29 for
//     K < 32768 x 4), and is reduced to a residue of A'B' mod m_l.
//   * Garner's algorithm turns the N residues into the mixed-radix digits of
//     the unique C' in (-M/2, M/2); C' is assembled exactly in 128-bit integer
//     arithmetic, converted to FP64, scaled by 2^-(alpha_i+beta_j) and added to
//     the FP64 diagonal-block part.
// The error comes only from rounding A and B to integers: about
// sqrt(K)/||A'_i|| relative, i.e. ~1e-13 for 13 moduli at K = 32768.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "uvt_f64_tile.hpp"
#include "uvt_runtime.hpp"

using namespace uvt;

namespace {

constexpr int kMaxMods = 16;
constexpr int kMaxSlices = 14;  // the two 64-bit halves of the reconstruction stay below 2^56
constexpr int kTileI = 256;  // int8 tile rows = row-tile granularity of the box/diagonal split
constexpr int kI8BK = 128;   // k bytes per LDS stage
constexpr int kI8S = kI8BK + 16;  // LDS row stride (bytes): conflict-free 16-byte fragment reads
constexpr int kI8Stage = 2 * kTileI * kI8S;

// Pairwise coprime moduli, largest first (distinct prime factors).
constexpr int kModuli[kMaxMods] = {256, 255, 253, 251, 247, 241, 239, 233,
                                   229, 227, 223, 217, 211, 199, 197, 193};

struct Mods {
    int n;
    int m[kMaxMods];
    float invf[kMaxMods];
    double invd[kMaxMods];
    uint8_t prodmod[kMaxMods][kMaxMods];  // (m_0 ... m_{p-1}) mod m_l, p < l
    uint8_t qinv[kMaxMods];               // (m_0 ... m_{l-1})^-1 mod m_l
    uint32_t lift[kMaxMods];              // multiple of m_l >= 2^20 (keeps sums >= 0)
    uint32_t magic[kMaxMods];             // ceil(2^39 / m_l): x mod m_l for x < 2^22
    int split;                            // digits [0, split) form the low word
    uint64_t wlo;                         // m_0 ... m_{split-1}   (< 2^56)
    uint64_t qhi;                         // m_split ... m_{n-1}  (< 2^56)
    double wlo_hi, wlo_lo;                // wlo as a double-word
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int64_t row_box_end(int64_t i) { return (i / kTileI) * kTileI; }

// ---- scales ---------------------------------------------------------------

// alpha_i from the 2-norm of A[i][0:kb(i)); one wave per row.
__global__ __launch_bounds__(256) void k_split_alpha(int64_t n, const double* __restrict__ A,
                                                     int64_t lda, double target, int* __restrict__ alpha)
{
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int64_t kb = row_box_end(row);
    const double* a = A + row * lda;
    double s = 0.0;
    for (int64_t k = lane; k < kb; k += 64) s = fma(a[k], a[k], s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        const double nrm = sqrt(s);
        alpha[row] = nrm > 0.0 ? ilogb(target / nrm) : 0;
    }
}

// Partial column sums of B^2 over 256-row chunks of [0, kmax).
__global__ __launch_bounds__(256) void k_split_beta_part(int64_t kmax, int64_t m,
                                                         const double* __restrict__ B, int64_t ldb,
                                                         double* __restrict__ part)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.y * 256;
    if (j >= m) return;
    const int64_t k1 = k0 + 256 < kmax ? k0 + 256 : kmax;
    double s = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
        const double b = B[k * ldb + j];
        s = fma(b, b, s);
    }
    part[(int64_t)blockIdx.y * m + j] = s;
}

__global__ __launch_bounds__(256) void k_split_beta(int64_t m, int64_t chunks,
                                                    const double* __restrict__ part, double target,
                                                    int* __restrict__ beta)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    double s = 0.0;
    for (int64_t c = 0; c < chunks; ++c) s += part[c * m + j];
    const double nrm = sqrt(s);
    beta[j] = nrm > 0.0 ? ilogb(target / nrm) : 0;
}

// ---- residues ---------------------------------------------------------------

// Symmetric residue of an exact integer-valued double a (|a| < 2^53).
__device__ __forceinline__ int sym_residue(double a, int m, double invm)
{
    const double q = rint(a * invm);
    int r = (int)fma(-q, (double)m, a);
    if (r > (m - 1) / 2) r -= m;
    if (r < -(m / 2)) r += m;
    return r;
}

// Ar[l][i][k] for k < kb(i): 16 consecutive k per thread.
__global__ __launch_bounds__(256) void k_split_resid_a(int64_t n, const double* __restrict__ A,
                                                       int64_t lda, const int* __restrict__ alpha,
                                                       int8_t* __restrict__ Ar, int64_t ldr,
                                                       int64_t mod_stride, Mods mods)
{
    for (int64_t i = blockIdx.y; i < n; i += gridDim.y) {
        const int64_t kb = row_box_end(i);
        const int al = alpha[i];
        for (int64_t k0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 16; k0 < kb;
             k0 += (int64_t)gridDim.x * 256 * 16) {
            double v[16];
            const double* a = A + i * lda + k0;
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = rint(ldexp(a[e], al));
            for (int l = 0; l < mods.n; ++l) {
                uint32_t w[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t x = 0;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        x |= (uint32_t)(uint8_t)(int8_t)sym_residue(v[q * 4 + e], mods.m[l],
                                                                     mods.invd[l])
                             << (8 * e);
                    w[q] = x;
                }
                *reinterpret_cast<v4i*>(Ar + l * mod_stride + i * ldr + k0) =
                    v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
            }
        }
    }
}

// Br[l][j][k] (B transposed) for k < kmax: 64 x 64 tiles through LDS.
__global__ __launch_bounds__(256) void k_split_resid_b(int64_t kmax, int64_t m,
                                                       const double* __restrict__ B, int64_t ldb,
                                                       const int* __restrict__ beta,
                                                       int8_t* __restrict__ Br, int64_t ldr,
                                                       int64_t mod_stride, Mods mods)
{
    __shared__ __attribute__((aligned(16))) int8_t t[64][64 + 16];
    const int64_t k0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
    const int jj = threadIdx.x & 63, kq = threadIdx.x >> 6;
    const int64_t j = j0 + jj;
    const bool jv = j < m;
    const int be = jv ? beta[j] : 0;
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int64_t k = k0 + kq + 4 * q;  // kmax is a multiple of 256: k < kmax always
        v[q] = jv ? rint(ldexp(B[k * ldb + j], be)) : 0.0;
    }
    const int wr = threadIdx.x >> 2, ws = (threadIdx.x & 3) * 16;
    for (int l = 0; l < mods.n; ++l) {
#pragma unroll
        for (int q = 0; q < 16; ++q) t[jj][kq + 4 * q] = (int8_t)sym_residue(v[q], mods.m[l], mods.invd[l]);
        __syncthreads();
        if (j0 + wr < m) {
            const v4i x = *reinterpret_cast<const v4i*>(&t[wr][ws]);
            *reinterpret_cast<v4i*>(Br + l * mod_stride + (j0 + wr) * ldr + k0 + ws) = x;
        }
        __syncthreads();
    }
}

// ---- int8 modular products ---------------------------------------------------

// The k loop of one 256 x 256 tile: A' and B' residue panels (128 bytes of k
// per stage) through a double-buffered LDS tile; 8 waves of 64 x 128, each a
// 2 x 4 block of v_mfma_i32_32x32x32_i8.  FULL: the tile lies inside C, so the
// loads need no bounds checks.
template <bool FULL>
__device__ __forceinline__ void i8_tile_loop(const int8_t* __restrict__ A, int64_t lda_r,
                                             const int8_t* __restrict__ B, int64_t ldb_r, int64_t rows_left,
                                             int64_t cols_left, int K, int8_t* smem, v16i (&acc)[2][4])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    v4i ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + q * 512;
            const int row = c >> 3, seg = (c & 7) * 16;
            if (FULL || row < rows_left) ra[q] = *reinterpret_cast<const v4i*>(A + row * lda_r + k0 + seg);
            else ra[q] = v4i{0};
            if (FULL || row < cols_left) rb[q] = *reinterpret_cast<const v4i*>(B + row * ldb_r + k0 + seg);
            else rb[q] = v4i{0};
        }
    };
    auto store = [&](int buf) {
        int8_t* as = smem + buf * kI8Stage;
        int8_t* bs = as + kTileI * kI8S;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + q * 512;
            const int row = c >> 3, seg = (c & 7) * 16;
            *reinterpret_cast<v4i*>(as + row * kI8S + seg) = ra[q];
            *reinterpret_cast<v4i*>(bs + row * kI8S + seg) = rb[q];
        }
    };
    auto compute = [&](int buf) {
        const int8_t* as = smem + buf * kI8Stage + (wm * 64 + (lane & 31)) * kI8S + (lane >> 5) * 16;
        const int8_t* bs = smem + buf * kI8Stage + kTileI * kI8S + (wn * 128 + (lane & 31)) * kI8S +
                           (lane >> 5) * 16;
#pragma unroll
        for (int ks = 0; ks < kI8BK / 32; ++ks) {
            v4i a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const v4i*>(as + i * 32 * kI8S + ks * 32);
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const v4i*>(bs + j * 32 * kI8S + ks * 32);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    };
    const int nk = K / kI8BK;  // K is a multiple of 256
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const bool more = kt + 1 < nk;
        if (more) load((kt + 1) * kI8BK);
        compute(kt & 1);
        if (more) store((kt + 1) & 1);
        __syncthreads();
    }
}

// One workgroup: a 256 x 256 tile of (A' B') mod m_l over k in [0, r0).
__global__ __launch_bounds__(512, 1) void k_split_gemm_i8(
    const int8_t* __restrict__ Ar, int64_t lda_r, int64_t a_mod_stride,
    const int8_t* __restrict__ Br, int64_t ldb_r, int64_t b_mod_stride,
    uint8_t* __restrict__ R, int64_t ldr, int64_t r_mod_stride, int64_t n, int64_t m, int nrt,
    int nct, Mods mods)
{
    __shared__ __attribute__((aligned(16))) int8_t smem[2 * kI8Stage];
    // XCD-aware mapping.  Workgroups go to the 8 XCDs round-robin, so XCD x
    // runs blocks x, x+8, ...; its s-th block takes tile s%32 of tile group
    // (s/32)*8 + x.  A group is 4 row tiles x 8 column tiles of one modulus,
    // i.e. the 32 workgroups one XCD holds at once, so they share 4 A panels
    // and 8 B panels through that XCD's L2.  Groups run longest-k first.
    const int xcd = (int)(blockIdx.x & 7), slot = (int)(blockIdx.x >> 3);
    const int g = (slot >> 5) * 8 + xcd, t = slot & 31;
    const int ngc = (nct + 7) / 8, per_rg = ngc * mods.n;
    const int rg = g / per_rg, grem = g % per_rg;
    const int l = grem / ngc, ct = (grem % ngc) * 8 + (t & 7);
    const int rt = nrt - 1 - (rg * 4 + (t >> 3));
    if (rt < 1 || ct >= nct) return;  // padding of the last groups (uniform per block)
    const int64_t r0 = (int64_t)rt * kTileI, c0 = (int64_t)ct * kTileI;
    const int8_t* A = Ar + l * a_mod_stride + r0 * lda_r;
    const int8_t* B = Br + l * b_mod_stride + c0 * ldb_r;

    v16i acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = v16i{0};
    if (r0 + kTileI <= n && c0 + kTileI <= m)
        i8_tile_loop<true>(A, lda_r, B, ldb_r, n - r0, m - c0, (int)r0, smem, acc);
    else
        i8_tile_loop<false>(A, lda_r, B, ldb_r, n - r0, m - c0, (int)r0, smem, acc);

    // D layout of v_mfma_i32_32x32x32_i8 (measured on gfx950): lane l, element
    // v holds row 8*(v/4) + 4*(l/32) + v%4, column l%32.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int mod = mods.m[l];
    const float invf = mods.invf[l];
    uint8_t* out = R + l * r_mod_stride;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t col = c0 + wn * 128 + j * 32 + (lane & 31);
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int64_t row = r0 + wm * 64 + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row < n && col < m) {
                    const int x = acc[i][j][v];
                    int r = x - mod * (int)floorf((float)x * invf);
                    if (r < 0) r += mod;
                    if (r >= mod) r -= mod;
                    out[row * ldr + col] = (uint8_t)r;
                }
            }
        }
}

// ---- FP64 diagonal blocks ---------------------------------------------------

// Rows [r0, r0+128) of C = A[rows, t0:row] * B[t0:row] where t0 is the start of
// the 256-row tile holding r0 (lower mask), written (not accumulated).
__global__ __launch_bounds__(512, 2) void k_split_diag_f64(int64_t n, int64_t m, int nct,
                                                           const double* __restrict__ A, int64_t lda,
                                                           const double* __restrict__ B, int64_t ldb,
                                                           double* __restrict__ C, int64_t ldc, int vec)
{
    __shared__ double smem[F64Cfg<128, 128, 4, 2>::SMEM_DOUBLES];
    const int rt = (int)(blockIdx.x / nct), ct = (int)(blockIdx.x % nct);
    const int r0 = rt * 128, c0 = ct * 128;
    const int rows = (int)(n - r0 < 128 ? n - r0 : 128);
    const int k_lo = (r0 / kTileI) * kTileI;
    if (vec)
        f64_tile<128, 128, true, 4, 2>(r0, rows, c0, k_lo, r0 + rows, kTileLower, A, lda, B, ldb, C,
                                       ldc, m, smem);
    else
        f64_tile<128, 128, false, 4, 2>(r0, rows, c0, k_lo, r0 + rows, kTileLower, A, lda, B, ldb,
                                        C, ldc, m, smem);
}

// ---- reconstruction -----------------------------------------------------------

// x mod m for 0 <= x < 2^22: with magic = ceil(2^39/m), floor(x/m) equals
// (x * magic) >> 39 exactly on that range (the rounding up of magic adds
// less than 2^-17 < 1/m to x/m).
__device__ __forceinline__ uint32_t mod_magic(uint32_t x, uint32_t m, uint32_t magic)
{
    return x - (__umulhi(x, magic) >> 7) * m;
}

// C[i][j] += C'_ij * 2^-(alpha_i + beta_j) for rows past the first 256-row
// tile.  Per element: mixed-radix digits v_l of C' from its residues
//   v_l = (r_l - sum_{p<l} v_p (m_0...m_{p-1} mod m_l)) (m_0...m_{l-1})^-1 mod m_l,
// then x = X_lo + W X_hi with X_lo = digits [0, split) and X_hi = the rest,
// both exact in 64 bits, and the signed value (x, or x - M when x > M/2)
// rounded once to FP64 with a double-word W.  Four columns per thread.
template <bool VEC4>
__global__ __launch_bounds__(256) void k_split_reconstruct(int64_t n, int64_t m,
                                                           const uint8_t* __restrict__ R, int64_t ldr,
                                                           int64_t r_mod_stride,
                                                           const int* __restrict__ alpha,
                                                           const int* __restrict__ beta,
                                                           double* __restrict__ C, int64_t ldc,
                                                           Mods mods)
{
    for (int64_t i = kTileI + blockIdx.y; i < n; i += gridDim.y)
    for (int64_t j4 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4; j4 < m;
         j4 += (int64_t)gridDim.x * 256 * 4) {
        const int ne = m - j4 < 4 ? (int)(m - j4) : 4;
        uint32_t res[kMaxMods][4];
#pragma unroll
        for (int l = 0; l < kMaxMods; ++l) {
            if (l < mods.n) {
                const uint8_t* r = R + l * r_mod_stride + i * ldr + j4;
                if (VEC4 && ne == 4) {
                    const uint32_t w = *reinterpret_cast<const uint32_t*>(r);
#pragma unroll
                    for (int e = 0; e < 4; ++e) res[l][e] = (w >> (8 * e)) & 255u;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) res[l][e] = e < ne ? r[e] : 0u;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) res[l][e] = 0u;
            }
        }
        const int ai = alpha[i];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (e >= ne) break;
            uint32_t v[kMaxMods];
#pragma unroll
            for (int l = 0; l < kMaxMods; ++l) {
                uint32_t t = 0;
                if (l < mods.n) {
                    uint32_t acc = mods.lift[l] + res[l][e];
#pragma unroll
                    for (int p = 0; p < l; ++p) acc -= v[p] * (uint32_t)mods.prodmod[p][l];
                    t = mod_magic(acc, (uint32_t)mods.m[l], mods.magic[l]);
                    t = mod_magic(t * (uint32_t)mods.qinv[l], (uint32_t)mods.m[l], mods.magic[l]);
                }
                v[l] = t;
            }
            uint64_t xlo = 0, xhi = 0;
#pragma unroll
            for (int l = kMaxMods - 1; l >= 0; --l) {
                if (l < mods.split) xlo = xlo * (uint64_t)mods.m[l] + v[l];
                else if (l < mods.n) xhi = xhi * (uint64_t)mods.m[l] + v[l];
            }
            double h, low;
            if (2 * xhi > mods.qhi) {  // x > M/2: value x - M = (X_hi - Q + 1) W + (X_lo - W)
                h = (double)((int64_t)xhi - (int64_t)mods.qhi + 1);
                low = (double)((int64_t)xlo - (int64_t)mods.wlo);
            } else {
                h = (double)xhi;
                low = (double)xlo;
            }
            const double d = fma(h, mods.wlo_hi, fma(h, mods.wlo_lo, low));
            const int64_t j = j4 + e;
            C[i * ldc + j] += ldexp(d, -(ai + beta[j]));
        }
    }
}

// ---- host side ------------------------------------------------------------------

struct Layout {
    int64_t kmax, lda_r, ldb_r, chunks;
    size_t off_alpha, off_beta, off_part, off_ar, off_br, off_r, total;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

static Layout layout(int64_t n, int64_t m, int nmod)
{
    Layout L;
    L.kmax = ((n - 1) / kTileI) * kTileI;  // largest box depth (multiple of 256)
    L.lda_r = L.kmax;
    L.ldb_r = L.kmax;
    L.chunks = (L.kmax + 255) / 256;
    size_t o = 0;
    L.off_alpha = o; o = align256(o + (size_t)n * 4);
    L.off_beta = o;  o = align256(o + (size_t)m * 4);
    L.off_part = o;  o = align256(o + (size_t)(L.chunks ? L.chunks : 1) * m * 8);
    L.off_ar = o;    o = align256(o + (size_t)nmod * n * L.lda_r);
    L.off_br = o;    o = align256(o + (size_t)nmod * m * L.ldb_r);
    L.off_r = o;     o = align256(o + (size_t)nmod * n * m);
    L.total = o;
    return L;
}

static int inv_mod(int a, int m)
{
    a %= m;
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return 0;
}

static Mods make_mods(int nmod, double* log2M)
{
    Mods md;
    memset(&md, 0, sizeof md);
    md.n = nmod;
    double lg = 0.0;
    for (int l = 0; l < nmod; ++l) {
        md.m[l] = kModuli[l];
        md.invf[l] = 1.0f / (float)kModuli[l];
        md.invd[l] = 1.0 / (double)kModuli[l];
        md.magic[l] = (uint32_t)(((1ull << 39) + kModuli[l] - 1) / kModuli[l]);
        lg += log2((double)kModuli[l]);
    }
    for (int l = 0; l < nmod; ++l) {
        int w = 1;  // (m_0 ... m_{p-1}) mod m_l
        for (int p = 0; p < l; ++p) {
            md.prodmod[p][l] = (uint8_t)w;
            w = (w * md.m[p]) % md.m[l];
        }
        md.qinv[l] = (uint8_t)inv_mod(w, md.m[l]);
        md.lift[l] = (uint32_t)(((1 << 20) / md.m[l] + 1) * md.m[l]);
    }
    md.split = nmod - 1 < 7 ? nmod - 1 : 7;
    md.wlo = 1;
    for (int l = 0; l < md.split; ++l) md.wlo *= (uint64_t)md.m[l];
    md.qhi = 1;
    for (int l = md.split; l < nmod; ++l) md.qhi *= (uint64_t)md.m[l];
    md.wlo_hi = (double)md.wlo;
    md.wlo_lo = (double)((int64_t)md.wlo - (int64_t)(uint64_t)md.wlo_hi);
    *log2M = lg;
    return md;
}

}  // namespace

extern "C" uvt_status uvt_trmm_split_workspace(int64_t n, int64_t m, int32_t slices, size_t* bytes)
{
    if (!bytes || n <= 0 || m <= 0 || slices < 2 || slices > kMaxSlices) return UVT_ERR_ARG;
    if (n > 131072 || m > 0x3FFFFFFF) return UVT_ERR_SIZE;  // int32 accumulation bound on k
    *bytes = layout(n, m, slices).total;
    return UVT_OK;
}

extern "C" uvt_status uvt_trmm_split(uvt_handle h, int64_t n, int64_t m, int32_t slices,
                                     const double* dA, int64_t lda, const double* dB, int64_t ldb,
                                     double* dC, int64_t ldc, void* dwork, size_t work_bytes,
                                     void* stream)
{
    if (!h || !dA || !dB || !dC || n <= 0 || m <= 0 || lda < n || ldb < m || ldc < m)
        return UVT_ERR_ARG;
    size_t need = 0;
    uvt_status st = uvt_trmm_split_workspace(n, m, slices, &need);
    if (st != UVT_OK) return st;
    if (!dwork || work_bytes < need || ((uintptr_t)dwork & 255)) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    const Layout L = layout(n, m, slices);
    double log2M = 0.0;
    const Mods mods = make_mods(slices, &log2M);
    // |A'_i| |B'_j| must stay below M/2 and every scaled entry below 2^53.
    double lt = (log2M - 2.0) / 2.0;
    if (lt > 52.0) lt = 52.0;
    const double target = ldexp(1.0, (int)floor(lt)) - 0.5 * sqrt((double)(L.kmax + 1));
    char* w = (char*)dwork;
    int* alpha = (int*)(w + L.off_alpha);
    int* beta = (int*)(w + L.off_beta);
    double* part = (double*)(w + L.off_part);
    int8_t* Ar = (int8_t*)(w + L.off_ar);
    int8_t* Br = (int8_t*)(w + L.off_br);
    uint8_t* R = (uint8_t*)(w + L.off_r);
    const bool vec = ((uintptr_t)dA % 16 == 0) && ((uintptr_t)dB % 16 == 0) && lda % 2 == 0 &&
                     ldb % 2 == 0;

    // FP64 diagonal blocks (store every row of C).
    const int nct128 = (int)((m + 127) / 128), nrt128 = (int)((n + 127) / 128);
    hipLaunchKernelGGL(k_split_diag_f64, dim3((unsigned)(nrt128 * nct128)), dim3(512), 0, s, n, m,
                       nct128, dA, lda, dB, ldb, dC, ldc, vec ? 1 : 0);
    UVT_HIP(hipGetLastError());
    if (L.kmax == 0) return UVT_OK;  // n <= 256: no box part

    hipLaunchKernelGGL(k_split_alpha, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, n, dA, lda,
                       target, alpha);
    hipLaunchKernelGGL(k_split_beta_part, dim3((unsigned)((m + 255) / 256), (unsigned)L.chunks),
                       dim3(256), 0, s, L.kmax, m, dB, ldb, part);
    hipLaunchKernelGGL(k_split_beta, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                       L.chunks, part, target, beta);
    {
        int64_t gx = (L.kmax / 16 + 255) / 256;
        if (gx > 64) gx = 64;
        const int64_t gy = n < 16384 ? n : 16384;
        hipLaunchKernelGGL(k_split_resid_a, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, s, n, dA,
                           lda, alpha, Ar, L.lda_r, (int64_t)n * L.lda_r, mods);
    }
    hipLaunchKernelGGL(k_split_resid_b, dim3((unsigned)(L.kmax / 64), (unsigned)((m + 63) / 64)),
                       dim3(256), 0, s, L.kmax, m, dB, ldb, beta, Br, L.ldb_r,
                       (int64_t)m * L.ldb_r, mods);
    UVT_HIP(hipGetLastError());
    const int nrt = (int)((n + kTileI - 1) / kTileI), nct = (int)((m + kTileI - 1) / kTileI);
    const int64_t groups = (int64_t)((nrt - 1 + 3) / 4) * ((nct + 7) / 8) * slices;
    const int64_t blocks = (groups + 7) / 8 * 8 * 32;
    if (blocks > 0x7FFFFFFF) return UVT_ERR_SIZE;
    hipLaunchKernelGGL(k_split_gemm_i8, dim3((unsigned)blocks), dim3(512), 0, s, Ar, L.lda_r,
                       (int64_t)n * L.lda_r, Br, L.ldb_r, (int64_t)m * L.ldb_r, R, m, n * m, n, m, nrt,
                       nct, mods);
    UVT_HIP(hipGetLastError());
    {
        int64_t gx = (m + 1023) / 1024;
        if (gx > 32) gx = 32;
        const int64_t gy = n - kTileI < 32768 ? n - kTileI : 32768;
        if (m % 4 == 0)
            hipLaunchKernelGGL(k_split_reconstruct<true>, dim3((unsigned)gx, (unsigned)gy), dim3(256),
                               0, s, n, m, R, m, n * m, alpha, beta, dC, ldc, mods);
        else
            hipLaunchKernelGGL(k_split_reconstruct<false>, dim3((unsigned)gx, (unsigned)gy), dim3(256),
                               0, s, n, m, R, m, n * m, alpha, beta, dC, ldc, mods);
    }
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}
