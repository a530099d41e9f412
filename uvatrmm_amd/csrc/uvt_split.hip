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
//     product of residues with int32 accumulation is exact (|sum| <= 128^2 K
//     < 2^31 for the depths K <= 130816 that n <= 131072 allows), and is
//     reduced to a residue of A'B' mod m_l.
//   * The N residues determine the unique C' in (-M/2, M/2) (Chinese
//     remainder theorem); C' is formed in FP64 from exact 40-bit limb sums,
//     scaled by 2^-(alpha_i+beta_j) and added to the FP64 diagonal-block
//     part.
// The error comes only from rounding A and B to integers (and the FP64
// diagonal blocks): about sqrt(K)/||A'_i|| relative, ~1e-13 for 13 moduli
// at K = 32768.  Each call also states its own error (see "error
// statement" below): a rigorous bound and the expected value, both relative
// to ||C||_F, computed on the device from the scales and norms.
//
// Scales.  The 2-norms come from sums of squares split by magnitude
// (uvt_xd.hpp), so rows or columns of any finite magnitude get exact power-
// of-two scales (no overflow of the squares near 1e154, no underflow to a
// zero norm below 1e-162).  A NaN or infinity in the box part of row i of A
// (or in rows < kmax of column j of B) cannot be represented in integers:
// that row of A' (those entries of B') are zeros, and the reconstruction
// writes NaN into every C[i][j] whose box product reads a non-finite entry
// (the FP64 path would give +-Inf or NaN there).
//
// Every stage is a batched kernel over a list of problems: a workgroup finds
// its problem from a per-stage lookup table into a device descriptor table,
// or takes the descriptor passed by value when the launch is for one problem
// (uvt_trmm_split).  The batched runner (uvt_batch.hip) always goes through
// the device table, so uvt_batch_update can rebind pointers under a captured
// graph.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "uvt_f64_tile.hpp"
#include "uvt_runtime.hpp"
#include "uvt_split.hpp"
#include "uvt_xd.hpp"

using namespace uvt;

namespace {

constexpr int kMaxMods = 16;
constexpr int kMaxSlices = 14;  // limb sums of the reconstruction stay below 14 * 256 * 2^40 < 2^52
constexpr int kTileI = 256;  // int8 tile rows = row-tile granularity of the box/diagonal split
constexpr int kI8BK = 64;    // k bytes per LDS stage (one v_mfma_i32_16x16x64_i8 step)
constexpr int kI8Stage = kTileI * kI8BK;  // one B' stage in LDS: 16 fragments of 1 KiB
constexpr int kFrag = 1024;       // one MFMA operand fragment: 16 rows x 64 k bytes
constexpr int kFragSkew = 256;    // extra bytes per 16-row block of A' (see a_tile_stride)
constexpr int kNonFinite = 0x40000000;    // alpha of a row whose box part holds a NaN / infinity
constexpr int kAllFinite = 0x7FFFFFFF;    // bnf of a column whose box rows are all finite
// B' residues of more than this many bytes are built and multiplied in two
// groups of moduli (one buffer reused), which halves their footprint for one
// extra pass over B.
constexpr size_t kBGroupBytes = (size_t)512 << 20;

// Pairwise coprime moduli, largest first (distinct prime factors).
constexpr int kModuli[kMaxMods] = {256, 255, 253, 251, 247, 241, 239, 233,
                                   229, 227, 223, 217, 211, 199, 197, 193};

// Moduli of the call, and the group [l0, l0 + nl) a residue-B / product
// launch works on.
struct Mods {
    int n, l0, nl;
    int m[kMaxMods];
    double invd[kMaxMods];
    float invf[kMaxMods];
};

// Stages, in launch order.
enum { K_DIAG = 0, K_ALPHA, K_BPART, K_BETA, K_BROWS, K_BCOLS, K_RA, K_RB, K_GEMM, K_REC, K_BFIN, K_NUM };
static_assert(K_NUM == UVT_SPLIT_STAGES, "stage count");

// Per row of A: 1-norm and sum of squares of the box part, sum of squares of
// the diagonal-block part (f NaN: a non-finite entry there).
struct RowQ {
    xd s1a, s2a, s2d;
};
// Per (256-row chunk, column) of B: sum of squares (sf, se) and of |b|
// (af, ae); af NaN: the chunk holds a non-finite entry of that column.
struct Part {
    double sf, af;
    int se, ae;
};
// Error-bound terms summed over one 256-row tile (see k_bound_final).
struct TileRow {
    xd ru2, ru1, rw, ra2, rd2;
    int nonfinite, pad;
};
struct TileCol {
    xd cq1, cq2, cb, cv;
    int nonfinite, pad;
};

// One problem as the kernels see it.
struct Desc {
    const double* A;
    const double* B;
    double* C;
    int64_t lda, ldb, ldc;
    int64_t n, m, kmax;       // kmax: deepest box (multiple of 256, 0 if n <= 256)
    int* alpha;
    int* beta;
    int* bnf;                 // per column: first 256-row chunk of B with a non-finite entry
    RowQ* rowq;
    Part* part;               // [nrt][m]
    xd* pre;                  // [nrt][m]: sum over chunks c < t of |B[chunk c][j]|_1
    TileRow* trow;            // [nrt]
    TileCol* tcol;            // [nrt]
    xd* cpart;                // [count[K_REC]]: sum of squares of C per reconstruction workgroup
    double* err;              // [4]: bound, estimate (relative), ||C||_F, absolute bound
    int8_t* Ar;               // [mod][row tile t][16 blocks of a_tile_stride(t) bytes] (see k_resid_a)
    int8_t* Br;               // [mod in group][k block][column block] fragments (see k_resid_b)
    uint8_t* R;               // [mod][n - 256][m]
    double target;            // row / column 2-norm target before rounding
    double c_max;             // bound on |C'_ij| = |A'_i B'_j|
    int32_t nrt, nct;         // 256-row/column tiles
    int32_t nrt128, nct128;   // 128 tiles of the FP64 diagonal part
    int32_t gxa, gxr;         // resid_a workgroups per 32-row block (128 k each), reconstruct per row
    int32_t vec;              // 16-byte FP64 loads allowed
    int32_t gr;               // row tiles per XCD group of the int8 products (x 32/gr columns)
    int32_t start[K_NUM];     // first workgroup of this problem in each stage
    int32_t count[K_NUM];     // workgroups of this problem in each stage
};

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// 16 bytes at base + off as a global load (scalar base + 32-bit lane offset;
// base comes from a descriptor, see gbl()).
__device__ __forceinline__ v4i ld_global16(const int8_t* base, uint32_t off)
{
    const UVT_GLOBAL int8_t* g = (const UVT_GLOBAL int8_t*)base;  // address-space cast
    return *(const UVT_GLOBAL v4i*)(g + off);
}

// The problem of this workgroup (a copy in scalar registers: every field is
// uniform over the workgroup).
__device__ __forceinline__ Desc fetch(const Desc& one, const Desc* table, const int* lut)
{
    if (!table) return one;
    // word by word through a global pointer (global, not flat, loads)
    static_assert(sizeof(Desc) % 4 == 0, "Desc is copied in 4-byte words");
    Desc d;
    const UVT_GLOBAL int* src = (const UVT_GLOBAL int*)gbl(&table[gbl(lut)[blockIdx.x]]);
    int* dst = reinterpret_cast<int*>(&d);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(Desc) / 4); ++w) dst[w] = src[w];
    return d;
}

__device__ __forceinline__ int64_t row_box_end(int64_t i) { return (i / kTileI) * kTileI; }
// Residue arrays are padded with zero rows to whole 256-row tiles, so every
// product tile loads without bounds checks.
__host__ __device__ __forceinline__ int64_t pad_tile(int64_t x) { return (x + kTileI - 1) / kTileI * kTileI; }
// A' residues are stored per 256-row tile t, only over the box that tile
// multiplies (k < 256 t): 16 blocks of 16 rows, each of 4t fragments plus a
// 256-byte skew.  Without the skew every row block would start on a 4 KiB
// boundary, and the waves of an XCD, which read the same k block of
// different row blocks at once, would crowd into a quarter of the L2
// channels (measured 3-11 % slower products).
__host__ __device__ __forceinline__ int64_t a_tile_stride(int64_t t) { return t * 4 * kFrag + kFragSkew; }
// First byte of row tile t (the sum of the strides of the tiles above).
__host__ __device__ __forceinline__ int64_t a_tile_base(int64_t t)
{
    return 16 * (2 * kFrag * t * (t - 1) + kFragSkew * t);
}
// Bytes of A' residues per modulus.
__host__ __device__ __forceinline__ int64_t a_mod_bytes(int64_t n) { return a_tile_base(pad_tile(n) / kTileI); }

// Sum over the threads of a workgroup of up to 16 waves (lds: one slot per
// wave); every thread gets it.
__device__ __forceinline__ xd block_sum(xd v, xd* lds)
{
    v = xd_wave_sum(v);
    __syncthreads();  // lds may still be read from a previous call
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    xd t = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = xd_add(t, lds[w]);
    return t;
}
// The workgroup total of the squares a NormAcc gathered (NaN if any entry
// was not finite); every thread gets it.
__device__ __forceinline__ xd block_sum_sq(NormAcc& acc, xd* lds)
{
    acc.wave_reduce();
    const xd v = acc.nonfinite ? xd{__builtin_nan(""), 0} : acc.sum_sq();
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    xd t = lds[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) t = xd_add(t, lds[w]);
    return t;
}
// An xd from device memory with one 16-byte global load.
__device__ __forceinline__ xd load_xd(const xd* p)
{
    static_assert(sizeof(xd) == 16, "xd is 16 bytes");
    const v4i r = *(const UVT_GLOBAL v4i*)gbl(p);
    return xd{__builtin_bit_cast(double, (uint64_t)(uint32_t)r.x | ((uint64_t)(uint32_t)r.y << 32)), r.z};
}
__device__ __forceinline__ int block_or(int v, int* lds)
{
    __syncthreads();
    if (threadIdx.x == 0) *lds = 0;
    __syncthreads();
    if (v) *lds = 1;  // benign race: every writer stores 1
    __syncthreads();
    return *lds;
}

__device__ __forceinline__ double nan_value() { return __builtin_nan(""); }

// 2^alpha with 2^alpha * ||row|| <= target < 2^(alpha+1) * ||row|| for a
// sum of squares s2 > 0 (exact: the norm's exponent is carried apart).
__device__ __forceinline__ int scale_exponent(double target, xd s2)
{
    const xd nrm = xd_sqrt(s2);
    return ilogb(target / nrm.f) - nrm.e;
}

// ---- scales and error-bound terms ---------------------------------------------

// alpha_i from the 2-norm of A[i][0:kb(i)), and the row's bound terms; one
// wave per row.
__global__ __launch_bounds__(256) void k_alpha(Desc one, const Desc* table, const int* lut)
{
    const Desc d = fetch(one, table, lut);
    const int64_t wb = (int64_t)blockIdx.x - d.start[K_ALPHA];
    const int64_t row = wb * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= d.n) return;
    const int64_t kb = row_box_end(row);
    const UVT_GLOBAL double* a = gbl(d.A) + row * d.lda;
    NormAcc box, dg;
    // Four loads in flight per lane, then the adds in the same k order as one
    // at a time (the sums, and so the scales, do not change).
    int64_t k = lane;
    for (; k + 192 < kb; k += 256) {
        const double v0 = a[k], v1 = a[k + 64], v2 = a[k + 128], v3 = a[k + 192];
        box.add(v0);
        box.add(v1);
        box.add(v2);
        box.add(v3);
    }
    for (; k < kb; k += 64) box.add(a[k]);
    for (int64_t k = kb + lane; k <= row; k += 64) dg.add(a[k]);
    box.wave_reduce();
    dg.wave_reduce();
    if (lane == 0) {
        const xd s2 = box.sum_sq();
        int al = 0;
        if (box.nonfinite) al = kNonFinite;
        else if (s2.f > 0.0) al = scale_exponent(d.target, s2);
        gbl(d.alpha)[row] = al;
        RowQ q;
        q.s1a = box.sum_abs();
        q.s2a = s2;
        q.s2d = dg.nonfinite ? xd{nan_value(), 0} : dg.sum_sq();
        d.rowq[row] = q;
    }
}

// Per (256-row chunk, column) of B over all of B's rows: sum of squares and
// of absolute values.
__global__ __launch_bounds__(256) void k_beta_part(Desc one, const Desc* table, const int* lut)
{
    const Desc d = fetch(one, table, lut);
    const int64_t wb = (int64_t)blockIdx.x - d.start[K_BPART];
    const int64_t ncb = (d.m + 255) / 256;
    const int64_t chunk = wb / ncb;
    const int64_t j = (wb % ncb) * 256 + threadIdx.x;
    if (j >= d.m) return;
    const int64_t k0 = chunk * 256;
    const int64_t k1 = k0 + 256 < d.n ? k0 + 256 : d.n;
    NormAcc acc;
    const UVT_GLOBAL double* b = gbl(d.B) + j;
    int64_t k = k0;
    for (; k + 3 < k1; k += 4) {  // four loads in flight, added in k order
        const double v0 = b[k * d.ldb], v1 = b[(k + 1) * d.ldb], v2 = b[(k + 2) * d.ldb],
                     v3 = b[(k + 3) * d.ldb];
        acc.add(v0);
        acc.add(v1);
        acc.add(v2);
        acc.add(v3);
    }
    for (; k < k1; ++k) acc.add(b[k * d.ldb]);
    const xd s2 = acc.sum_sq(), s1 = acc.sum_abs();
    d.part[chunk * d.m + j] = Part{s2.f, acc.nonfinite ? nan_value() : s1.f, s2.e, s1.e};
}

// beta_j from the 2-norm of B[0:kmax][j]; the column's first non-finite
// chunk; and the prefix 1-norms pre[t][j] = sum_{c < t} |B[chunk c][j]|_1.
__global__ __launch_bounds__(256) void k_beta(Desc one, const Desc* table, const int* lut)
{
    const Desc d = fetch(one, table, lut);
    const int64_t wb = (int64_t)blockIdx.x - d.start[K_BETA];
    const int64_t j = wb * 256 + threadIdx.x;
    if (j >= d.m) return;
    const Part* part = d.part + j;  // small structs: generic (flat) accesses
    xd s2{0.0, 0};
    xd pre{0.0, 0};
    int nf = kAllFinite;
    for (int t = 1; t < d.nrt; ++t) {  // chunk c = t - 1 < kmax / 256
        const Part p = part[(int64_t)(t - 1) * d.m];
        s2 = xd_add(s2, xd{p.sf, p.se});
        if (p.af != p.af) nf = nf == kAllFinite ? t - 1 : nf;
        else pre = xd_add(pre, xd{p.af, p.ae});
        d.pre[(int64_t)t * d.m + j] = pre;
    }
    gbl(d.beta)[j] = s2.f > 0.0 ? scale_exponent(d.target, s2) : 0;
    gbl(d.bnf)[j] = nf;
}

// Row side of the error bound, one workgroup per 256-row tile t (one thread
// per row): with v_i = |A_box,i|_1 / 2 + kb_i 2^-alpha_i / 4,
//   ru2 = sum v_i^2, ru1 = sum 2^-alpha_i v_i, rw = sum 4^-alpha_i,
//   ra2 = sum |A_box,i|^2 (tiles t >= 1), rd2 = sum |A_diag,i|^2 (all tiles).
__global__ __launch_bounds__(256) void k_bound_rows(Desc one, const Desc* table, const int* lut)
{
    __shared__ xd lds[4];
    __shared__ int flag;
    const Desc d = fetch(one, table, lut);
    const int t = (int)((int64_t)blockIdx.x - d.start[K_BROWS]);
    const int64_t i = (int64_t)t * kTileI + threadIdx.x;
    xd ru2{0.0, 0}, ru1{0.0, 0}, rw{0.0, 0}, ra2{0.0, 0}, rd2{0.0, 0};
    int nf = 0;
    if (i < d.n) {
        const int al = gbl(d.alpha)[i];
        const RowQ q = d.rowq[i];
        if (al == kNonFinite || q.s2d.f != q.s2d.f) {
            nf = 1;
        } else {
            rd2 = q.s2d;
            if (t >= 1) {
                const xd h = xd_pow2(-al);
                const xd v = xd_add(xd{q.s1a.f, q.s1a.e - 1}, xd_mul(xd_from((double)(t * kTileI)), xd_pow2(-al - 2)));
                ru2 = xd_mul(v, v);
                ru1 = xd_mul(h, v);
                rw = xd_mul(h, h);
                ra2 = q.s2a;
            }
        }
    }
    TileRow r;
    r.ru2 = block_sum(ru2, lds);
    r.ru1 = block_sum(ru1, lds);
    r.rw = block_sum(rw, lds);
    r.ra2 = block_sum(ra2, lds);
    r.rd2 = block_sum(rd2, lds);
    r.nonfinite = block_or(nf, &flag);
    r.pad = 0;
    if (threadIdx.x == 0) d.trow[t] = r;
}

// Column side, one workgroup per 256-row tile t (threads over columns):
//   cq1 = sum_j 2^-beta_j pre[t][j], cq2 = sum_j pre[t][j]^2 (t >= 1),
//   cb = |B[chunk t]|_F^2, and for t = 0 cv = sum_j 4^-beta_j.
__global__ __launch_bounds__(256) void k_bound_cols(Desc one, const Desc* table, const int* lut)
{
    __shared__ xd lds[4];
    __shared__ int flag;
    const Desc d = fetch(one, table, lut);
    const int t = (int)((int64_t)blockIdx.x - d.start[K_BCOLS]);
    xd cq1{0.0, 0}, cq2{0.0, 0}, cb{0.0, 0}, cv{0.0, 0};
    int nf = 0;
    for (int64_t j = threadIdx.x; j < d.m; j += 256) {
        const Part p = d.part[(int64_t)t * d.m + j];
        if (p.af != p.af) nf = 1;
        cb = xd_add(cb, xd{p.sf, p.se});
        const int be = gbl(d.beta)[j];
        if (t >= 1) {
            const xd q = d.pre[(int64_t)t * d.m + j];
            cq1 = xd_add(cq1, xd_mul(xd_pow2(-be), q));
            cq2 = xd_add(cq2, xd_mul(q, q));
        } else {
            cv = xd_add(cv, xd_pow2(-2 * be));
        }
    }
    TileCol c;
    c.cq1 = block_sum(cq1, lds);
    c.cq2 = block_sum(cq2, lds);
    c.cb = block_sum(cb, lds);
    c.cv = block_sum(cv, lds);
    c.nonfinite = block_or(nf, &flag);
    c.pad = 0;
    if (threadIdx.x == 0) d.tcol[t] = c;
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

// Rows of A'.  The residues go out in MFMA operand order: for row i of tile
// t = i/256, block b = (i/16)%16 and k block kb = k/64, fragment (b, kb) is
// 1024 contiguous bytes at a_tile_base(t) + b * a_tile_stride(t) + kb * 1024,
// where lane l = i%16 + 16*((k%64)/16) holds k bytes 16*(l/16) .. +16 of row
// i, i.e. the 16 bytes lane l feeds v_mfma_i32_16x16x64_i8.  A wave of the
// product kernel then loads a fragment with one coalesced 1 KiB read.  A
// workgroup covers 32 rows (rows >= 256) x 128 k, i.e. four fragments.  The
// scaled, rounded rows go through LDS: read from A with the 8 lanes of a row
// on 128 contiguous bytes (16 per lane), then taken back as 16 consecutive k
// per thread (row t/8), so a wave writes 8 whole 128-byte lines per modulus.
// Rows n .. pad_tile(n) are zeros, and so are rows with a non-finite entry.
__global__ __launch_bounds__(256) void k_resid_a(Desc one, const Desc* table, const int* lut, Mods mods)
{
    __shared__ __attribute__((aligned(16))) double tv[32][128 + 4];
    const Desc d = fetch(one, table, lut);
    const int64_t wb = (int64_t)blockIdx.x - d.start[K_RA];
    const int64_t rb = kTileI / 32 + wb / d.gxa, kq = wb % d.gxa;
    if (kq * 128 >= row_box_end(rb * 32)) return;  // boxes end on multiples of 256
    const int t = threadIdx.x, c = t & 7, r = t >> 3;
    const int64_t i = rb * 32 + r;
    const int al = i < d.n ? gbl(d.alpha)[i] : kNonFinite;
    const bool iv = al != kNonFinite;  // a real row with finite entries
    {
        const UVT_GLOBAL double* a = gbl(d.A) + i * d.lda + kq * 128 + 2 * c;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            d2 x = {0.0, 0.0};
            if (iv) {
                if (d.vec)  // 16-byte aligned: A and lda even
                    x = *reinterpret_cast<const UVT_GLOBAL d2*>(a + 16 * e);
                else
                    x = d2{a[16 * e], a[16 * e + 1]};
            }
            *reinterpret_cast<d2*>(&tv[r][16 * e + 2 * c]) = d2{rint(ldexp(x.x, al)), rint(ldexp(x.y, al))};
        }
    }
    __syncthreads();
    double v[16];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const d2 x = *reinterpret_cast<const d2*>(&tv[r][16 * c + 2 * e]);
        v[2 * e] = x.x;
        v[2 * e + 1] = x.y;
    }
    const int64_t mod_stride = a_mod_bytes(d.n);
    const int64_t tile = i / kTileI;
    UVT_GLOBAL int8_t* out = gbl(d.Ar) + a_tile_base(tile) + ((i >> 4) & 15) * a_tile_stride(tile) +
                             (kq * 2 + (c >> 2)) * kFrag + ((i & 15) + 16 * (c & 3)) * 16;
    for (int l = 0; l < mods.n; ++l) {
        uint32_t q4[4];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            uint32_t x = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                x |= (uint32_t)(uint8_t)(int8_t)sym_residue(v[w * 4 + e], mods.m[l], mods.invd[l])
                     << (8 * e);
            q4[w] = x;
        }
        *reinterpret_cast<UVT_GLOBAL v4i*>(out + l * mod_stride) =
            v4i{(int)q4[0], (int)q4[1], (int)q4[2], (int)q4[3]};
    }
}

// Residues of B' for k < kmax and the moduli l0 .. l0+nl-1 of the launch, in
// MFMA operand order like A': for column block cb = j/16 and k block
// kb = k/64, fragment (kb, cb) is 1024 contiguous bytes at
// (kb * pad_tile(m)/16 + cb) * 1024, where lane l = j%16 + 16*((k%64)/16)
// holds k bytes 16*(l/16) .. +16 of column j (the 16 bytes lane l feeds
// v_mfma_i32_16x16x64_i8 as its B operand).  A product workgroup then moves
// each fragment from memory to LDS and from LDS to its waves as one
// contiguous 1 KiB per wave.  A workgroup converts 128 k x 64 columns: thread
// (j = t%64, kq = t/64) converts k = 4 (kq + 4p) + e of column j (p < 8,
// e < 4) and packs whole 4-byte words, which go to LDS k-word-major and come
// back so that wave w writes fragments (k block w%2, column blocks w/2 and
// w/2 + 2), one lane's 16 bytes each.  Two LDS buffers, so one barrier per
// modulus.  Columns m .. pad_tile(m) are zeros, and so are non-finite
// entries.
__global__ __launch_bounds__(256) void k_resid_b(Desc one, const Desc* table, const int* lut, Mods mods)
{
    constexpr int kW = 32;          // 4-byte words per column of the tile (128 k)
    constexpr int kPitch = 64 + 4;  // 4 * kPitch = 16 (mod 32) banks: the readback's lanes hit distinct banks
    __shared__ uint32_t t[2][kW][kPitch];  // [buffer][k word][column]
    const Desc d = fetch(one, table, lut);
    const int64_t wb = (int64_t)blockIdx.x - d.start[K_RB];
    const int64_t nkx = d.kmax / 128;  // kmax is a multiple of 256
    const int64_t k0 = (wb % nkx) * 128, j0 = (wb / nkx) * 64;
    const int jj = threadIdx.x & 63, kq = threadIdx.x >> 6;
    const int64_t j = j0 + jj;
    const bool jv = j < d.m;
    const int be = jv ? gbl(d.beta)[j] : 0;
    double v[kW];
#pragma unroll
    for (int p = 0; p < kW / 4; ++p)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t k = k0 + 4 * (kq + 4 * p) + e;  // < kmax
            const double b = jv ? gbl(d.B)[k * d.ldb + j] : 0.0;
            v[p * 4 + e] = fabs(b) <= 1.79769313486231570815e308 ? rint(ldexp(b, be)) : 0.0;
        }
    // readback: wave w, lane = c + 16 q -> column block w/2 (+2), k block w%2
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, c = lane & 15, q = lane >> 4;
    const int64_t ncb = pad_tile(d.m) / 16;
    const int64_t mod_stride = pad_tile(d.m) * d.kmax;
    for (int l = 0; l < mods.nl; ++l) {
        const int lm = mods.l0 + l;
        uint32_t(*tb)[kPitch] = t[l & 1];
#pragma unroll
        for (int p = 0; p < kW / 4; ++p) {
            uint32_t x = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                x |= (uint32_t)(uint8_t)(int8_t)sym_residue(v[p * 4 + e], mods.m[lm], mods.invd[lm])
                     << (8 * e);
            tb[kq + 4 * p][jj] = x;
        }
        // One barrier per modulus: buffer l & 1 is written again only at
        // l + 2, after every thread has passed the barrier of l + 1 and so
        // finished reading it here.
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int cbl = (w >> 1) + 2 * h, kbl = w & 1;
            const int col = cbl * 16 + c, kw = kbl * 16 + 4 * q;
            const v4i x = {(int)tb[kw][col], (int)tb[kw + 1][col], (int)tb[kw + 2][col], (int)tb[kw + 3][col]};
            const int64_t frag = (k0 / 64 + kbl) * ncb + j0 / 16 + cbl;
            *reinterpret_cast<UVT_GLOBAL v4i*>(gbl(d.Br) + l * mod_stride + frag * kFrag + lane * 16) = x;
        }
    }
}

// ---- int8 modular products ---------------------------------------------------

// The k loop of one 256 x 256 tile; 8 waves of 64 x 128, each a 4 x 8 block
// of v_mfma_i32_16x16x64_i8.  A' fragments come straight from memory into
// registers (fragment order, one coalesced 1 KiB load each); B' fragments go
// through a double-buffered LDS stage of 64 k (16 fragments: each wave loads
// two and stores them as they are, and every wave reads its eight; all
// contiguous 1 KiB per wave, so no LDS bank conflicts).  Loads run two k stages ahead:
// the products are bound by memory latency, not bandwidth (one stage of
// prefetch left every CU waiting on about 48 KiB in flight, measured), so
// A' has two register sets, even and odd stages, and each fragment slot is
// refilled for stage kt + 2 as soon as its products are issued; B' has two
// register sets and is written to LDS the stage before its use.  The loop
// is unrolled by two so every buffer has a fixed name.  The residue arrays
// are padded to whole tiles, so no load needs a bounds check.
__device__ __forceinline__ void i8_tile_loop(const int8_t* __restrict__ Af, int64_t astr,
                                             const int8_t* __restrict__ B, int64_t bstr, int K,
                                             int8_t* smem, v4i (&acc)[4][8])
{
    // Addresses are a wave-uniform base (scalar registers) plus one 32-bit
    // per-lane offset, which keeps the vector registers for the operands.
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    // this lane's 16 bytes of fragment (row block 4*wm + i, k block kt)
    const int8_t* Awave = Af + (int64_t)(4 * wm) * astr;
    const uint32_t aoff = (uint32_t)lane * 16;
    // B' fragments 2*wave and 2*wave + 1 of a stage (bstr bytes per k block)
    const int8_t* Bwave = B + (int64_t)(2 * wave) * kFrag;
    const uint32_t boff = (uint32_t)lane * 16;
    const uint32_t soff = (uint32_t)(2 * wave) * kFrag + (uint32_t)lane * 16;
    v4i aE[4], aO[4], bE[2], bO[2];
    auto load_a = [&](int kt, int i) { return ld_global16(Awave + (int64_t)i * astr + (int64_t)kt * kFrag, aoff); };
    auto load_b = [&](int kt, v4i (&d)[2]) {
#pragma unroll
        for (int q = 0; q < 2; ++q) d[q] = ld_global16(Bwave + (int64_t)kt * bstr + q * kFrag, boff);
    };
    auto store_b = [&](int buf, const v4i (&d)[2]) {
        int8_t* bs = smem + buf * kI8Stage + soff;
#pragma unroll
        for (int q = 0; q < 2; ++q) *reinterpret_cast<v4i*>(bs + q * kFrag) = d[q];
    };
    // Products of stage kt from LDS buffer buf; refills a for stage kn.
    auto compute = [&](int buf, v4i (&a)[4], int kn) {
        const int8_t* bs = smem + buf * kI8Stage + wn * 8 * kFrag + lane * 16;
        v4i b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) b[j] = *reinterpret_cast<const v4i*>(bs + j * kFrag);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
            a[i] = load_a(kn, i);
        }
    };
    // Every load and store is unconditional (past the end they fetch the
    // last stage again and the values go unused): a load under a branch
    // makes the compiler's wait counts at the join assume the branch not
    // taken, i.e. wait for the loads just issued.
    // The prologue issues its loads in the loop's order (pinned): the wait
    // counts at the loop head must hold for the first iteration too.
    const int nk = K / kI8BK;  // K is a multiple of 256, so nk is a multiple of 4
    load_b(0, bE);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) aE[i] = load_a(0, i);
    __builtin_amdgcn_sched_barrier(0);
    load_b(1, bO);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) aO[i] = load_a(1, i);
    __builtin_amdgcn_sched_barrier(0);
    store_b(0, bE);
    __syncthreads();
    // The scheduling barriers keep the B' loads at the head of their stage
    // and each LDS store after the products of its stage, so the B' data of
    // a stage has two stages to arrive.
    for (int kt = 0; kt < nk; kt += 2) {
        const int k2 = min(kt + 2, nk - 1), k3 = min(kt + 3, nk - 1);
        load_b(k2, bE);
        __builtin_amdgcn_sched_barrier(0);
        compute(0, aE, k2);
        __builtin_amdgcn_sched_barrier(0);
        store_b(1, bO);
        __syncthreads();
        load_b(k3, bO);
        __builtin_amdgcn_sched_barrier(0);
        compute(1, aO, k3);
        __builtin_amdgcn_sched_barrier(0);
        store_b(0, bE);
        __syncthreads();
    }
}

// One workgroup: a 256 x 256 tile of (A' B') mod m_l over k in [0, r0), for
// the launch's moduli l0 .. l0+nl-1.
__global__ __launch_bounds__(512, 1) void k_gemm_i8(Desc one, const Desc* table, const int* lut, Mods mods)
{
    __shared__ __attribute__((aligned(16))) int8_t smem[2 * kI8Stage];  // B' only
    // Only the fields used here are read (from the kernel argument or the
    // descriptor table), so the descriptor does not occupy registers.
    const Desc* dp = table ? &table[lut[blockIdx.x]] : nullptr;
#define UVT_FIELD(f) (dp ? dp->f : one.f)
    const int64_t wb = (int64_t)blockIdx.x - UVT_FIELD(start[K_GEMM]);
    // XCD-aware mapping.  Workgroups go to the 8 XCDs round-robin (every
    // problem starts at a multiple of 256 workgroups), so XCD x runs local
    // blocks x, x+8, ...; its s-th block takes tile s%32 of tile group
    // (s/32)*8 + x.  A group is gr row tiles x 32/gr column tiles of one
    // modulus (8 x 4 for large problems), i.e. the 32 workgroups one XCD holds
    // at once, so they share their A and B panels through that XCD's L2.
    // Groups run longest-k first.
    const int xcd = (int)(wb & 7), slot = (int)(wb >> 3);
    const int g = (slot >> 5) * 8 + xcd, t = slot & 31;
    const int nct = UVT_FIELD(nct), nrt = UVT_FIELD(nrt);
    const int gr = UVT_FIELD(gr), gc = 32 / gr;
    const int ngc = (nct + gc - 1) / gc, per_rg = ngc * mods.nl;
    const int rg = g / per_rg, grem = g % per_rg;
    // Uniform over the workgroup by construction; readfirstlane tells the
    // compiler, so the depth and the panel bases stay in scalar registers
    // (the divisions above otherwise leave them in vector registers, and
    // every k step's clamp and addresses become per-lane arithmetic).
    const int l = __builtin_amdgcn_readfirstlane(grem / ngc);
    const int ct = __builtin_amdgcn_readfirstlane((grem % ngc) * gc + t % gc);
    const int rt = __builtin_amdgcn_readfirstlane(nrt - 1 - (rg * gr + t / gc));
    if (rt < 1 || ct >= nct) return;  // padding of the last groups (uniform per block)
    const int64_t r0 = (int64_t)rt * kTileI, c0 = (int64_t)ct * kTileI;
    const int64_t n = UVT_FIELD(n), m = UVT_FIELD(m), kmax = UVT_FIELD(kmax);
    const int lm = mods.l0 + l;  // the modulus
    const int8_t* Af = UVT_FIELD(Ar) + lm * a_mod_bytes(n) + a_tile_base(rt);
    const int8_t* B = UVT_FIELD(Br) + l * (pad_tile(m) * kmax) + (c0 / 16) * kFrag;
    UVT_GLOBAL uint8_t* const Rout = gbl(UVT_FIELD(R));
#undef UVT_FIELD

    v4i acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = v4i{0, 0, 0, 0};
    i8_tile_loop(Af, a_tile_stride(rt), B, pad_tile(m) / 16 * kFrag, (int)r0, smem, acc);

    // D layout of v_mfma_i32_16x16x64_i8: lane l, element v holds row
    // 4*(l/16) + v, column l%16.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int mod = mods.m[lm];
    const float invf = mods.invf[lm];
    UVT_GLOBAL uint8_t* out = Rout + lm * ((n - kTileI) * m);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t col = c0 + wn * 128 + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = r0 + wm * 64 + i * 16 + 4 * (lane >> 4) + v;
                if (row < n && col < m) {
                    const int x = acc[i][j][v];
                    int r = x - mod * (int)floorf((float)x * invf);
                    if (r < 0) r += mod;
                    if (r >= mod) r -= mod;
                    out[(row - kTileI) * m + col] = (uint8_t)r;
                }
            }
        }
}

// ---- FP64 diagonal blocks ---------------------------------------------------

// Rows [r0, r0+128) of C = A[rows, t0:row] * B[t0:row] where t0 is the start of
// the 256-row tile holding r0 (lower mask), written (not accumulated).
__global__ __launch_bounds__(512, 4) void k_diag_f64(Desc one, const Desc* table, const int* lut)
{
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<128, 128, 4, 2>::SMEM_DOUBLES];
    const Desc* dp = table ? &table[lut[blockIdx.x]] : nullptr;
#define UVT_FIELD(f) (dp ? dp->f : one.f)
    const int64_t wb = (int64_t)blockIdx.x - UVT_FIELD(start[K_DIAG]);
    const int nct128 = UVT_FIELD(nct128);
    const int64_t n = UVT_FIELD(n);
    const int rt = (int)(wb / nct128), ct = (int)(wb % nct128);
    const int r0 = rt * 128, c0 = ct * 128;
    const int rows = (int)(n - r0 < 128 ? n - r0 : 128);
    const int k_lo = (r0 / kTileI) * kTileI;
    const UVT_GLOBAL double* A = gbl(UVT_FIELD(A));
    const UVT_GLOBAL double* B = gbl(UVT_FIELD(B));
    UVT_GLOBAL double* C = gbl(UVT_FIELD(C));
    if (UVT_FIELD(vec))
        f64_tile<128, 128, true, 4, 2>(r0, rows, c0, k_lo, r0 + rows, kTileLower, A, UVT_FIELD(lda),
                                       B, UVT_FIELD(ldb), C, UVT_FIELD(ldc), UVT_FIELD(m), smem);
    else
        f64_tile<128, 128, false, 4, 2>(r0, rows, c0, k_lo, r0 + rows, kTileLower, A, UVT_FIELD(lda),
                                        B, UVT_FIELD(ldb), C, UVT_FIELD(ldc), UVT_FIELD(m), smem);
#undef UVT_FIELD
}

// ---- reconstruction -----------------------------------------------------------

// Chinese remainder constants for the first N moduli, evaluated at compile
// time (128-bit integers) so that the reconstruction kernel sees them as
// immediates.  With M = m_0 ... m_{N-1}, x = sum_l r_l e_l mod M for
// e_l = (M/m_l) * ((M/m_l)^-1 mod m_l) < M.  Every e_l and M are split in
// three 40-bit limbs, so that sum_l r_l E_{l,k} < 14 * 256 * 2^40 < 2^52 stays
// an exact FP64 integer for each limb k.
template <int N>
struct CrtTab {
    double e[3][N];      // limbs of e_l: e_l = e[2] 2^80 + e[1] 2^40 + e[0]
    double mk[3];        // limbs of M
    double inv_mhi;      // 1 / (M / 2^80)
    constexpr CrtTab() : e(), mk(), inv_mhi(0.0)
    {
        using u128 = unsigned __int128;
        const u128 mask = ((u128)1 << 40) - 1;
        u128 M = 1;
        for (int l = 0; l < N; ++l) M *= (u128)kModuli[l];
        for (int l = 0; l < N; ++l) {
            const uint32_t ml = (uint32_t)kModuli[l];
            const u128 W = M / ml;
            const uint32_t w = (uint32_t)(W % ml);
            uint32_t inv = 1;
            for (uint32_t x = 1; x < ml; ++x)
                if ((w * x) % ml == 1) {
                    inv = x;
                    break;
                }
            const u128 el = W * inv;
            e[0][l] = (double)(uint64_t)(el & mask);
            e[1][l] = (double)(uint64_t)((el >> 40) & mask);
            e[2][l] = (double)(uint64_t)(el >> 80);
        }
        mk[0] = (double)(uint64_t)(M & mask);
        mk[1] = (double)(uint64_t)((M >> 40) & mask);
        mk[2] = (double)(uint64_t)(M >> 80);
        const double two40 = 1099511627776.0;
        inv_mhi = 1.0 / (mk[2] + mk[1] / two40 + mk[0] / (two40 * two40));
    }
};

// C[i][j] += C'_ij * 2^-(alpha_i + beta_j) for rows past the first 256-row
// tile, with N moduli.  Per element, all in FP64 and exact up to the last
// step: the limb sums S_k = sum_l r_l E_{l,k}; q = rint(S / M), which the
// scale target makes unambiguous (|C'| < M/2 by a relative margin above
// 1e-6, and S / M is formed to about 2^-40); X_k = S_k - q M_k; a carry that
// brings X_0 into [0, 2^40), so that X_2 2^40 + X_1 is the integer
// floor(C' / 2^40); and C' = (X_2 2^40 + X_1) 2^40 + X_0 with two roundings
// (fma), exact whenever C' is an FP64 number.  A workgroup takes 1024
// columns of a row, four per thread, and adds the sum of squares of the
// finished C values to the error statement's ||C||_F (workgroups of rows
// 0 .. 255 only read C for it).
template <int N>
__global__ __launch_bounds__(256) void k_reconstruct(Desc one, const Desc* table, const int* lut)
{
    constexpr CrtTab<N> T{};
    constexpr double two40 = 1099511627776.0, inv40 = 1.0 / two40, inv80 = inv40 * inv40;
    __shared__ xd lds[4];
    const Desc d = fetch(one, table, lut);
    const int64_t wb = (int64_t)blockIdx.x - d.start[K_REC];
    const int64_t i = wb / d.gxr;
    const int64_t j4 = (wb % d.gxr) * 1024 + threadIdx.x * 4;
    const int64_t m = d.m;
    NormAcc acc;
    if (i < d.n && j4 < m) {
        const int ne = m - j4 < 4 ? (int)(m - j4) : 4;
        UVT_GLOBAL double* crow = gbl(d.C) + i * d.ldc + j4;
        double cv[4] = {0.0, 0.0, 0.0, 0.0};
        if (i < kTileI) {  // no box part: C is final
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (e < ne) cv[e] = crow[e];
        } else {
            const bool vec4 = (m & 3) == 0 && ne == 4;
            const int64_t mod_stride = (d.n - kTileI) * m;
            uint32_t res[N][4];
#pragma unroll
            for (int l = 0; l < N; ++l) {
                const UVT_GLOBAL uint8_t* r = gbl(d.R) + l * mod_stride + (i - kTileI) * m + j4;
                if (vec4) {
                    const uint32_t x = *reinterpret_cast<const UVT_GLOBAL uint32_t*>(r);
#pragma unroll
                    for (int e = 0; e < 4; ++e) res[l][e] = (x >> (8 * e)) & 255u;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) res[l][e] = e < ne ? r[e] : 0u;
                }
            }
            const int ai = gbl(d.alpha)[i];
            const int tile = (int)(i / kTileI);  // chunks c < tile feed this row's box
            double xs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int l = 0; l < N; ++l) {
                    const double r = (double)res[l][e];
                    s0 = fma(r, T.e[0][l], s0);
                    s1 = fma(r, T.e[1][l], s1);
                    s2 = fma(r, T.e[2][l], s2);
                }
                const double q = rint(fma(s1, inv40, fma(s0, inv80, s2)) * T.inv_mhi);
                double x0 = fma(-q, T.mk[0], s0), x1 = fma(-q, T.mk[1], s1), x2 = fma(-q, T.mk[2], s2);
                const double c0 = floor(x0 * inv40);
                x0 = fma(-c0, two40, x0);
                x1 += c0;
                xs[e] = fma(fma(x2, two40, x1), two40, x0);
            }
            // C, beta and the non-finite markers in 16-byte pieces where
            // aligned (the row start decides, the same for the whole
            // workgroup), else element by element.
            const UVT_GLOBAL int* brow = gbl(d.beta) + j4;
            const UVT_GLOBAL int* frow = gbl(d.bnf) + j4;
            if (ne == 4 && ((uintptr_t)crow & 15) == 0) {
                const v4i b = *reinterpret_cast<const UVT_GLOBAL v4i*>(brow);
                const v4i f = *reinterpret_cast<const UVT_GLOBAL v4i*>(frow);
                d2 c01 = *reinterpret_cast<const UVT_GLOBAL d2*>(crow);
                d2 c23 = *reinterpret_cast<const UVT_GLOBAL d2*>(crow + 2);
                const bool bad = ai == kNonFinite;
                c01.x += (bad || f.x < tile) ? nan_value() : ldexp(xs[0], -(ai + b.x));
                c01.y += (bad || f.y < tile) ? nan_value() : ldexp(xs[1], -(ai + b.y));
                c23.x += (bad || f.z < tile) ? nan_value() : ldexp(xs[2], -(ai + b.z));
                c23.y += (bad || f.w < tile) ? nan_value() : ldexp(xs[3], -(ai + b.w));
                *reinterpret_cast<UVT_GLOBAL d2*>(crow) = c01;
                *reinterpret_cast<UVT_GLOBAL d2*>(crow + 2) = c23;
                cv[0] = c01.x;
                cv[1] = c01.y;
                cv[2] = c23.x;
                cv[3] = c23.y;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (e < ne) {
                        const bool bad = ai == kNonFinite || frow[e] < tile;
                        cv[e] = crow[e] + (bad ? nan_value() : ldexp(xs[e], -(ai + brow[e])));
                        crow[e] = cv[e];
                    }
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) acc.add(cv[e]);  // zeros past the row end add nothing
    }
    const xd s = block_sum_sq(acc, lds);
    if (threadIdx.x == 0) d.cpart[wb] = s;
}

template <int N>
void launch_reconstruct(unsigned grid, hipStream_t s, const Desc& one, const Desc* table, const int* lut)
{
    hipLaunchKernelGGL(k_reconstruct<N>, dim3(grid), dim3(256), 0, s, one, table, lut);
}

// ---- error statement ----------------------------------------------------------

// The error of the result, one workgroup per problem, in two figures
// relative to ||C||_F (both NaN when any input entry is non-finite):
//
// bound (rigorous).  With A'_ik = A_ik 2^alpha_i + d_ik, B'_kj = B_kj
// 2^beta_j + e_kj, |d|, |e| <= 1/2, the integer part is exact, so the box
// error of C_ij is sum_k (A_ik e_kj 2^-beta_j + d_ik B_kj 2^-alpha_i +
// d_ik e_kj 2^-(alpha_i+beta_j)), at most
//   2^-(alpha_i+beta_j) (u_i + q_tj / 2),
//   u_i = 2^alpha_i |A_box,i|_1 / 2 + kb_i / 4,  q_tj = 2^beta_j |B[0:256t, j]|_1
// for row i in tile t.  Its Frobenius norm is summed tile by tile from the
// row and column terms of k_bound_rows / k_bound_cols:
//   E_box^2 = sum_t [ru2_t cv + ru1_t cq1_t + rw_t cq2_t / 4].
// The FP64 diagonal blocks (at most 257 products per element) err by at most
// gamma_257 |A_d,i|_2 |B_t,j|_2 per element: gamma_257 sqrt(sum_t rd2_t cb_t).
// Forming C' in FP64 (two roundings), and adding it to the diagonal part
// (one), add 3u |C_box| + u |C_diag| per element, with |C_box,ij| <= c_max
// 2^-(alpha_i+beta_j) and |C_diag,ij| <= |A_d,i| |B_t,j|.  The absolute
// bound b is the sum of these norms (plus 2^-1073 sqrt(nm) for results in the
// subnormal range), and the relative one b / (||C||_F - b) >= the error
// relative to the exact product's norm.
//
// estimate (expected value).  The same first-order terms with d, e taken as
// independent and uniform on [-1/2, 1/2] (variance 1/12):
//   E ||E_box||_F^2 = (sum_i |A_box,i|^2 cv + rw cn) / 12,  cn = |B[0:kmax]|_F^2,
// plus the FP64 parts as independent roundings of relative size u: the
// diagonal blocks' dot products (u^2/6 sum_t rd2_t cb_t) and the final
// roundings (u^2 ||C||_F^2).  Measured errors of random inputs sit at about
// 0.9x of it, where the bound is typically 30-150x the error (sqrt(K) from
// the worst-case signs).
__global__ __launch_bounds__(1024) void k_bound_final(Desc one, const Desc* table, const int* lut)
{
    __shared__ xd lds[16];
    __shared__ int flag;
    const Desc d = fetch(one, table, lut);
    int nf = 0;
    // ||C||^2 from the reconstruction's partial sums: eight independent
    // loads in flight per thread
    xd c2{0.0, 0};
    const int64_t np = d.count[K_REC];
    for (int64_t k0 = (int64_t)threadIdx.x * 8; k0 < np; k0 += 1024 * 8) {
        xd v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = k0 + u < np ? load_xd(d.cpart + k0 + u) : xd{0.0, 0};
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (v[u].f != v[u].f) nf = 1;
            else c2 = xd_add(c2, v[u]);
        }
    }
    const xd cv = d.tcol[0].cv;
    xd e2{0.0, 0}, ra2{0.0, 0}, rw{0.0, 0}, cn{0.0, 0}, ds{0.0, 0};
    for (int t = threadIdx.x; t < d.nrt; t += 1024) {
        const TileRow r = d.trow[t];
        const TileCol c = d.tcol[t];
        nf |= r.nonfinite | c.nonfinite;
        if (t >= 1) {
            e2 = xd_add(e2, xd_add(xd_mul(r.ru2, cv),
                                   xd_add(xd_mul(r.ru1, c.cq1), xd_mul(r.rw, xd{c.cq2.f, c.cq2.e - 2}))));
            rw = xd_add(rw, r.rw);
        }
        ra2 = xd_add(ra2, r.ra2);
        if (t < d.nrt - 1) cn = xd_add(cn, c.cb);
        ds = xd_add(ds, xd_mul(r.rd2, c.cb));
    }
    c2 = block_sum(c2, lds);
    e2 = block_sum(e2, lds);
    ra2 = block_sum(ra2, lds);
    rw = block_sum(rw, lds);
    cn = block_sum(cn, lds);
    ds = block_sum(ds, lds);
    nf = block_or(nf, &flag);
    if (threadIdx.x != 0) return;
    const double u = 0x1p-53;
    const double g257 = 257.0 * u / (1.0 - 257.0 * u);
    const xd sd = xd_sqrt(ds);
    xd babs = xd_sqrt(e2);
    babs = xd_add(babs, xd_mul(xd_from(g257 + 1.01 * u), sd));
    babs = xd_add(babs, xd_mul(xd_from(3.0 * u * d.c_max), xd_sqrt(xd_mul(rw, cv))));
    babs = xd_add(babs, xd_mul(xd_pow2(-1073), xd_sqrt(xd_from((double)d.n * (double)d.m))));
    const xd cnorm = xd_sqrt(c2);
    // expected squares: integer roundings (variance 1/12 each), FP64 dot
    // products of the diagonal blocks (~u^2/6 |A_d,i|^2 |B_t,j|^2 per
    // element), and the three final roundings (~u^2/3 |C_ij|^2 each)
    xd est2 = xd_mul(xd_add(xd_mul(ra2, cv), xd_mul(rw, cn)), xd_from(1.0 / 12.0));
    est2 = xd_add(est2, xd_mul(ds, xd_from(u * u / 6.0)));
    est2 = xd_add(est2, xd_mul(c2, xd_from(u * u)));
    const xd est = xd_sqrt(est2);
    double bound, estimate;
    if (nf) {
        bound = estimate = nan_value();
    } else {
        bound = xd_lt(babs, cnorm) ? xd_val(xd_div(babs, xd_add(cnorm, xd{-babs.f, babs.e})))
                                   : __builtin_inf();
        estimate = cnorm.f > 0.0 ? xd_val(xd_div(est, cnorm)) : (est.f > 0.0 ? __builtin_inf() : 0.0);
    }
    UVT_GLOBAL double* e = gbl(d.err);
    e[0] = bound;
    e[1] = estimate;
    e[2] = xd_val(cnorm);
    e[3] = xd_val(babs);
}

// ---- host side ------------------------------------------------------------------

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

Mods make_mods(int nmod, int l0, int nl)
{
    Mods md;
    memset(&md, 0, sizeof md);
    md.n = nmod;
    md.l0 = l0;
    md.nl = nl;
    for (int l = 0; l < nmod; ++l) {
        md.m[l] = kModuli[l];
        md.invf[l] = 1.0f / (float)kModuli[l];
        md.invd[l] = 1.0 / (double)kModuli[l];
    }
    return md;
}

// Largest 2-norm (2^lt) of a scaled and rounded row of A' or column of B'
// for nmod moduli: lt = (log2 M - 1)/2 (real, not rounded down), so
// |A'_i B'_j| <= |A'_i| |B'_j| <= 2^(2 lt) < M/2 and the residues determine
// the signed product uniquely; every entry stays below 2^52.
double split_lt(int nmod)
{
    double log2M = 0.0;
    for (int l = 0; l < nmod; ++l) log2M += log2((double)kModuli[l]);
    double lt = (log2M - 1.0) / 2.0 - 1e-6;  // margin for the rounding of log2M
    if (lt > 52.0) lt = 52.0;
    return lt;
}
// Norm every scaled row of A' and column of B' stays below before rounding
// to integers (rounding k entries adds at most sqrt(k)/2 to a 2-norm).
double split_target(int64_t kmax, int nmod) { return exp2(split_lt(nmod)) - 0.5 * sqrt((double)(kmax + 1)); }

// Fewer moduli than this problem needs: the scaled operands would keep
// fewer than 6 bits.
bool too_few_moduli(int64_t n, int nmod)
{
    const int64_t kmax = ((n - 1) / kTileI) * kTileI;
    return kmax > 0 && split_target(kmax, nmod) < 64.0;
}

// Row tiles per XCD group of the int8 products.  8 x 4 groups share panels
// best through an XCD's L2 (measured); another shape is taken only when it
// leaves at least 15 % fewer group slots over rtiles x ctiles tiles.
int group_rows(int64_t rtiles, int64_t ctiles)
{
    static const int pref[] = {8, 4, 16, 2, 32, 1};
    auto slots = [&](int gr) {
        const int64_t gc = 32 / gr;
        return (rtiles + gr - 1) / gr * ((ctiles + gc - 1) / gc);
    };
    int64_t best = INT64_MAX;
    for (int gr : pref) best = std::min(best, slots(gr));
    for (int gr : pref)
        if (slots(gr) * 100 <= best * 115) return gr;
    return 8;
}

int64_t kmax_of(int64_t n) { return ((n - 1) / kTileI) * kTileI; }

// Workgroups of one product launch over nl moduli: XCD groups of gr x 32/gr
// tiles per modulus, 32 workgroups each, padded to a multiple of 256 (8
// groups) so that every problem starts on XCD 0.
int64_t gemm_blocks(const Desc& d, int nl)
{
    const int64_t groups = (int64_t)((d.nrt - 1 + d.gr - 1) / d.gr) * ((d.nct + 32 / d.gr - 1) / (32 / d.gr)) * nl;
    return (groups + 7) / 8 * 8 * 32;
}

// Moduli per residue-B buffer of a single-problem call: all of them, or half
// when the residues of B would exceed kBGroupBytes.
int moduli_per_group(int64_t n, int64_t m, int nmod)
{
    const size_t all = (size_t)nmod * (size_t)pad_tile(m) * (size_t)kmax_of(n);
    return all > kBGroupBytes ? (nmod + 1) / 2 : nmod;
}

// Workspace layout of one problem; every piece 256-byte aligned.
struct Layout {
    size_t alpha, beta, bnf, rowq, part, pre, trow, tcol, cpart, err, Ar, Br, R, total;
};

Layout layout(int64_t n, int64_t m, int nmod, int per_group)
{
    const int64_t kmax = kmax_of(n);
    const int64_t nrt = (n + kTileI - 1) / kTileI;
    const int64_t gxr = (m + 1023) / 1024;
    Layout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.alpha = take((size_t)n * 4);
    L.beta = take((size_t)m * 4);
    L.bnf = take((size_t)m * 4);
    L.rowq = take((size_t)n * sizeof(RowQ));
    L.part = take((size_t)nrt * m * sizeof(Part));
    L.pre = take((size_t)nrt * m * sizeof(xd));
    L.trow = take((size_t)nrt * sizeof(TileRow));
    L.tcol = take((size_t)nrt * sizeof(TileCol));
    L.cpart = take((size_t)n * gxr * sizeof(xd));
    L.err = take(4 * sizeof(double));
    L.Ar = take(kmax > 0 ? (size_t)nmod * a_mod_bytes(n) : 0);
    L.Br = take(kmax > 0 ? (size_t)per_group * pad_tile(m) * kmax : 0);
    L.R = take(kmax > 0 ? (size_t)nmod * (n - kTileI) * m : 0);
    L.total = o;
    return L;
}

uvt_status check_problem(const uvt_split_problem& p)
{
    if (!p.A || !p.B || !p.C || p.n <= 0 || p.m <= 0 || p.lda < p.n || p.ldb < p.m || p.ldc < p.m)
        return UVT_ERR_ARG;
    if (p.n > 131072 || p.m > 0x3FFFFFFF) return UVT_ERR_SIZE;  // int32 accumulation bound on k
    return UVT_OK;
}

// Fills the descriptors (workspace pointers from `work` on, per_group moduli
// per residue-B buffer) and the stage workgroup counts; returns the
// per-stage totals.
void describe(const uvt_split_problem* probs, int count, int nmod, int per_group, char* work,
              Desc* out, int64_t totals[K_NUM])
{
    for (int k = 0; k < K_NUM; ++k) totals[k] = 0;
    size_t off = 0;
    const double lt = split_lt(nmod);
    for (int p = 0; p < count; ++p) {
        const uvt_split_problem& q = probs[p];
        Desc& d = out[p];
        memset(&d, 0, sizeof d);
        d.A = q.A;
        d.B = q.B;
        d.C = q.C;
        d.lda = q.lda;
        d.ldb = q.ldb;
        d.ldc = q.ldc;
        d.n = q.n;
        d.m = q.m;
        d.kmax = kmax_of(q.n);
        d.target = split_target(d.kmax, nmod);
        d.c_max = exp2(2.0 * lt) * (1.0 + 1e-6);
        d.nrt = (int32_t)((q.n + kTileI - 1) / kTileI);
        d.nct = (int32_t)((q.m + kTileI - 1) / kTileI);
        d.nrt128 = (int32_t)((q.n + 127) / 128);
        d.nct128 = (int32_t)((q.m + 127) / 128);
        d.gxa = (int32_t)(d.kmax / 128);
        d.gxr = (int32_t)((q.m + 1023) / 1024);
        d.vec = ((uintptr_t)q.A % 16 == 0) && ((uintptr_t)q.B % 16 == 0) && q.lda % 2 == 0 &&
                q.ldb % 2 == 0;
        const Layout L = layout(q.n, q.m, nmod, per_group);
        char* w = work ? work + off : nullptr;
        d.alpha = (int*)(w + L.alpha);
        d.beta = (int*)(w + L.beta);
        d.bnf = (int*)(w + L.bnf);
        d.rowq = (RowQ*)(w + L.rowq);
        d.part = (Part*)(w + L.part);
        d.pre = (xd*)(w + L.pre);
        d.trow = (TileRow*)(w + L.trow);
        d.tcol = (TileCol*)(w + L.tcol);
        d.cpart = (xd*)(w + L.cpart);
        d.err = (double*)(w + L.err);
        d.Ar = (int8_t*)(w + L.Ar);
        d.Br = (int8_t*)(w + L.Br);
        d.R = (uint8_t*)(w + L.R);
        off += L.total;
        const int64_t ncb = (q.m + 255) / 256;
        d.count[K_DIAG] = d.nrt128 * d.nct128;
        d.count[K_ALPHA] = (int32_t)((q.n + 3) / 4);
        d.count[K_BPART] = (int32_t)(ncb * d.nrt);
        d.count[K_BETA] = (int32_t)ncb;
        d.count[K_BROWS] = d.nrt;
        d.count[K_BCOLS] = d.nrt;
        d.count[K_REC] = (int32_t)(q.n * d.gxr);
        d.count[K_BFIN] = 1;
        if (d.kmax > 0) {
            d.count[K_RA] = (int32_t)((pad_tile(q.n) - kTileI) / 32 * d.gxa);
            d.count[K_RB] = (int32_t)((d.kmax / 128) * (pad_tile(q.m) / 64));
            d.gr = group_rows(d.nrt - 1, d.nct);
            d.count[K_GEMM] = (int32_t)gemm_blocks(d, per_group);
        }
        for (int k = 0; k < K_NUM; ++k) {
            d.start[k] = (int32_t)totals[k];
            totals[k] += d.count[k];
        }
    }
}

}  // namespace

// ---- shared with uvt_batch.hip and uvt_serve.hip -------------------------------

bool uvt_split_supported(int64_t n, int nmod)
{
    return n > 0 && n <= 131072 && nmod >= 2 && nmod <= kMaxSlices && !too_few_moduli(n, nmod);
}

size_t uvt_split_desc_bytes(int count) { return align256((size_t)count * sizeof(Desc)); }

size_t uvt_split_err_offset(int64_t n, int64_t m, int nmod)
{
    return layout(n, m, nmod, moduli_per_group(n, m, nmod)).err;
}

uvt_status uvt_split_batch_sizes(const uvt_split_problem* probs, int count, int nmod,
                                 size_t* work_bytes, int64_t* lut_entries)
{
    if (nmod < 2 || nmod > kMaxSlices || count < 0) return UVT_ERR_ARG;
    size_t w = 0;
    for (int p = 0; p < count; ++p) {
        uvt_status st = check_problem(probs[p]);
        if (st != UVT_OK) return st;
        if (too_few_moduli(probs[p].n, nmod)) return UVT_ERR_SIZE;
        w += layout(probs[p].n, probs[p].m, nmod, nmod).total;
    }
    std::vector<Desc> d(count);
    int64_t tot[K_NUM];
    describe(probs, count, nmod, nmod, nullptr, d.data(), tot);
    int64_t lut = 0;
    for (int k = 0; k < K_NUM; ++k) {
        if (tot[k] > 0x7FFFFFFF) return UVT_ERR_SIZE;
        lut += tot[k];
    }
    *work_bytes = w;
    *lut_entries = lut;
    return UVT_OK;
}

uvt_status uvt_split_batch_build(const uvt_split_problem* probs, int count, int nmod, void* dwork,
                                 std::vector<char>& desc_img, std::vector<int>& lut_img,
                                 int64_t totals[UVT_SPLIT_STAGES], int64_t lut_off[UVT_SPLIT_STAGES],
                                 std::vector<double*>* err)
{
    std::vector<Desc> d(count);
    describe(probs, count, nmod, nmod, (char*)dwork, d.data(), totals);
    desc_img.assign((const char*)d.data(), (const char*)d.data() + count * sizeof(Desc));
    lut_img.clear();
    for (int k = 0; k < K_NUM; ++k) {
        lut_off[k] = (int64_t)lut_img.size();
        for (int p = 0; p < count; ++p)
            for (int32_t b = 0; b < d[p].count[k]; ++b) lut_img.push_back(p);
    }
    if (err) {
        err->clear();
        for (int p = 0; p < count; ++p) err->push_back(d[p].err);
    }
    return UVT_OK;
}

uvt_status uvt_split_batch_launch(const uvt_split_problem* single, int count, int nmod,
                                  void* dwork, const void* ddesc, const int* dlut,
                                  const int64_t totals_in[UVT_SPLIT_STAGES],
                                  const int64_t lut_off[UVT_SPLIT_STAGES], hipStream_t s, int stages)
{
    Desc one;
    memset(&one, 0, sizeof one);
    int64_t totals[K_NUM];
    const Desc* table = (const Desc*)ddesc;
    // one problem launched directly may split the residues of B in two
    // groups of moduli; a batch keeps them all (one product launch)
    int per_group = nmod;
    if (!table) {
        if (count != 1) return UVT_ERR_ARG;
        per_group = moduli_per_group(single->n, single->m, nmod);
        describe(single, 1, nmod, per_group, (char*)dwork, &one, totals);
    } else {
        for (int k = 0; k < K_NUM; ++k) totals[k] = totals_in[k];
    }
    const Mods mods = make_mods(nmod, 0, nmod);
    auto lut = [&](int k) { return table ? dlut + lut_off[k] : (const int*)nullptr; };
    auto launch = [&](auto kernel, int k, auto... args) {
        if (totals[k])
            hipLaunchKernelGGL(kernel, dim3((unsigned)totals[k]), dim3(256), 0, s, one, table, lut(k), args...);
    };
    // The FP64 diagonal part stores every row of C first; the box part adds to it.
    if ((stages & UVT_SPLIT_DIAG) && totals[K_DIAG])
        hipLaunchKernelGGL(k_diag_f64, dim3((unsigned)totals[K_DIAG]), dim3(512), 0, s, one, table, lut(K_DIAG));
    if (stages & UVT_SPLIT_PREP) {
        launch(k_alpha, K_ALPHA);
        launch(k_beta_part, K_BPART);
        launch(k_beta, K_BETA);
        launch(k_bound_rows, K_BROWS);
        launch(k_bound_cols, K_BCOLS);
        launch(k_resid_a, K_RA, mods);
        if (per_group == nmod) launch(k_resid_b, K_RB, mods);
    }
    if (stages & UVT_SPLIT_GEMM) {
        for (int l0 = 0; l0 < nmod; l0 += per_group) {
            const int nl = std::min(per_group, nmod - l0);
            const Mods g = make_mods(nmod, l0, nl);
            if (per_group != nmod) launch(k_resid_b, K_RB, g);
            // a short last group has fewer product workgroups than described
            const int64_t ng = nl == per_group ? totals[K_GEMM] : gemm_blocks(one, nl);
            if (ng)
                hipLaunchKernelGGL(k_gemm_i8, dim3((unsigned)ng), dim3(512), 0, s, one, table, lut(K_GEMM), g);
        }
    }
    if ((stages & UVT_SPLIT_REC) && totals[K_REC]) {
        const unsigned grid = (unsigned)totals[K_REC];
        switch (nmod) {
        case 2: launch_reconstruct<2>(grid, s, one, table, lut(K_REC)); break;
        case 3: launch_reconstruct<3>(grid, s, one, table, lut(K_REC)); break;
        case 4: launch_reconstruct<4>(grid, s, one, table, lut(K_REC)); break;
        case 5: launch_reconstruct<5>(grid, s, one, table, lut(K_REC)); break;
        case 6: launch_reconstruct<6>(grid, s, one, table, lut(K_REC)); break;
        case 7: launch_reconstruct<7>(grid, s, one, table, lut(K_REC)); break;
        case 8: launch_reconstruct<8>(grid, s, one, table, lut(K_REC)); break;
        case 9: launch_reconstruct<9>(grid, s, one, table, lut(K_REC)); break;
        case 10: launch_reconstruct<10>(grid, s, one, table, lut(K_REC)); break;
        case 11: launch_reconstruct<11>(grid, s, one, table, lut(K_REC)); break;
        case 12: launch_reconstruct<12>(grid, s, one, table, lut(K_REC)); break;
        case 13: launch_reconstruct<13>(grid, s, one, table, lut(K_REC)); break;
        default: launch_reconstruct<14>(grid, s, one, table, lut(K_REC)); break;
        }
        if (totals[K_BFIN])
            hipLaunchKernelGGL(k_bound_final, dim3((unsigned)totals[K_BFIN]), dim3(1024), 0, s, one, table,
                               lut(K_BFIN));
    }
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

extern "C" uvt_status uvt_trmm_split_workspace(int64_t n, int64_t m, int32_t slices, size_t* bytes)
{
    if (!bytes || n <= 0 || m <= 0 || slices < 2 || slices > kMaxSlices) return UVT_ERR_ARG;
    if (n > 131072 || m > 0x3FFFFFFF) return UVT_ERR_SIZE;  // int32 accumulation bound on k
    if (too_few_moduli(n, slices)) return UVT_ERR_SIZE;
    *bytes = layout(n, m, slices, moduli_per_group(n, m, slices)).total;
    return UVT_OK;
}

extern "C" uvt_status uvt_trmm_split(uvt_handle h, int64_t n, int64_t m, int32_t slices,
                                     const double* dA, int64_t lda, const double* dB, int64_t ldb,
                                     double* dC, int64_t ldc, void* dwork, size_t work_bytes,
                                     void* stream)
{
    if (!h) return UVT_ERR_ARG;
    const uvt_split_problem p = {dA, lda, dB, ldb, dC, ldc, n, m};
    uvt_status st = check_problem(p);
    if (st != UVT_OK) return st;
    size_t need = 0;
    st = uvt_trmm_split_workspace(n, m, slices, &need);
    if (st != UVT_OK) return st;
    if (!dwork || work_bytes < need || ((uintptr_t)dwork & 255)) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    // One stream, stage after stage: running the FP64 diagonal blocks beside
    // the HBM-bound scale and residue stages saves nothing (measured), and
    // the int8 products want every CU to themselves.
    const int64_t zero[K_NUM] = {0};
    return uvt_split_batch_launch(&p, 1, slices, dwork, nullptr, nullptr, zero, zero,
                                  (hipStream_t)stream, UVT_SPLIT_ALL);
}

extern "C" uvt_status uvt_trmm_split_error(uvt_handle h, int64_t n, int64_t m, int32_t slices,
                                           const void* dwork, size_t work_bytes, double* bound,
                                           double* estimate, void* stream)
{
    if (!h || !dwork) return UVT_ERR_ARG;
    size_t need = 0;
    uvt_status st = uvt_trmm_split_workspace(n, m, slices, &need);
    if (st != UVT_OK) return st;
    if (work_bytes < need) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    double e[2];
    UVT_HIP(hipStreamSynchronize((hipStream_t)stream));
    UVT_HIP(hipMemcpy(e, (const char*)dwork + uvt_split_err_offset(n, m, slices), sizeof e,
                      hipMemcpyDeviceToHost));
    if (bound) *bound = e[0];
    if (estimate) *estimate = e[1];
    return UVT_OK;
}
