// uvt_blas.hip -- device-resident BLAS forms of the two routines the
// reference calls per band (UVaTRMM/src/mm_triangular.c:347-360 cblas_dgemm,
// :374-387 cblas_dtrmm), row-major, every transpose / side / triangle / diagonal
// variant.  All of them run on the FP64 MFMA tile kernel of uvt_f64_tile.hpp:
//
//   dgemm  C = alpha op(A) op(B) + beta C
//          op(A), op(B) packed row-major when transposed (LDS-tiled transpose),
//          then the rectangular tile walk (no triangle mask), then the alpha /
//          beta epilogue (C is not read when beta == 0).
//   dtrmm  B := alpha op(A) B  (left)  or  alpha B op(A)  (right), in place
//          reduced to  R = tril(L) X  with
//            right side:   X = B^T, and op(A) -> op(A)^T
//            upper T:      P T P is lower for the row/column reversal P, so
//                          T X = P (P T P) (P X): rows of X and of the result
//                          are reversed on packing / unpacking
//          L packed with its strict upper part zeroed and a unit diagonal
//          written in when asked; the result unpacked into B with alpha.
// Layout-only reductions (column-major input) are the caller's (the CBLAS
// shim): a column-major matrix is the transpose of a row-major one.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "uvt_f64_tile.hpp"
#include "uvt_internal.h"
#include "uvt_runtime.hpp"

using namespace uvt;

namespace {

constexpr int PT = 64;  // pack tile edge

// dst[i][j] = alpha * S(rmap(i), cmap(j)) for i < rows, j < cols, where
// S(a, b) = src[a*lds + b] (trans = 0) or src[b*lds + a] (trans = 1),
// rmap(i) = rev_r ? rows-1-i : i, cmap(j) = rev_c ? cols-1-j : j.
// tri: only the lower triangle (j <= i) is read; above it dst = 0, and with
// unit the diagonal is 1.  The transposed read goes through an LDS tile so
// both the global reads and the writes are row-contiguous.
__global__ __launch_bounds__(256) void k_pack(const double* __restrict__ src, int64_t lds,
                                              double* __restrict__ dst, int64_t ldd, int64_t rows,
                                              int64_t cols, int trans, int rev_r, int rev_c,
                                              double alpha, int tri, int unit)
{
    __shared__ double tile[PT][PT + 1];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int64_t bi = blockIdx.y; bi * PT < rows; bi += gridDim.y) {
        const int64_t i0 = bi * PT, j0 = (int64_t)blockIdx.x * PT;
        if (tri && j0 > i0 + PT - 1) {  // strictly above the diagonal: zeros
            for (int r = ty; r < PT; r += 4) {
                const int64_t i = i0 + r, j = j0 + tx;
                if (i < rows && j < cols) dst[i * ldd + j] = 0.0;
            }
            continue;
        }
        if (!trans) {
            for (int r = ty; r < PT; r += 4) {
                const int64_t i = i0 + r, j = j0 + tx;
                if (i >= rows || j >= cols) continue;
                double v;
                if (tri && j > i) v = 0.0;
                else if (unit && i == j) v = 1.0;
                else {
                    const int64_t a = rev_r ? rows - 1 - i : i, b = rev_c ? cols - 1 - j : j;
                    v = alpha * src[a * lds + b];
                }
                dst[i * ldd + j] = v;
            }
        } else {
            // read S(rmap(i), cmap(j)) = src[cmap(j)*lds + rmap(i)]: for a fixed
            // j the reads over i are contiguous (reversed when rev_r)
            for (int c = ty; c < PT; c += 4) {
                const int64_t j = j0 + c, i = i0 + tx;
                double v = 0.0;
                if (i < rows && j < cols && !(tri && j > i)) {
                    const int64_t a = rev_r ? rows - 1 - i : i, b = rev_c ? cols - 1 - j : j;
                    v = src[b * lds + a];
                }
                tile[c][tx] = v;
            }
            __syncthreads();
            for (int r = ty; r < PT; r += 4) {
                const int64_t i = i0 + r, j = j0 + tx;
                if (i >= rows || j >= cols) continue;
                double v;
                if (tri && j > i) v = 0.0;
                else if (unit && i == j) v = 1.0;
                else v = alpha * tile[tx][r];
                dst[i * ldd + j] = v;
            }
            __syncthreads();
        }
    }
}

// C = alpha * P + beta * C (P == nullptr: alpha * 0).  beta == 0: C is not read.
__global__ __launch_bounds__(256) void k_axpby(int64_t rows, int64_t cols, double alpha,
                                               const double* __restrict__ P, int64_t ldp,
                                               double beta, double* __restrict__ C, int64_t ldc)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= cols) return;
    for (int64_t i = blockIdx.y; i < rows; i += gridDim.y) {
        const double p = P ? alpha * P[i * ldp + j] : 0.0;
        double* c = C + i * ldc + j;
        *c = beta == 0.0 ? p : p + beta * *c;
    }
}

// Rectangle C[0:M, 0:N] = A[0:M, 0:K] * B[0:K, 0:N] on the FP64 tile.
template <int TM, int TN, bool VEC>
__global__ __launch_bounds__((64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N),
                             (TileWaves<TM, TN>::WAVES_PER_EU))
void k_gemm_f64(int64_t M, int64_t N, int K, int nct, const double* __restrict__ A, int64_t lda,
                const double* __restrict__ B, int64_t ldb, double* __restrict__ C, int64_t ldc)
{
    using W = TileWaves<TM, TN>;
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<TM, TN, W::M, W::N>::SMEM_DOUBLES];
    const int rt = (int)(blockIdx.x / nct), ct = (int)(blockIdx.x % nct);
    const int r0 = rt * TM;
    const int rows = (int)(M - r0 < TM ? M - r0 : TM);
    f64_tile<TM, TN, VEC, W::M, W::N>(r0, rows, ct * TN, 0, K, 0, gbl(A), lda, gbl(B), ldb, gbl(C),
                                      ldc, N, smem);
}

template <int TM, int TN>
void launch_gemm(bool vec, int64_t M, int64_t N, int64_t K, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double* C, int64_t ldc, hipStream_t s)
{
    const int nrt = (int)((M + TM - 1) / TM), nct = (int)((N + TN - 1) / TN);
    const dim3 grid((unsigned)((int64_t)nrt * nct));
    const dim3 block(64 * TileWaves<TM, TN>::M * TileWaves<TM, TN>::N);
    if (vec)
        hipLaunchKernelGGL((k_gemm_f64<TM, TN, true>), grid, block, 0, s, M, N, (int)K, nct, A, lda,
                           B, ldb, C, ldc);
    else
        hipLaunchKernelGGL((k_gemm_f64<TM, TN, false>), grid, block, 0, s, M, N, (int)K, nct, A,
                           lda, B, ldb, C, ldc);
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }
int64_t even(int64_t x) { return x + (x & 1); }
size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

void pack(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t rows, int64_t cols,
          bool trans, bool rev_r, bool rev_c, double alpha, bool tri, bool unit, hipStream_t s)
{
    const int64_t gy = std::min<int64_t>((rows + PT - 1) / PT, 32768);
    const dim3 grid((unsigned)((cols + PT - 1) / PT), (unsigned)gy);
    hipLaunchKernelGGL(k_pack, grid, dim3(256), 0, s, src, lds, dst, ldd, rows, cols, (int)trans,
                       (int)rev_r, (int)rev_c, alpha, (int)tri, (int)unit);
}

void axpby(int64_t rows, int64_t cols, double alpha, const double* P, int64_t ldp, double beta,
           double* C, int64_t ldc, hipStream_t s)
{
    const dim3 grid((unsigned)((cols + 255) / 256), (unsigned)std::min<int64_t>(rows, 32768));
    hipLaunchKernelGGL(k_axpby, grid, dim3(256), 0, s, rows, cols, alpha, P, ldp, beta, C, ldc);
}

bool sizes_ok(int64_t a, int64_t b, int64_t c)
{
    const int64_t lim = 0x3FFFFFFF;
    return a <= lim && b <= lim && c <= lim;
}

}  // namespace

uvt_status uvt_launch_gemm_f64(int32_t cfg, int64_t M, int64_t N, int64_t K, const double* A,
                               int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc,
                               hipStream_t s)
{
    if (M <= 0 || N <= 0 || K <= 0 || lda < K || ldb < N || ldc < N) return UVT_ERR_ARG;
    if (!sizes_ok(M, N, K)) return UVT_ERR_SIZE;
    const int64_t tm = uvt_cfg_tile_m(cfg), tn = uvt_cfg_tile_n(cfg);
    if (((M + tm - 1) / tm) * ((N + tn - 1) / tn) > 0x7FFFFFFF) return UVT_ERR_SIZE;
    const bool vec = aligned16(A) && aligned16(B) && lda % 2 == 0 && ldb % 2 == 0;
    if (cfg == UVT_CFG_64x64) launch_gemm<64, 64>(vec, M, N, K, A, lda, B, ldb, C, ldc, s);
    else if (cfg == UVT_CFG_128x32) launch_gemm<128, 32>(vec, M, N, K, A, lda, B, ldb, C, ldc, s);
    else launch_gemm<128, 128>(vec, M, N, K, A, lda, B, ldb, C, ldc, s);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

extern "C" uvt_status uvt_dgemm(uvt_handle h, uvt_trans ta, uvt_trans tb, int64_t m, int64_t n,
                                int64_t k, double alpha, const double* dA, int64_t lda,
                                const double* dB, int64_t ldb, double beta, double* dC,
                                int64_t ldc, void* stream)
{
    if (!h || m < 0 || n < 0 || k < 0) return UVT_ERR_ARG;
    if ((ta != UVT_NO_TRANS && ta != UVT_TRANS) || (tb != UVT_NO_TRANS && tb != UVT_TRANS))
        return UVT_ERR_ARG;
    const bool tA = ta == UVT_TRANS, tB = tb == UVT_TRANS;
    if (lda < std::max<int64_t>(1, tA ? m : k) || ldb < std::max<int64_t>(1, tB ? k : n) ||
        ldc < std::max<int64_t>(1, n))
        return UVT_ERR_ARG;
    if (m == 0 || n == 0) return UVT_OK;
    if (!dC || (k > 0 && alpha != 0.0 && (!dA || !dB))) return UVT_ERR_ARG;
    if (!sizes_ok(m, n, k)) return UVT_ERR_SIZE;
    UVT_HIP(hipSetDevice(h->device));
    hipStream_t s = (hipStream_t)stream;
    if (k == 0 || alpha == 0.0) {  // C = beta C, A and B not read
        axpby(m, n, 0.0, nullptr, 0, beta, dC, ldc, s);
        UVT_HIP(hipGetLastError());
        return UVT_OK;
    }
    const bool direct = alpha == 1.0 && beta == 0.0;
    const int64_t lap = even(k), lbp = even(n), lpp = even(n);
    const size_t a_bytes = tA ? up256((size_t)m * lap * 8) : 0;
    const size_t b_bytes = tB ? up256((size_t)k * lbp * 8) : 0;
    const size_t p_bytes = direct ? 0 : up256((size_t)m * lpp * 8);
    void* w = nullptr;
    if (a_bytes + b_bytes + p_bytes) {
        uvt_status st = uvt_handle_workspace(h, a_bytes + b_bytes + p_bytes, &w);
        if (st != UVT_OK) return st;
    }
    char* wp = (char*)w;
    const double* Ap = dA;
    int64_t la = lda;
    if (tA) {  // op(A) = A^T: A stored k x m
        pack(dA, lda, (double*)wp, lap, m, k, true, false, false, 1.0, false, false, s);
        Ap = (const double*)wp;
        la = lap;
        wp += a_bytes;
    }
    const double* Bp = dB;
    int64_t lb = ldb;
    if (tB) {  // op(B) = B^T: B stored n x k
        pack(dB, ldb, (double*)wp, lbp, k, n, true, false, false, 1.0, false, false, s);
        Bp = (const double*)wp;
        lb = lbp;
        wp += b_bytes;
    }
    double* P = direct ? dC : (double*)wp;
    const int64_t lp = direct ? ldc : lpp;
    uvt_status st = uvt_launch_gemm_f64(uvt_default_cfg(m, n), m, n, k, Ap, la, Bp, lb, P, lp, s);
    if (st != UVT_OK) return st;
    if (!direct) axpby(m, n, alpha, P, lp, beta, dC, ldc, s);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

// ---- dtrmm ---------------------------------------------------------------------

// Scratch of uvt_dtrmm_noalloc: a fixed device array, one copy per device,
// held until the process ends.  Host code never passes `scratch` itself to a
// kernel (in host code it is only a stand-in for the device symbol); it uses
// the address hipGetSymbolAddress returned for the handle's device, which
// uvt_noalloc_setup stores in the handle.
constexpr size_t kNoAllocScratchBytes = (size_t)UVT_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];  // per device, set once

// One dtrmm call, reduced to R = tril(L) X (see the comment at the top of
// this file): which operands are packed and the workspace bytes of each.
struct DtrmmCall {
    bool right, tA, rev, unit, tX, packL, packX;
    int64_t nt, w, ll, lx, lr;
    size_t l_bytes, x_bytes, r_bytes;
};

// Validates the arguments and sizes the workspace.  UVT_OK with
// *nothing = true when B is empty (no work at all).
uvt_status dtrmm_prepare(uvt_handle h, uvt_side side, uvt_uplo uplo, uvt_trans ta, uvt_diag diag,
                         int64_t m, int64_t n, double alpha, const double* dA, int64_t lda,
                         double* dB, int64_t ldb, DtrmmCall* c, bool* nothing)
{
    *nothing = false;
    if (!h || m < 0 || n < 0) return UVT_ERR_ARG;
    if ((side != UVT_LEFT && side != UVT_RIGHT) || (uplo != UVT_LOWER && uplo != UVT_UPPER) ||
        (ta != UVT_NO_TRANS && ta != UVT_TRANS) || (diag != UVT_NON_UNIT && diag != UVT_UNIT))
        return UVT_ERR_ARG;
    c->right = side == UVT_RIGHT;
    c->nt = c->right ? n : m;  // order of the triangle
    c->w = c->right ? m : n;   // columns of X
    if (lda < std::max<int64_t>(1, c->nt) || ldb < std::max<int64_t>(1, n)) return UVT_ERR_ARG;
    if (m == 0 || n == 0) {
        *nothing = true;
        return UVT_OK;
    }
    if (!dB || (alpha != 0.0 && !dA)) return UVT_ERR_ARG;
    if (!sizes_ok(m, n, 1)) return UVT_ERR_SIZE;
    // T = op(A) (left) or op(A)^T (right), the nt x nt factor applied to X
    c->tA = (ta == UVT_TRANS) != c->right;
    const bool t_lower = (uplo == UVT_LOWER) != c->tA;
    c->rev = !t_lower;  // upper T: reverse rows and columns
    c->unit = diag == UVT_UNIT;
    c->tX = c->right;   // X = B^T
    c->packL = c->tA || c->rev || c->unit;
    c->ll = even(c->nt);
    c->lx = even(c->w);
    c->lr = even(c->w);
    c->l_bytes = c->packL ? up256((size_t)c->nt * c->ll * 8) : 0;
    c->packX = c->tX || c->rev;  // else X is B itself (R is written elsewhere)
    c->x_bytes = c->packX ? up256((size_t)c->nt * c->lx * 8) : 0;
    c->r_bytes = up256((size_t)c->nt * c->lr * 8);
    if (alpha == 0.0) {  // no workspace: B = 0, A not read
        c->l_bytes = c->x_bytes = c->r_bytes = 0;
    }
    return UVT_OK;
}

size_t dtrmm_bytes(const DtrmmCall& c) { return c.l_bytes + c.x_bytes + c.r_bytes; }

// Enqueues the call on stream s with its workspace at wv.
uvt_status dtrmm_launch(const DtrmmCall& c, int64_t m, int64_t n, double alpha, const double* dA,
                        int64_t lda, double* dB, int64_t ldb, void* wv, hipStream_t s)
{
    if (alpha == 0.0) {  // B = 0, A not read
        axpby(m, n, 0.0, nullptr, 0, 0.0, dB, ldb, s);
        UVT_HIP(hipGetLastError());
        return UVT_OK;
    }
    char* wp = (char*)wv;
    const double* L = dA;
    int64_t la = lda;
    if (c.packL) {
        pack(dA, lda, (double*)wp, c.ll, c.nt, c.nt, c.tA, c.rev, c.rev, 1.0, true, c.unit, s);
        L = (const double*)wp;
        la = c.ll;
        wp += c.l_bytes;
    }
    const double* X = dB;
    int64_t lxx = ldb;
    if (c.packX) {
        pack(dB, ldb, (double*)wp, c.lx, c.nt, c.w, c.tX, c.rev, false, 1.0, false, false, s);
        X = (const double*)wp;
        lxx = c.lx;
        wp += c.x_bytes;
    }
    double* R = (double*)wp;
    uvt_status st = uvt_launch_trapezoid_f64(uvt_wide_tiles(uvt_default_cfg(c.nt, c.w), c.nt, c.w), c.nt, c.w, L, la, X, lxx,
                                             R, c.lr, s);
    if (st != UVT_OK) return st;
    // B <- alpha * unpack(R): B(i, j) = R(rmap(i), j) (left) or R(cmap(j), i) (right)
    if (c.tX) pack(R, c.lr, dB, ldb, m, n, true, false, c.rev, alpha, false, false, s);
    else pack(R, c.lr, dB, ldb, m, n, false, c.rev, false, alpha, false, false, s);
    UVT_HIP(hipGetLastError());
    return UVT_OK;
}

}  // namespace

// Looks up this device's copy of `scratch` once per device (with the device
// current, as uvt_create leaves it) and stores its address in the handle.
uvt_status uvt_noalloc_setup(uvt_handle h)
{
    if (h->device < 0 || h->device >= kMaxDevices) return UVT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[h->device]) {
        void* p = nullptr;
        UVT_HIP(hipSetDevice(h->device));
        UVT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return UVT_ERR_HIP;
        g_scratch_ptr[h->device] = p;
    }
    h->scratch = g_scratch_ptr[h->device];
    return UVT_OK;
}

extern "C" uvt_status uvt_dtrmm(uvt_handle h, uvt_side side, uvt_uplo uplo, uvt_trans ta,
                                uvt_diag diag, int64_t m, int64_t n, double alpha,
                                const double* dA, int64_t lda, double* dB, int64_t ldb,
                                void* stream)
{
    DtrmmCall c;
    bool nothing = false;
    uvt_status st = dtrmm_prepare(h, side, uplo, ta, diag, m, n, alpha, dA, lda, dB, ldb, &c, &nothing);
    if (st != UVT_OK || nothing) return st;
    UVT_HIP(hipSetDevice(h->device));
    void* wv = nullptr;
    if (dtrmm_bytes(c)) {
        st = uvt_handle_workspace(h, dtrmm_bytes(c), &wv);
        if (st != UVT_OK) return st;
    }
    return dtrmm_launch(c, m, n, alpha, dA, lda, dB, ldb, wv, (hipStream_t)stream);
}

extern "C" uvt_status uvt_dtrmm_noalloc(uvt_handle h, uvt_side side, uvt_uplo uplo, uvt_trans ta,
                                        uvt_diag diag, int64_t m, int64_t n, double alpha,
                                        const double* dA, int64_t lda, double* dB, int64_t ldb,
                                        void* stream)
{
    DtrmmCall c;
    bool nothing = false;
    uvt_status st = dtrmm_prepare(h, side, uplo, ta, diag, m, n, alpha, dA, lda, dB, ldb, &c, &nothing);
    if (st != UVT_OK || nothing) return st;
    if (dtrmm_bytes(c) > kNoAllocScratchBytes) return UVT_ERR_SIZE;
    if (!h->scratch) return UVT_ERR_STATE;
    UVT_HIP(hipSetDevice(h->device));
    return dtrmm_launch(c, m, n, alpha, dA, lda, dB, ldb, h->scratch, (hipStream_t)stream);
}
