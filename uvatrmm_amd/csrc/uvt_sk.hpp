// uvt_sk.hpp -- stream-K geometry of the trapezoid walk (host and device).
//
// The trapezoid walk of C = tril(A) * B with TM x TN tiles has nrt row tiles
// and nct column tiles; row tile j reduces over k in [0, min(n, (j+1) TM)),
// which is D(j) = ceil(that / 16) panels of 16.  Stream-K (Osama, Merrill,
// Cecka, Garland, Owens, PPoPP 2023) lays every tile's panels end to end --
// row tiles in increasing order, the nct column tiles of a row tile one after
// another, each tile's panels in k order -- into one sequence of P panels,
// and gives workgroup b of G the panels [b P / G, (b+1) P / G).  A workgroup
// thus runs a few whole tiles, possibly a tail of one tile before them and a
// head of one after.  Every workgroup does the same number of panels, so a
// launch whose tiles differ in depth (narrow B: few, deep tiles) finishes
// together instead of waiting for its deepest tile.
//
// Where a tile's panels are split between workgroups:
//   - the piece that starts at k = 0 writes C;
//   - every other piece is the first piece of its workgroup b (a workgroup's
//     later pieces start their tiles), and writes partial tile b;
//   - afterwards one workgroup per split tile adds C + partial tiles in
//     panel (k) order, so the result does not depend on timing.
#pragma once

#include <stdint.h>

namespace uvt {

constexpr int kSkPanel = 16;  // k panel of the FP64 tile kernels
constexpr int kXcds = 8;      // MI355X: workgroup i of a launch runs on XCD i % 8

struct SkGeo {
    int64_t n, m;  // C = tril(A) * B: A n x n, B and C n x m
    int64_t P;     // panels of the whole walk
    int32_t nrt, nct, tm, tn;
    int64_t G;     // workgroups, 1 <= G <= P
};

// Panels before row tile j, 0 <= j <= nrt - 1 (row tiles before the last
// are full: (r + 1) * tm / 16 panels each).
__host__ __device__ inline int64_t sk_before(const SkGeo& g, int j)
{
    return (int64_t)g.nct * (g.tm / kSkPanel) * ((int64_t)j * (j + 1) / 2);
}

// Panels of each tile of row tile j.
__host__ __device__ inline int64_t sk_depth(const SkGeo& g, int j)
{
    return j < g.nrt - 1 ? (int64_t)(j + 1) * (g.tm / kSkPanel) : (g.n + kSkPanel - 1) / kSkPanel;
}

// The row tile that holds panel x, 0 <= x < P (binary search over the
// closed-form prefix counts).
__host__ __device__ inline int sk_row(const SkGeo& g, int64_t x)
{
    int lo = 0, hi = g.nrt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (sk_before(g, mid) <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// First panel of workgroup b (b = G gives P).
__host__ __device__ inline int64_t sk_start(const SkGeo& g, int64_t b) { return b * g.P / g.G; }

// One piece of a tile: rows [r0, r0 + rows), columns from c0, k in
// [k_lo, k_hi); p0 > 0 marks a piece that starts inside its tile.
struct SkPiece {
    int r0, rows, c0, k_lo, k_hi;
    int64_t p0, next;  // panel offset inside the tile; the sequence position after the piece
};

// The piece that starts at panel x of the sequence and ends at panel xe at
// the latest (x < xe <= P).
__host__ __device__ inline SkPiece sk_piece(const SkGeo& g, int64_t x, int64_t xe)
{
    const int j = sk_row(g, x);
    const int64_t d = sk_depth(g, j), off = x - sk_before(g, j);
    const int64_t ct = off / d, p0 = off % d, tile_end = x - p0 + d;
    const int64_t x1 = xe < tile_end ? xe : tile_end;
    SkPiece s;
    s.r0 = j * g.tm;
    s.rows = (int)(g.n - s.r0 < g.tm ? g.n - s.r0 : g.tm);
    s.c0 = (int)(ct * g.tn);
    s.k_lo = (int)(p0 * kSkPanel);
    s.k_hi = x1 == tile_end ? s.r0 + s.rows : (int)((p0 + (x1 - x)) * kSkPanel);
    s.p0 = p0;
    s.next = x1;
    return s;
}

// The geometry of an n x m walk in tm x tn tiles over G workgroups (G is
// capped at P so that every workgroup has at least one panel).
inline SkGeo sk_geometry(int64_t n, int64_t m, int tm, int tn, int64_t G)
{
    SkGeo g;
    g.n = n;
    g.m = m;
    g.tm = tm;
    g.tn = tn;
    g.nrt = (int32_t)((n + tm - 1) / tm);
    g.nct = (int32_t)((m + tn - 1) / tn);
    g.G = 1;
    g.P = sk_before(g, g.nrt - 1) + (int64_t)g.nct * sk_depth(g, g.nrt - 1);
    g.G = G < 1 ? 1 : (G > g.P ? g.P : G);
    return g;
}

}  // namespace uvt
