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
// Column groups (ncg > 1): for narrow B the nct column tiles of a row tile
// share every A panel, so instead of laying them end to end the sequence
// holds each row tile's panels once (nct = 1 in the arithmetic below) and a
// launch of G groups x ncg workgroups runs it: member c of group b runs group
// b's share of the sequence in column tile c.  The ncg members of a group
// read the same A panels at about the same time, from one XCD's L2.
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
    int64_t P;     // panels of the whole walk (of one column tile when ncg > 1)
    int32_t nrt, nct, tm, tn;
    int64_t G;     // workgroups (groups when ncg > 1), 1 <= G <= P
    int32_t ncg;   // column tiles per group: 1 (none), or every column tile
};

// Workgroups of the launch.
__host__ __device__ inline int64_t sk_launch(const SkGeo& g) { return g.G * g.ncg; }

// The launch index w that hardware workgroup i runs (workgroup i of a launch
// runs on XCD i % 8): XCD x takes a contiguous run of w, so consecutive w --
// the members of a group, and neighbouring shares of the sequence -- share
// an XCD's L2.  A permutation of [0, nw).
__host__ __device__ inline int64_t sk_xcd_order(int64_t i, int64_t nw)
{
    const int64_t x = i % kXcds, q = nw / kXcds, r = nw % kXcds;
    return x * q + (x < r ? x : r) + i / kXcds;
}

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

// The geometry of an n x m walk in tm x tn tiles over at most W workgroups
// (every workgroup gets at least one panel).  group_cols: column groups
// where there are at least kXcds of them (G = W / nct groups of nct), else
// none.
inline SkGeo sk_geometry(int64_t n, int64_t m, int tm, int tn, int64_t W, bool group_cols = false)
{
    SkGeo g;
    g.n = n;
    g.m = m;
    g.tm = tm;
    g.tn = tn;
    g.nrt = (int32_t)((n + tm - 1) / tm);
    const int32_t nct = (int32_t)((m + tn - 1) / tn);
    const bool grouped = group_cols && nct > 1 && W / nct >= kXcds;
    g.nct = grouped ? 1 : nct;
    g.ncg = grouped ? nct : 1;
    g.G = 1;
    g.P = sk_before(g, g.nrt - 1) + (int64_t)g.nct * sk_depth(g, g.nrt - 1);
    const int64_t G = grouped ? W / nct : W;
    g.G = G < 1 ? 1 : (G > g.P ? g.P : G);
    return g;
}

}  // namespace uvt
