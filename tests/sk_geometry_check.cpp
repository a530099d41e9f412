// Host-only check of the stream-K geometry of the trapezoid walk (built and
// run by tests/test_sk_host.py; no GPU), over tile shapes, ragged and tiny
// sizes and workgroup counts from 1 to more than the panels:
//   - the workgroups' pieces cover every k panel of every tile exactly once,
//     each piece inside its tile's k range [0, min(n, r0 + rows)) and on
//     16-deep panel boundaries except at the tile's end;
//   - a piece that starts inside its tile is the first piece of its
//     workgroup, and every tile has exactly one piece that starts at k = 0
//     (the one that writes C);
//   - the fixup rule of k_sk_fixup (run here on the host) adds each partial
//     tile to the right tile exactly once, in k order, and touches no tile
//     without partial tiles;
//   - the XCD mapping of k_trmm_sk_f64 is a permutation of the workgroups.
#include <stdio.h>

#include <algorithm>
#include <vector>

#define __host__
#define __device__
#include "uvt_sk.hpp"

using namespace uvt;

static int fails = 0;
#define CHECK(c, ...)                                          \
    do {                                                       \
        if (!(c)) {                                            \
            if (fails++ < 20) {                                \
                fprintf(stderr, "line %d: %s: ", __LINE__, #c); \
                fprintf(stderr, __VA_ARGS__);                  \
                fprintf(stderr, "\n");                         \
            }                                                  \
        }                                                      \
    } while (0)

// k_sk_fixup's choice for workgroup b: the partial tiles it adds (empty if
// none) and the tile they belong to.
static std::vector<int64_t> fixup(const SkGeo& g, int64_t b, int& j_out, int64_t& ct_out)
{
    std::vector<int64_t> adds;
    if (b == 0) return adds;
    const int64_t x = sk_start(g, b);
    const int j = sk_row(g, x);
    const int64_t d = sk_depth(g, j), off = x - sk_before(g, j), p0 = off % d;
    if (p0 == 0) return adds;
    const int64_t tile0 = x - p0;
    if (sk_start(g, b - 1) > tile0) return adds;
    int64_t cnt = 1;
    while (b + cnt < g.G && sk_start(g, b + cnt) < tile0 + d) ++cnt;
    for (int64_t q = 0; q < cnt; ++q) adds.push_back(b + q);
    j_out = j;
    ct_out = off / d;
    return adds;
}

static void check(int64_t n, int64_t m, int tm, int tn, int64_t G0)
{
    const SkGeo g = sk_geometry(n, m, tm, tn, G0);
    const int64_t tiles = (int64_t)g.nrt * g.nct;
    std::vector<int64_t> first(g.nrt + 1, 0);  // panel index of each row tile's first tile
    int64_t P = 0;
    for (int j = 0; j < g.nrt; ++j) {
        const int64_t khi = std::min<int64_t>(n, (int64_t)(j + 1) * tm);
        CHECK(sk_depth(g, j) == (khi + 15) / 16, "depth n=%ld j=%d", (long)n, j);
        CHECK(sk_before(g, j) == P, "prefix n=%ld m=%ld j=%d", (long)n, (long)m, j);
        first[j] = P;
        P += sk_depth(g, j) * g.nct;
    }
    CHECK(P == g.P, "P n=%ld m=%ld", (long)n, (long)m);
    CHECK(g.G >= 1 && g.G <= g.P && (G0 < 1 || g.G == std::min(G0, g.P)), "G");
    std::vector<int> cover(P, 0);
    std::vector<int> c_writes(tiles, 0);
    std::vector<std::vector<std::pair<int, int64_t>>> parts(tiles);  // (k_lo, workgroup)
    for (int64_t b = 0; b < g.G; ++b) {
        int64_t x = sk_start(g, b);
        const int64_t xe = sk_start(g, b + 1);
        CHECK(x < xe, "empty workgroup %ld", (long)b);
        bool first_piece = true;
        while (x < xe) {
            const SkPiece s = sk_piece(g, x, xe);
            CHECK(s.next > x && s.next <= xe, "next");
            const int j = s.r0 / tm;
            const int64_t ct = s.c0 / tn;
            CHECK(s.r0 % tm == 0 && s.c0 % tn == 0 && j < g.nrt && ct < g.nct, "tile");
            CHECK(s.rows == std::min<int64_t>(tm, n - s.r0), "rows");
            CHECK(s.k_lo == s.p0 * 16 && s.k_lo < s.k_hi && s.k_hi <= s.r0 + s.rows, "k range");
            CHECK(s.k_hi == s.r0 + s.rows || s.k_hi % 16 == 0, "k end");
            CHECK(s.p0 == 0 || first_piece, "a later piece starts inside its tile (b=%ld)", (long)b);
            const int64_t base = first[j] + ct * sk_depth(g, j);
            for (int64_t p = s.k_lo / 16; p < (s.k_hi + 15) / 16; ++p) cover[base + p]++;
            const int64_t t = (int64_t)j * g.nct + ct;
            if (s.p0 > 0) parts[t].push_back({s.k_lo, b});
            else c_writes[t]++;
            x = s.next;
            first_piece = false;
        }
    }
    for (int64_t p = 0; p < P; ++p) CHECK(cover[p] == 1, "panel %ld covered %d times", (long)p, cover[p]);
    for (int64_t t = 0; t < tiles; ++t) CHECK(c_writes[t] == 1, "tile %ld: %d k=0 pieces", (long)t, c_writes[t]);
    std::vector<std::vector<int64_t>> added(tiles);
    for (int64_t b = 0; b < g.G; ++b) {
        int j = -1;
        int64_t ct = -1;
        const std::vector<int64_t> a = fixup(g, b, j, ct);
        if (a.empty()) continue;
        const int64_t t = (int64_t)j * g.nct + ct;
        CHECK(added[t].empty(), "tile %ld fixed up twice", (long)t);
        added[t] = a;
    }
    for (int64_t t = 0; t < tiles; ++t) {
        std::vector<int64_t> want;
        int last = -1;
        for (const auto& pr : parts[t]) {
            CHECK(pr.first > last, "k order");
            last = pr.first;
            want.push_back(pr.second);
        }
        CHECK(added[t] == want, "fixup of tile %ld (n=%ld m=%ld G=%ld)", (long)t, (long)n, (long)m, (long)g.G);
    }
    if (g.G % kXcds == 0) {  // the kernel's workgroup -> sequence mapping
        std::vector<int> seen(g.G, 0);
        for (int64_t i = 0; i < g.G; ++i) seen[(i % kXcds) * (g.G / kXcds) + i / kXcds]++;
        for (int64_t b = 0; b < g.G; ++b) CHECK(seen[b] == 1, "xcd map");
    }
}

int main()
{
    const int shapes[][2] = {{128, 128}, {256, 128}, {64, 64}, {128, 32}};
    const int64_t ns[] = {1, 15, 16, 17, 64, 127, 128, 129, 300, 1000, 2048, 5000, 8165, 16384};
    const int64_t ms[] = {1, 31, 32, 33, 128, 200, 939, 2000};
    const int64_t gs[] = {1, 2, 7, 512, 768, 2048, 1 << 30};
    int cases = 0;
    for (const auto& sh : shapes)
        for (int64_t n : ns)
            for (int64_t m : ms)
                for (int64_t G : gs) {
                    check(n, m, sh[0], sh[1], G);
                    ++cases;
                }
    check(32768, 8192, 128, 128, 512);  // the largest plans
    check(65536, 1024, 128, 128, 512);
    if (fails) {
        fprintf(stderr, "%d failures in %d cases\n", fails, cases);
        return 1;
    }
    printf("ok\n");
    return 0;
}
