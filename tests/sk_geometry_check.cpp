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
//   - all of it again with column groups (a group's members run its share of
//     the sequence in one column tile each), and the launch never exceeds the
//     workgroups asked for;
//   - the XCD mapping of k_trmm_sk_f64 is a permutation of the launch.
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

// k_sk_fixup's choice for workgroup w (group b, column tile cx): the partial
// tiles it adds (empty if none) and the tile they belong to.
static std::vector<int64_t> fixup(const SkGeo& g, int64_t w, int& j_out, int64_t& ct_out)
{
    std::vector<int64_t> adds;
    const int64_t b = w / g.ncg, cx = w % g.ncg;
    if (b == 0) return adds;
    const int64_t x = sk_start(g, b);
    const int j = sk_row(g, x);
    const int64_t d = sk_depth(g, j), off = x - sk_before(g, j), p0 = off % d;
    if (p0 == 0) return adds;
    const int64_t tile0 = x - p0;
    if (sk_start(g, b - 1) > tile0) return adds;
    int64_t cnt = 1;
    while (b + cnt < g.G && sk_start(g, b + cnt) < tile0 + d) ++cnt;
    for (int64_t q = 0; q < cnt; ++q) adds.push_back((b + q) * g.ncg + cx);
    j_out = j;
    ct_out = off / d + cx;
    return adds;
}

static void check(int64_t n, int64_t m, int tm, int tn, int64_t G0, bool grouped)
{
    const SkGeo g = sk_geometry(n, m, tm, tn, G0, grouped);
    const int64_t nct = (m + tn - 1) / tn;  // column tiles of C
    const int64_t tiles = (int64_t)g.nrt * nct;
    const bool gr = grouped && nct > 1 && G0 / nct >= kXcds;
    CHECK(g.ncg == (gr ? nct : 1) && g.nct * g.ncg == nct, "groups n=%ld m=%ld G=%ld", (long)n, (long)m, (long)G0);
    std::vector<int64_t> first(g.nrt + 1, 0);  // panel index of each row tile's first tile (all columns)
    int64_t P = 0, Pall = 0;
    for (int j = 0; j < g.nrt; ++j) {
        const int64_t khi = std::min<int64_t>(n, (int64_t)(j + 1) * tm);
        CHECK(sk_depth(g, j) == (khi + 15) / 16, "depth n=%ld j=%d", (long)n, j);
        CHECK(sk_before(g, j) == P, "prefix n=%ld m=%ld j=%d", (long)n, (long)m, j);
        first[j] = Pall;
        P += sk_depth(g, j) * g.nct;
        Pall += sk_depth(g, j) * nct;
    }
    CHECK(P == g.P, "P n=%ld m=%ld", (long)n, (long)m);
    const int64_t want_G = gr ? G0 / nct : G0;
    CHECK(g.G >= 1 && g.G <= g.P && (want_G < 1 || g.G == std::min(want_G, g.P)), "G");
    CHECK(G0 < 1 || sk_launch(g) <= G0, "launch %ld above %ld", (long)sk_launch(g), (long)G0);
    std::vector<int> cover(Pall, 0);
    std::vector<int> c_writes(tiles, 0);
    std::vector<std::vector<std::pair<int, int64_t>>> parts(tiles);  // (k_lo, partial tile slot)
    for (int64_t w = 0; w < sk_launch(g); ++w) {
        const int64_t b = w / g.ncg, cx = w % g.ncg;
        int64_t x = sk_start(g, b);
        const int64_t xe = sk_start(g, b + 1);
        CHECK(x < xe, "empty workgroup %ld", (long)w);
        bool first_piece = true;
        while (x < xe) {
            const SkPiece s = sk_piece(g, x, xe);
            CHECK(s.next > x && s.next <= xe, "next");
            const int j = s.r0 / tm;
            const int64_t ct = s.c0 / tn + cx;
            CHECK(s.r0 % tm == 0 && s.c0 % tn == 0 && j < g.nrt && ct < nct, "tile");
            CHECK(s.rows == std::min<int64_t>(tm, n - s.r0), "rows");
            CHECK(s.k_lo == s.p0 * 16 && s.k_lo < s.k_hi && s.k_hi <= s.r0 + s.rows, "k range");
            CHECK(s.k_hi == s.r0 + s.rows || s.k_hi % 16 == 0, "k end");
            CHECK(s.p0 == 0 || first_piece, "a later piece starts inside its tile (w=%ld)", (long)w);
            const int64_t base = first[j] + ct * sk_depth(g, j);
            for (int64_t p = s.k_lo / 16; p < (s.k_hi + 15) / 16; ++p) cover[base + p]++;
            const int64_t t = (int64_t)j * nct + ct;
            if (s.p0 > 0) parts[t].push_back({s.k_lo, w});
            else c_writes[t]++;
            x = s.next;
            first_piece = false;
        }
    }
    for (int64_t p = 0; p < Pall; ++p) CHECK(cover[p] == 1, "panel %ld covered %d times", (long)p, cover[p]);
    for (int64_t t = 0; t < tiles; ++t) CHECK(c_writes[t] == 1, "tile %ld: %d k=0 pieces", (long)t, c_writes[t]);
    std::vector<std::vector<int64_t>> added(tiles);
    for (int64_t w = 0; w < sk_launch(g); ++w) {
        int j = -1;
        int64_t ct = -1;
        const std::vector<int64_t> a = fixup(g, w, j, ct);
        if (a.empty()) continue;
        const int64_t t = (int64_t)j * nct + ct;
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
    {  // the kernels' hardware workgroup -> launch index mapping
        const int64_t nw = sk_launch(g);
        std::vector<int> seen(nw, 0);
        for (int64_t i = 0; i < nw; ++i) {
            const int64_t w = sk_xcd_order(i, nw);
            CHECK(w >= 0 && w < nw, "xcd map range");
            if (w >= 0 && w < nw) seen[w]++;
        }
        for (int64_t w = 0; w < nw; ++w) CHECK(seen[w] == 1, "xcd map");
        // a group's members, where a group lies inside one XCD's run, share it
        for (int64_t i = 0; i + kXcds < nw; ++i)
            CHECK(sk_xcd_order(i + kXcds, nw) == sk_xcd_order(i, nw) + 1, "xcd run");
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
                for (int64_t G : gs)
                    for (bool grouped : {false, true}) {
                        check(n, m, sh[0], sh[1], G, grouped);
                        ++cases;
                    }
    for (bool grouped : {false, true}) {
        check(32768, 8192, 128, 128, 512, grouped);  // the largest plans
        check(65536, 1024, 128, 128, 512, grouped);
        check(6772, 1548, 128, 128, 512, grouped);  // 13 column tiles: 39 groups of 13
    }
    if (fails) {
        fprintf(stderr, "%d failures in %d cases\n", fails, cases);
        return 1;
    }
    printf("ok\n");
    return 0;
}
