// Host-only check of the serving arena and its admission rule (built and run
// by tests/test_serve_host.py; no GPU): a fake executor admits requests in
// FIFO order while a slot is free and one arena block holds the request,
// completes them in random order, and checks that blocks never overlap, the
// budget is never exceeded, requests are admitted in order, nothing
// deadlocks, and the arena coalesces back to one free block.
#include <stdio.h>
#include <stdlib.h>

#include <deque>
#include <map>
#include <random>
#include <vector>

#include "uvt_serve_arena.hpp"

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                   \
        }                                                               \
    } while (0)

static int basic()
{
    uvt::OffsetArena a(1000);  // rounds up to 1024
    CHECK(a.size() == 1024);
    const size_t x = a.alloc(100), y = a.alloc(300), z = a.alloc(256);
    CHECK(x == 0 && y == 256 && z == 768);
    CHECK(a.used() == 1024 && a.alloc(1) == SIZE_MAX);
    a.free(y, 300);
    CHECK(a.largest_free() == 512 && a.alloc(600) == SIZE_MAX);
    a.free(x, 100);  // merges with the hole after it
    CHECK(a.free_blocks() == 1 && a.largest_free() == 768);
    a.free(z, 256);
    CHECK(a.free_blocks() == 1 && a.largest_free() == 1024 && a.used() == 0);
    return 0;
}

struct Req {
    int id;
    size_t bytes;
};

static int fake_executor(unsigned seed, size_t budget, int slots, int count)
{
    std::mt19937_64 rng(seed);
    uvt::OffsetArena arena(budget);
    std::deque<Req> pending;
    for (int i = 0; i < count; ++i) {
        // mostly small, some close to the whole budget
        const double u = std::uniform_real_distribution<double>(0, 1)(rng);
        size_t b = u < 0.9 ? (size_t)(u * budget / 8) + 1 : (size_t)(u * budget * 0.95);
        b = b > arena.size() ? arena.size() : b;
        pending.push_back(Req{i, b});
    }
    std::map<size_t, std::pair<size_t, int>> live;  // offset -> (granule bytes, id)
    int next_admit = 0, done = 0;
    while (done < count) {
        // admission, FIFO
        while (!pending.empty() && (int)live.size() < slots) {
            const size_t off = arena.alloc(pending.front().bytes);
            if (off == SIZE_MAX) break;
            CHECK(pending.front().id == next_admit);
            ++next_admit;
            const size_t g = uvt::OffsetArena::granule(pending.front().bytes);
            // no overlap with the blocks in flight
            auto it = live.lower_bound(off);
            if (it != live.end()) CHECK(off + g <= it->first);
            if (it != live.begin()) {
                auto pr = std::prev(it);
                CHECK(pr->first + pr->second.first <= off);
            }
            CHECK(off + g <= arena.size());
            live[off] = {g, pending.front().id};
            pending.pop_front();
            CHECK(arena.used() <= arena.size());
        }
        // nothing in flight: the head must have been admitted (no deadlock)
        CHECK(!live.empty());
        // complete one request in flight, chosen at random
        auto it = live.begin();
        std::advance(it, std::uniform_int_distribution<size_t>(0, live.size() - 1)(rng));
        arena.free(it->first, it->second.first);
        live.erase(it);
        ++done;
    }
    CHECK(arena.used() == 0 && arena.free_blocks() == 1 && arena.largest_free() == arena.size());
    return 0;
}

int main()
{
    if (basic()) return 1;
    for (unsigned s = 0; s < 20; ++s)
        if (fake_executor(s, (size_t)1 << (20 + s % 10), 1 + (int)(s % 16), 3000)) return 1;
    printf("ok\n");
    return 0;
}
