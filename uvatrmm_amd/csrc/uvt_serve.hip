// uvt_serve.hip -- a TRMM server: requests of mixed sizes and tolerance
// classes arrive over time, are admitted by HBM budget and free slots, run
// concurrently on per-slot streams, and are timed from arrival to result.
// No counterpart in the reference (host code only; the kernels are those of
// uvt_api.hip and uvt_split.hip).
//
// One scheduler thread owns every HIP call of the server:
//   loop: collect finished slots (event query) -> free their arena blocks ->
//         admit pending requests in FIFO order while a slot is free and the
//         next request's operands fit in the arena (a large request held back
//         by the large-request cap lets later ones pass) -> launch them
//         (inputs, product, optional copy-back) on the slot's stream -> sleep
//         until a submission or ~50 us.
// A request's kernels depend only on its own (n, m, class), so its result is
// the same alone and in the mix.  Besides the arena, each slot holds a fixed
// buffer for the partial tiles of a stream-K FP64 product (64 MiB on a GPU of
// 256 CUs), reused by the slot's next request once the last one is done.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "uvt_internal.h"
#include "uvt_runtime.hpp"
#include "uvt_serve_arena.hpp"
#include "uvt_split.hpp"

namespace {

size_t align_up(size_t x) { return uvt::OffsetArena::granule(x); }

// The serving arena: one device allocation of the budget, carved by offsets.
class Arena {
public:
    bool init(size_t bytes)
    {
        offsets_.reset(bytes);
        if (hipMalloc(&base_, offsets_.size()) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            return false;
        }
        return true;
    }
    void release()
    {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
        offsets_.reset(0);
    }
    size_t alloc(size_t bytes) { return offsets_.alloc(bytes); }
    void free(size_t off, size_t bytes) { offsets_.free(off, bytes); }
    char* at(size_t off) const { return (char*)base_ + off; }
    size_t used() const { return offsets_.used(); }
    size_t size() const { return offsets_.size(); }

private:
    void* base_ = nullptr;
    uvt::OffsetArena offsets_;
};

struct Pending {
    uint64_t ticket;
    uvt_request r;
    double t_submit;
    size_t bytes;
};

struct Slot {
    bool busy = false;
    bool big = false;  // counts against the large-request cap
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Pending p{};
    double t_admit = 0.0;
    double* dC = nullptr;
    size_t off = SIZE_MAX;  // the request's arena block (p.bytes long)
    uvt_status status = UVT_OK;
    const char* err = nullptr;  // int8 path: the error statement in the workspace
    double* parts = nullptr;    // partial tiles of a stream-K FP64 product (the slot's own)
};


}  // namespace

struct uvt_server_s {
    uvt_handle h = nullptr;
    uvt_serve_config cfg{};
    std::vector<uvt_tune_entry> tune;  // empty: built-in tile-shape choice
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_drain;
    std::deque<Pending> pending;
    std::vector<uvt_completion> done;
    std::vector<Slot> slots;
    Arena arena;
    bool stop = false;
    uint64_t next_ticket = 1;
    int inflight = 0;
    int big_inflight = 0;
    int hwq = 4;  // hardware queues HIP multiplexes the slot streams onto
    // statistics
    int64_t submitted = 0, completed = 0, rejected = 0;
    std::vector<double> latencies;
    double first_submit = -1.0, last_done = 0.0, flops = 0.0;
    size_t peak_bytes = 0;
    int peak_inflight = 0;
    std::chrono::steady_clock::time_point t0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Whether a request's box part runs on the int8 path: split class, measured
// faster alone at this size, and a size the int8 path supports.
static bool runs_int8(const uvt_serve_config& cfg, int64_t n, int64_t m, int32_t prec)
{
    return prec == UVT_PREC_SPLIT && uvt_split_faster_alone(n, m) && uvt_split_supported(n, cfg.slices);
}

static size_t request_bytes(const uvt_serve_config& cfg, int64_t n, int64_t m, int32_t prec,
                            bool ab_in_arena, bool c_in_arena)
{
    size_t b = 0;
    if (ab_in_arena) b += align_up((size_t)n * n * 8) + align_up((size_t)n * m * 8);
    if (c_in_arena) b += align_up((size_t)n * m * 8);
    if (runs_int8(cfg, n, m, prec)) {
        size_t w = 0;
        if (uvt_trmm_split_workspace(n, m, cfg.slices, &w) == UVT_OK) b += align_up(w);
    }
    return b;
}

static bool ab_in_arena(const uvt_request& r) { return r.seeded || !(r.flags & UVT_REQ_AB_DEVICE); }
static bool c_in_arena(const uvt_request& r) { return !(r.C && (r.flags & UVT_REQ_C_DEVICE)); }

extern "C" uvt_status uvt_serve_request_bytes(const uvt_serve_config* cfg, int64_t n, int64_t m,
                                              int32_t prec, size_t* bytes)
{
    if (!cfg || !bytes || n <= 0 || m <= 0) return UVT_ERR_ARG;
    *bytes = request_bytes(*cfg, n, m, prec, true, true);
    return UVT_OK;
}

// Enqueues one admitted request on its slot (scheduler thread, lock not
// held).  Its operands are carved, in request_bytes order, out of the one
// arena block admission reserved for it.
static uvt_status launch(uvt_server s, Slot& sl)
{
    const uvt_request& r = sl.p.r;
    const int64_t n = r.n, m = r.m;
    char* cursor = s->arena.at(sl.off);
    auto carve = [&](size_t bytes) -> char* {
        char* p = cursor;
        cursor += align_up(bytes);
        return p;
    };
    const bool own_ab = ab_in_arena(r);
    const double* dA = own_ab ? (const double*)carve((size_t)n * n * 8) : r.A;
    const double* dB = own_ab ? (const double*)carve((size_t)n * m * 8) : r.B;
    double* dC = c_in_arena(r) ? (double*)carve((size_t)n * m * 8) : r.C;
    const bool int8 = runs_int8(s->cfg, n, m, r.prec);
    size_t wbytes = 0;
    void* work = nullptr;
    sl.err = nullptr;
    if (int8) {
        if (uvt_trmm_split_workspace(n, m, s->cfg.slices, &wbytes) != UVT_OK) return UVT_ERR_ARG;
        work = carve(wbytes);
        sl.err = (const char*)work + uvt_split_err_offset(n, m, s->cfg.slices);
    }
    if (cursor > s->arena.at(sl.off) + sl.p.bytes) return UVT_ERR_NOMEM;  // cannot happen
    sl.dC = dC;
    hipStream_t st = sl.stream;
    uvt_status e = UVT_OK;
    if (r.seeded) {
        e = uvt_fill_seeded(s->h, n, m, r.seed, r.spread, (double*)dA, n, (double*)dB, m, st);
    } else if (own_ab) {
        UVT_HIP(hipMemcpyAsync((void*)dA, r.A, (size_t)n * n * 8, hipMemcpyHostToDevice, st));
        UVT_HIP(hipMemcpyAsync((void*)dB, r.B, (size_t)n * m * 8, hipMemcpyHostToDevice, st));
    }
    if (e != UVT_OK) return e;
    if (int8) {
        e = uvt_trmm_split(s->h, n, m, s->cfg.slices, dA, n, dB, m, dC, m, work, wbytes, st);
    } else {
        const int tile = uvt_wide_tiles(s->tune.empty() ? uvt_default_cfg(n, m) : uvt_tune_pick(s->tune, n, m), n, m);
        const int64_t sk = uvt_sk_workgroups(tile, n, m, s->h->cus);
        e = sk > 0 ? uvt_launch_trapezoid_sk_f64(tile, n, m, dA, n, dB, m, dC, m, sk, sl.parts, st)
                   : uvt_launch_trapezoid_f64(tile, n, m, dA, n, dB, m, dC, m, st);
    }
    if (e != UVT_OK) return e;
    UVT_HIP(hipEventRecord(sl.done, st));
    return UVT_OK;
}

static bool is_big(const uvt_serve_config& cfg, const uvt_request& r)
{
    return cfg.max_big > 0 && (double)r.m * (double)r.n * (double)r.n >= cfg.big_flops;
}

// The large request to admit next when they go shortest first: the oldest
// one that has waited big_wait_ms or longer, else the one of fewest flops
// (the earliest among equals).  Some large request is pending.
static std::deque<Pending>::iterator next_big(uvt_server s)
{
    const double now = s->now(), wait = s->cfg.big_wait_ms * 1e-3;
    auto best = s->pending.end();
    double best_flops = 0.0;
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it) {
        if (!is_big(s->cfg, it->r)) continue;
        if (now - it->t_submit >= wait) return it;  // the queue is in arrival order
        const double f = (double)it->r.m * (double)it->r.n * (double)it->r.n;
        if (best == s->pending.end() || f < best_flops) {
            best = it;
            best_flops = f;
        }
    }
    return best;
}

// A free slot on the least busy hardware queue.  HIP spreads streams over
// its hardware queues in turn (GPU_MAX_HW_QUEUES of them, 4 by default), so
// slot i shares a queue with slots i +- hwq; a kernel queued behind another
// request's on the same queue waits for all of it.  Ties go to the lowest
// slot.
static Slot* pick_slot(uvt_server s)
{
    const int ns = (int)s->slots.size();
    Slot* best = nullptr;
    int best_busy = 0;
    for (int i = 0; i < ns; ++i) {
        if (s->slots[i].busy) continue;
        int busy = 0;
        for (int j = i % s->hwq; j < ns; j += s->hwq) busy += s->slots[j].busy ? 1 : 0;
        if (!best || busy < best_busy) {
            best = &s->slots[i];
            best_busy = busy;
        }
    }
    return best;
}

static void scheduler(uvt_server s)
{
    (void)hipSetDevice(s->h->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progress = false;
        // 1. finished slots
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = sl.status == UVT_OK ? hipEventQuery(sl.done) : hipSuccess;
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                (void)hipGetLastError();
                sl.status = UVT_ERR_HIP;
            }
            if (sl.status == UVT_OK && sl.p.r.C && c_in_arena(sl.p.r)) {
                lk.unlock();
                const hipError_t c = hipMemcpy(sl.p.r.C, sl.dC, (size_t)sl.p.r.n * sl.p.r.m * 8,
                                               hipMemcpyDeviceToHost);
                lk.lock();
                if (c != hipSuccess) {
                    (void)hipGetLastError();
                    sl.status = UVT_ERR_HIP;
                }
            }
            // the int8 path's error statement, read before its workspace goes
            double err[2] = {-1.0, -1.0};
            if (sl.status == UVT_OK && sl.err) {
                lk.unlock();
                const hipError_t c = hipMemcpy(err, sl.err, sizeof err, hipMemcpyDeviceToHost);
                lk.lock();
                if (c != hipSuccess) {
                    (void)hipGetLastError();
                    sl.status = UVT_ERR_HIP;
                }
            }
            const double t = s->now();
            s->arena.free(sl.off, sl.p.bytes);
            sl.off = SIZE_MAX;
            uvt_completion c;
            memset(&c, 0, sizeof c);
            c.ticket = sl.p.ticket;
            c.tag = sl.p.r.tag;
            c.status = sl.status;
            c.path = sl.err ? 1 : 0;
            c.t_submit = sl.p.t_submit;
            c.t_admit = sl.t_admit;
            c.t_done = t;
            c.err_bound = err[0];
            c.err_estimate = err[1];
            s->done.push_back(c);
            s->completed++;
            if (sl.status == UVT_OK) {
                s->latencies.push_back(t - sl.p.t_submit);
                s->flops += (double)sl.p.r.m * (double)sl.p.r.n * (double)sl.p.r.n;
            }
            s->last_done = std::max(s->last_done, t);
            sl.busy = false;
            if (sl.big) s->big_inflight--;
            sl.big = false;
            s->inflight--;
            progress = true;
        }
        if (progress) s->cv_drain.notify_all();
        // 2. admission: FIFO, except that a large request held back by the
        // cap lets later requests pass (large ones keep their order, or go
        // shortest first with big_wait_ms); a request that does not fit in the
        // arena stops admission, so memory is granted in queue order
        for (;;) {
            auto it = s->pending.begin();
            while (it != s->pending.end() && is_big(s->cfg, it->r) && s->big_inflight >= s->cfg.max_big)
                ++it;
            if (it == s->pending.end()) break;
            if (is_big(s->cfg, it->r) && s->cfg.big_wait_ms > 0) it = next_big(s);
            Slot* free_slot = pick_slot(s);
            if (!free_slot) break;
            // one contiguous block per request; none free yet: wait for
            // completions (with nothing in flight the arena is one free block,
            // and every queued request fits the budget)
            const size_t off = s->arena.alloc(it->bytes);
            if (off == SIZE_MAX) break;
            free_slot->busy = true;
            free_slot->big = is_big(s->cfg, it->r);
            free_slot->off = off;
            free_slot->p = *it;
            free_slot->t_admit = s->now();
            free_slot->status = UVT_OK;
            s->pending.erase(it);  // before unlocking: submissions may grow the queue meanwhile
            s->inflight++;
            if (free_slot->big) s->big_inflight++;
            s->peak_inflight = std::max(s->peak_inflight, s->inflight);
            // the arena and the slot belong to this thread: launch unlocked, so
            // submit and poll never wait behind copies or kernel launches
            lk.unlock();
            const uvt_status ls = launch(s, *free_slot);
            // a failed launch may have enqueued part of its work: let it finish
            // before the blocks go back to the arena
            if (ls != UVT_OK) (void)hipStreamSynchronize(free_slot->stream);
            lk.lock();
            free_slot->status = ls;
            s->peak_bytes = std::max(s->peak_bytes, s->arena.used());
            progress = true;
        }
        if (s->stop && s->pending.empty() && s->inflight == 0) break;
        if (!progress) {
            if (s->inflight > 0) s->cv_work.wait_for(lk, std::chrono::microseconds(50));
            else s->cv_work.wait(lk);
        }
    }
}

extern "C" uvt_status uvt_serve_create(uvt_handle h, const uvt_serve_config* cfg, uvt_server* out)
{
    if (!h || !cfg || !out || cfg->max_slots < 1 || cfg->max_slots > 64 || cfg->slices < 2 ||
        cfg->slices > 14 || cfg->hbm_budget < 4096 || !(cfg->big_flops >= 0.0) || cfg->max_big < 0 ||
        cfg->big_wait_ms < 0)
        return UVT_ERR_ARG;
    std::vector<uvt_tune_entry> tune;
    if (cfg->tune_table && !uvt_tune_load(cfg->tune_table, tune)) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    uvt_server s = new (std::nothrow) uvt_server_s;
    if (!s) return UVT_ERR_NOMEM;
    s->h = h;
    s->cfg = *cfg;
    s->cfg.tune_table = nullptr;  // the caller's string need not outlive create
    s->tune = tune;
    s->t0 = std::chrono::steady_clock::now();
    if (const char* q = getenv("GPU_MAX_HW_QUEUES")) s->hwq = std::max(1, atoi(q));
    if (!s->arena.init(cfg->hbm_budget)) {
        delete s;
        return UVT_ERR_NOMEM;
    }
    s->slots.resize(cfg->max_slots);
    const size_t parts_bytes = uvt_sk_max_parts_bytes(h->cus);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void**)&sl.parts, parts_bytes) != hipSuccess) {
            (void)hipGetLastError();
            for (Slot& x : s->slots) {
                if (x.stream) (void)hipStreamDestroy(x.stream);
                if (x.done) (void)hipEventDestroy(x.done);
                if (x.parts) (void)hipFree(x.parts);
            }
            s->arena.release();
            delete s;
            return UVT_ERR_HIP;
        }
    }
    s->th = std::thread(scheduler, s);
    *out = s;
    return UVT_OK;
}

extern "C" uvt_status uvt_serve_submit(uvt_server s, const uvt_request* r, uint64_t* ticket)
{
    if (!s || !r || r->n <= 0 || r->m <= 0) return UVT_ERR_ARG;
    if (r->prec != UVT_PREC_F64 && r->prec != UVT_PREC_SPLIT) return UVT_ERR_ARG;
    if (!r->seeded && (!r->A || !r->B)) return UVT_ERR_ARG;
    if (r->flags & ~(UVT_REQ_AB_DEVICE | UVT_REQ_C_DEVICE)) return UVT_ERR_ARG;
    if ((r->flags & UVT_REQ_C_DEVICE) && !r->C) return UVT_ERR_ARG;
    if (r->n > 65536 || r->m > 0x3FFFFFFF) return UVT_ERR_SIZE;
    const size_t bytes = request_bytes(s->cfg, r->n, r->m, r->prec, ab_in_arena(*r), c_in_arena(*r));
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return UVT_ERR_STATE;
    if (bytes > s->arena.size()) {
        s->rejected++;
        return UVT_ERR_BUDGET;
    }
    Pending p;
    p.ticket = s->next_ticket++;
    p.r = *r;
    p.t_submit = s->now();
    p.bytes = bytes;
    if (s->first_submit < 0) s->first_submit = p.t_submit;
    s->pending.push_back(p);
    s->submitted++;
    if (ticket) *ticket = p.ticket;
    s->cv_work.notify_one();
    return UVT_OK;
}

extern "C" uvt_status uvt_serve_poll(uvt_server s, uvt_completion* out, int32_t cap, int32_t* count)
{
    if (!s || !count || (cap > 0 && !out)) return UVT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    const int32_t k = (int32_t)std::min<size_t>(cap > 0 ? cap : 0, s->done.size());
    for (int32_t i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *count = k;
    return UVT_OK;
}

extern "C" uvt_status uvt_serve_drain(uvt_server s)
{
    if (!s) return UVT_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_drain.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
    return UVT_OK;
}

extern "C" uvt_status uvt_serve_stats_get(uvt_server s, uvt_serve_stats* st)
{
    if (!s || !st) return UVT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof *st);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->rejected = s->rejected;
    std::vector<double> v = s->latencies;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto rank = [&](double q) {  // nearest-rank percentile
            size_t k = (size_t)std::ceil(q * (double)v.size());
            return v[std::min(v.size(), std::max<size_t>(k, 1)) - 1];
        };
        st->p50_s = rank(0.50);
        st->p99_s = rank(0.99);
        double sum = 0.0;
        for (double x : v) sum += x;
        st->mean_s = sum / (double)v.size();
        st->max_s = v.back();
    }
    st->wall_s = s->first_submit >= 0 ? s->last_done - s->first_submit : 0.0;
    st->flops = s->flops;
    st->peak_bytes = (double)s->peak_bytes;
    st->peak_inflight = s->peak_inflight;
    return UVT_OK;
}

extern "C" uvt_status uvt_serve_destroy(uvt_server s)
{
    if (!s) return UVT_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(s->h->device);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.parts) (void)hipFree(sl.parts);
    }
    s->arena.release();
    delete s;
    return UVT_OK;
}
