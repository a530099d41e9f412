// uvt_batch.hip -- many uneven TRMMs in one hipGraph launch, and the
// autotuner that picks the FP64 tile shape per problem size.  No counterpart
// in the reference, whose ranks each run one N x N product.
//
// A batch is captured once into a graph:
//   zero the work counters (a fill kernel, not a memset node)
//   -> scales and residues of the split problems
//   -> fork: main branch: one persistent kernel draining one queue of the
//                         FP64 tiles of every problem, each tile with its
//                         problem's tile shape, most work first
//            side branch: the int8 residue products of the split problems
//   -> join -> the split reconstruction, which adds the integer box part to
//              the diagonal blocks the FP64 queue stored.
// The FP64 queue holds every tile of the FP64 problems and the diagonal-block
// tiles of the split problems.  Split-class problems run wholly in FP64,
// which meets their tolerance, unless their route asks for the int8 path or
// (route auto) the int8 path is measured faster in a batch
// (uvt_split_faster_in_batch); sizes the int8 path does not support always
// run in FP64.  The split stages read every problem from a device
// descriptor table, so uvt_batch_update reaches them under the captured
// graph.
// Every tile and every split stage works on one problem with a fixed
// k-order, so a problem's result is the same alone or in any batch.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "uvt_f64_tile.hpp"
#include "uvt_internal.h"
#include "uvt_runtime.hpp"
#include "uvt_split.hpp"

using namespace uvt;

namespace {

struct BatchTile {
    int32_t prob, r0, rows, c0, k_lo, k_hi, cfg;  // cfg: UVT_CFG_* tile shape
};

struct BatchProb {
    const double* A;
    const double* B;
    double* C;
    int64_t lda, ldb, ldc, m;
    int32_t vec, pad;
};

// Persistent FP64 tile kernel for every tile shape: every workgroup (8
// waves, LDS for the largest shape) takes the next tile of the queue (an
// atomic counter) and runs it with that tile's shape, until the queue is
// empty, so all waves exit once `ntiles` tiles have been handed out.  One
// queue for all shapes means one drain at the end of the batch instead of
// one per shape.  128x128 tiles run as 4 x 2 waves of 32x64, 64x64 tiles as
// 2 x 4 waves of 32x16, 128x32 tiles as 4 x 2 waves of 32x16 (the B panel
// has half as many pieces as threads).  The order of every sum does not
// depend on the wave layout, so a tile gives the same bits as the
// single-problem kernels' tiles of the same shape.
constexpr int kBatchThreads = 512;
template <bool VEC>
__global__ __launch_bounds__(kBatchThreads, 4)
void k_batch_f64(const BatchTile* __restrict__ tiles, int ntiles, const BatchProb* __restrict__ probs,
                 int* __restrict__ counter)
{
    __shared__ __attribute__((aligned(16))) double smem[F64Cfg<128, 128, 4, 2>::SMEM_DOUBLES];
    static_assert(F64Cfg<64, 64, 2, 4>::SMEM_DOUBLES <= F64Cfg<128, 128, 4, 2>::SMEM_DOUBLES &&
                      F64Cfg<128, 64, 4, 2>::SMEM_DOUBLES <= F64Cfg<128, 128, 4, 2>::SMEM_DOUBLES &&
                      F64Cfg<128, 32, 4, 2>::SMEM_DOUBLES <= F64Cfg<128, 128, 4, 2>::SMEM_DOUBLES,
                  "LDS sized for the largest shape");
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        __syncthreads();  // every thread has read `next` before it is overwritten
        if (t >= ntiles) break;
        const BatchTile bt = tiles[t];
        const BatchProb p = probs[bt.prob];
        // p.pad is always 0, but loaded per tile: the thread index built from
        // it keeps each shape's offset arithmetic inside its own branch
        const int tid = (int)threadIdx.x + p.pad;
        if (bt.cfg == UVT_CFG_64x64)
            f64_tile<64, 64, VEC, 2, 4, 16, true, true>(bt.r0, bt.rows, bt.c0, bt.k_lo, bt.k_hi, kTileLower, gbl(p.A),
                                        p.lda, gbl(p.B), p.ldb, gbl(p.C), p.ldc, p.m, smem, tid);
        else if (bt.cfg == UVT_CFG_128x64)
            f64_tile<128, 64, VEC, 4, 2, 16, true, true>(bt.r0, bt.rows, bt.c0, bt.k_lo, bt.k_hi, kTileLower, gbl(p.A),
                                         p.lda, gbl(p.B), p.ldb, gbl(p.C), p.ldc, p.m, smem, tid);
        else if (bt.cfg == UVT_CFG_128x32)
            f64_tile<128, 32, VEC, 4, 2, 16, true, true>(bt.r0, bt.rows, bt.c0, bt.k_lo, bt.k_hi, kTileLower, gbl(p.A),
                                         p.lda, gbl(p.B), p.ldb, gbl(p.C), p.ldc, p.m, smem, tid);
        else
            f64_tile<128, 128, VEC, 4, 2, 16, true, true>(bt.r0, bt.rows, bt.c0, bt.k_lo, bt.k_hi, kTileLower, gbl(p.A),
                                          p.lda, gbl(p.B), p.ldb, gbl(p.C), p.ldc, p.m, smem, tid);
        __syncthreads();  // the LDS tile is reused by the next tile
    }
}

__global__ void k_zero_ints(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Workgroups of the persistent kernel: every CU filled to its occupancy.
int persistent_grid(int cus)
{
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_batch_f64<true>, kBatchThreads, 0) != hipSuccess ||
        b < 1) {
        (void)hipGetLastError();
        b = 1;
    }
    return cus * b;
}

void launch_persistent(bool vec, int grid, const BatchTile* tiles, int ntiles, const BatchProb* probs,
                       int* counter, hipStream_t s)
{
    const dim3 g((unsigned)std::max(1, std::min(grid, ntiles))), block(kBatchThreads);
    if (vec) hipLaunchKernelGGL(k_batch_f64<true>, g, block, 0, s, tiles, ntiles, probs, counter);
    else hipLaunchKernelGGL(k_batch_f64<false>, g, block, 0, s, tiles, ntiles, probs, counter);
}

// ---- tuning table ------------------------------------------------------------

const char* cfg_name(int cfg)
{
    return cfg == UVT_CFG_64x64 ? "64x64" : cfg == UVT_CFG_128x32 ? "128x32" : "128x128";
}

int cfg_from_name(const std::string& s)
{
    if (s == "64x64") return UVT_CFG_64x64;
    if (s == "128x32") return UVT_CFG_128x32;
    if (s == "128x128") return UVT_CFG_128x128;
    return -1;
}

}  // namespace

// Reads the table written by uvt_autotune_f64: a JSON object whose "table"
// array holds {"n_max": N, "m_max": M, "cfg": "AxB", ...} entries.  Only
// those three keys are read; entries are matched in file order.
bool uvt_tune_load(const char* path, std::vector<uvt_tune_entry>& out)
{
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    std::string txt;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, got);
    fclose(f);
    size_t pos = txt.find("\"table\"");
    if (pos == std::string::npos) return false;
    out.clear();
    for (;;) {
        const size_t obj = txt.find('{', pos);
        if (obj == std::string::npos) break;
        const size_t end = txt.find('}', obj);
        if (end == std::string::npos) return false;
        const std::string e = txt.substr(obj, end - obj);
        auto num = [&](const char* key, int64_t* v) {
            const size_t k = e.find(key);
            if (k == std::string::npos) return false;
            const size_t c = e.find(':', k);
            if (c == std::string::npos) return false;
            *v = strtoll(e.c_str() + c + 1, nullptr, 10);
            return true;
        };
        uvt_tune_entry t;
        const size_t k = e.find("\"cfg\"");
        if (!num("\"n_max\"", &t.n_max) || !num("\"m_max\"", &t.m_max) || k == std::string::npos)
            return false;
        const size_t q1 = e.find('"', e.find(':', k) + 1);
        const size_t q2 = e.find('"', q1 + 1);
        if (q1 == std::string::npos || q2 == std::string::npos) return false;
        t.cfg = cfg_from_name(e.substr(q1 + 1, q2 - q1 - 1));
        if (t.cfg < 0) return false;
        out.push_back(t);
        pos = end + 1;
    }
    return !out.empty();
}

int uvt_tune_pick(const std::vector<uvt_tune_entry>& table, int64_t n, int64_t m)
{
    for (const uvt_tune_entry& t : table)
        if (n <= t.n_max && m <= t.m_max) return t.cfg;
    return uvt_default_cfg(n, m);
}

namespace {

// Tiles of the lower trapezoid of one problem.  diag_only: only the 256-row
// diagonal blocks (k from the start of the row's 256-row tile), the FP64 part
// of a split problem.
// With 128x128 tiles the last column block gets a tile as narrow as it
// allows (32, 64 or 128 columns), so a problem issues at most 31 columns of
// padding products instead of up to 127.
void trapezoid_tiles(int prob, int64_t n, int64_t m, int cfg, std::vector<BatchTile>& out,
                     bool diag_only = false)
{
    const int64_t tm = uvt_cfg_tile_m(cfg), tn = uvt_cfg_tile_n(cfg);
    for (int64_t r0 = 0; r0 < n; r0 += tm)
        for (int64_t c0 = 0; c0 < m; c0 += tn) {
            const int64_t rows = std::min(tm, n - r0);
            const int64_t k_lo = diag_only ? (r0 / 256) * 256 : 0;
            int c = cfg;
            if (cfg == UVT_CFG_128x128 && m - c0 <= 64) {
                c = m - c0 <= 32 ? UVT_CFG_128x32 : UVT_CFG_128x64;
            }
            out.push_back(BatchTile{prob, (int32_t)r0, (int32_t)rows, (int32_t)c0, (int32_t)k_lo,
                                    (int32_t)(r0 + rows), c});
        }
}

// Most work first (depth times tile area), so the queue drains with its
// smallest tiles.
void sort_longest_first(std::vector<BatchTile>& v)
{
    auto work = [](const BatchTile& t) {
        return (int64_t)(t.k_hi - t.k_lo) * uvt_cfg_tile_m(t.cfg) * uvt_cfg_tile_n(t.cfg);
    };
    std::stable_sort(v.begin(), v.end(), [&](const BatchTile& a, const BatchTile& b) { return work(a) > work(b); });
}

// Split-class problems whose integer box part is too small to pay for its
// conversions run in FP64 (which meets the split tolerance).

}  // namespace

struct uvt_batch_s {
    uvt_handle h;
    int32_t count, slices;
    std::vector<uvt_problem> probs;
    // FP64 part
    std::vector<int> f64_index;        // problem -> slot in the FP64 table
    std::vector<BatchProb> f64_probs;  // host image
    // one queue of every tile shape per 16-byte load path: index vec
    std::vector<BatchTile> tiles[2];
    BatchTile* d_tiles[2] = {nullptr, nullptr};
    BatchProb* d_probs = nullptr;
    int* d_counters = nullptr;
    // split part (problems that run the integer stages)
    std::vector<bool> int8;
    std::vector<uvt_split_problem> sp;
    void* d_split_work = nullptr;
    size_t split_work_bytes = 0;
    void* d_split_desc = nullptr;
    int* d_split_lut = nullptr;
    int64_t split_totals[UVT_SPLIT_STAGES] = {0}, split_lut_off[UVT_SPLIT_STAGES] = {0};
    std::vector<double*> split_err;  // per int8-path problem: its error statement (device)
    // graph
    hipStream_t cap = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    double flops = 0.0, int8_flops = 0.0;
};

static void batch_free(uvt_batch b)
{
    if (!b) return;
    (void)hipSetDevice(b->h->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->side) (void)hipStreamDestroy(b->side);
    for (auto& p : b->d_tiles)
        if (p) (void)hipFree(p);
    if (b->d_probs) (void)hipFree(b->d_probs);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_split_work) (void)hipFree(b->d_split_work);
    if (b->d_split_desc) (void)hipFree(b->d_split_desc);
    if (b->d_split_lut) (void)hipFree(b->d_split_lut);
    delete b;
}

static uvt_status check(const uvt_problem& p)
{
    if (!p.A || !p.B || !p.C || p.n <= 0 || p.m <= 0 || p.lda < p.n || p.ldb < p.m || p.ldc < p.m)
        return UVT_ERR_ARG;
    if (p.prec != UVT_PREC_F64 && p.prec != UVT_PREC_SPLIT) return UVT_ERR_ARG;
    if (p.route != UVT_ROUTE_AUTO && p.route != UVT_ROUTE_INT8 && p.route != UVT_ROUTE_FP64)
        return UVT_ERR_ARG;
    if (p.prec == UVT_PREC_F64 && p.route == UVT_ROUTE_INT8) return UVT_ERR_ARG;
    if (p.n > 0x3FFFFFFF || p.m > 0x3FFFFFFF) return UVT_ERR_SIZE;
    return UVT_OK;
}

// Whether a problem's box part runs on the int8 path: split class, a box to
// run (n > 256), a size the int8 path supports with these moduli (else FP64,
// which meets the split tolerance), and either asked for or measured faster
// in a batch.
static bool routes_to_int8(const uvt_problem& p, int slices)
{
    if (p.prec != UVT_PREC_SPLIT || p.route == UVT_ROUTE_FP64 || p.n <= 256) return false;
    if (!uvt_split_supported(p.n, slices)) return false;
    return p.route == UVT_ROUTE_INT8 || uvt_split_faster_in_batch(p.n, p.m);
}

static BatchProb f64_desc(const uvt_problem& p)
{
    BatchProb d;
    d.A = p.A;
    d.B = p.B;
    d.C = p.C;
    d.lda = p.lda;
    d.ldb = p.ldb;
    d.ldc = p.ldc;
    d.m = p.m;
    d.vec = ((uintptr_t)p.A % 16 == 0) && ((uintptr_t)p.B % 16 == 0) && p.lda % 2 == 0 &&
            p.ldb % 2 == 0;
    d.pad = 0;
    return d;
}

static uvt_status upload_split(uvt_batch b)
{
    if (b->sp.empty()) return UVT_OK;
    std::vector<char> desc;
    std::vector<int> lut;
    uvt_status st = uvt_split_batch_build(b->sp.data(), (int)b->sp.size(), b->slices, b->d_split_work,
                                          desc, lut, b->split_totals, b->split_lut_off, &b->split_err);
    if (st != UVT_OK) return st;
    UVT_HIP(hipMemcpy(b->d_split_desc, desc.data(), desc.size(), hipMemcpyHostToDevice));
    if (!lut.empty())
        UVT_HIP(hipMemcpy(b->d_split_lut, lut.data(), lut.size() * sizeof(int), hipMemcpyHostToDevice));
    return UVT_OK;
}

extern "C" uvt_status uvt_batch_create(uvt_handle h, const uvt_problem* probs, int32_t count,
                                       int32_t slices, const char* tune_table, uvt_batch* out)
{
    if (!h || !probs || count <= 0 || !out) return UVT_ERR_ARG;
    if (slices < 2 || slices > 14) return UVT_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        uvt_status st = check(probs[i]);
        if (st != UVT_OK) return st;
    }
    std::vector<uvt_tune_entry> table;
    if (tune_table && !uvt_tune_load(tune_table, table)) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    uvt_batch b = new (std::nothrow) uvt_batch_s;
    if (!b) return UVT_ERR_NOMEM;
    b->h = h;
    b->count = count;
    b->slices = slices;
    b->probs.assign(probs, probs + count);
    uvt_status st = UVT_OK;
    // classify: every problem has an FP64 descriptor (the whole product, or
    // the diagonal blocks of a split problem); split problems whose box part
    // pays off also enter the integer stages.
    for (int i = 0; i < count; ++i) {
        const uvt_problem& p = probs[i];
        b->flops += (double)p.m * (double)p.n * (double)p.n;
        const int slot = (int)b->f64_probs.size();
        b->f64_index.push_back(slot);
        b->f64_probs.push_back(f64_desc(p));
        const int cfg = table.empty() ? uvt_default_cfg(p.n, p.m) : uvt_tune_pick(table, p.n, p.m);
        const bool use_int8 = routes_to_int8(p, slices);
        trapezoid_tiles(slot, p.n, p.m, cfg, b->tiles[b->f64_probs.back().vec], use_int8);
        b->int8.push_back(use_int8);
        if (use_int8) b->int8_flops += (double)p.m * (double)p.n * (double)p.n;
        if (use_int8) b->sp.push_back(uvt_split_problem{p.A, p.lda, p.B, p.ldb, p.C, p.ldc, p.n, p.m});
    }
    for (auto& v : b->tiles) sort_longest_first(v);
    for (auto& v : b->tiles)
        if (v.size() > 0x7FFFFFFF) st = UVT_ERR_SIZE;
    if (st != UVT_OK) {
        batch_free(b);
        return st;
    }
#define BATCH_HIP(call)                  \
    do {                                 \
        if ((call) != hipSuccess) {      \
            (void)hipGetLastError();     \
            batch_free(b);               \
            return UVT_ERR_HIP;          \
        }                                \
    } while (0)
    // FP64 buffers
    for (int c = 0; c < 2; ++c)
        if (!b->tiles[c].empty()) {
            BATCH_HIP(hipMalloc(&b->d_tiles[c], b->tiles[c].size() * sizeof(BatchTile)));
            BATCH_HIP(hipMemcpy(b->d_tiles[c], b->tiles[c].data(), b->tiles[c].size() * sizeof(BatchTile),
                                hipMemcpyHostToDevice));
        }
    if (!b->f64_probs.empty()) {
        BATCH_HIP(hipMalloc(&b->d_probs, b->f64_probs.size() * sizeof(BatchProb)));
        BATCH_HIP(hipMemcpy(b->d_probs, b->f64_probs.data(), b->f64_probs.size() * sizeof(BatchProb),
                            hipMemcpyHostToDevice));
    }
    BATCH_HIP(hipMalloc(&b->d_counters, 128 * sizeof(int)));
    // split buffers
    if (!b->sp.empty()) {
        int64_t lut_entries = 0;
        st = uvt_split_batch_sizes(b->sp.data(), (int)b->sp.size(), slices, &b->split_work_bytes,
                                   &lut_entries);
        if (st != UVT_OK) {
            batch_free(b);
            return st;
        }
        if (hipMalloc(&b->d_split_work, b->split_work_bytes ? b->split_work_bytes : 256) != hipSuccess) {
            (void)hipGetLastError();
            batch_free(b);
            return UVT_ERR_NOMEM;
        }
        BATCH_HIP(hipMalloc(&b->d_split_desc, uvt_split_desc_bytes((int)b->sp.size())));
        BATCH_HIP(hipMalloc((void**)&b->d_split_lut, (size_t)(lut_entries ? lut_entries : 1) * sizeof(int)));
        st = upload_split(b);
        if (st != UVT_OK) {
            batch_free(b);
            return st;
        }
    }
    // capture (occupancy queries first: nothing but stream work inside the capture)
    const int grid = persistent_grid(h->cus);
    BATCH_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    BATCH_HIP(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    BATCH_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    BATCH_HIP(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
    BATCH_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(128), 0, b->cap, b->d_counters, 128);
    // scales and residues of the split problems first: short kernels that
    // would be starved behind the persistent FP64 workgroups
    if (!b->sp.empty())
        st = uvt_split_batch_launch(b->sp.data(), (int)b->sp.size(), slices, b->d_split_work,
                                    b->d_split_desc, b->d_split_lut, b->split_totals, b->split_lut_off,
                                    b->cap, UVT_SPLIT_PREP);
    BATCH_HIP(hipEventRecord(b->fork, b->cap));
    BATCH_HIP(hipStreamWaitEvent(b->side, b->fork, 0));
    // int8 residue products on the side branch, the FP64 queue on the main one
    if (st == UVT_OK && !b->sp.empty())
        st = uvt_split_batch_launch(b->sp.data(), (int)b->sp.size(), slices, b->d_split_work,
                                    b->d_split_desc, b->d_split_lut, b->split_totals, b->split_lut_off,
                                    b->side, UVT_SPLIT_GEMM);
    for (int q = 0; q < 2; ++q)
        if (!b->tiles[q].empty())
            launch_persistent(q == 1, grid, b->d_tiles[q], (int)b->tiles[q].size(), b->d_probs,
                              b->d_counters + 16 * q, b->cap);
    BATCH_HIP(hipEventRecord(b->join, b->side));
    BATCH_HIP(hipStreamWaitEvent(b->cap, b->join, 0));
    // the reconstruction adds to the diagonal blocks the FP64 queue stored
    if (st == UVT_OK && !b->sp.empty())
        st = uvt_split_batch_launch(b->sp.data(), (int)b->sp.size(), slices, b->d_split_work,
                                    b->d_split_desc, b->d_split_lut, b->split_totals, b->split_lut_off,
                                    b->cap, UVT_SPLIT_REC);
    const hipError_t cap_err = hipStreamEndCapture(b->cap, &b->graph);
    if (st != UVT_OK || cap_err != hipSuccess) {
        (void)hipGetLastError();
        batch_free(b);
        return st != UVT_OK ? st : UVT_ERR_HIP;
    }
    BATCH_HIP(hipGraphGetNodes(b->graph, nullptr, &b->nodes));
    BATCH_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
#undef BATCH_HIP
    *out = b;
    return UVT_OK;
}

extern "C" uvt_status uvt_batch_run(uvt_batch b, void* stream)
{
    if (!b) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(b->h->device));
    UVT_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return UVT_OK;
}

extern "C" uvt_status uvt_batch_update(uvt_batch b, const uvt_problem* probs, int32_t count)
{
    if (!b || !probs || count != b->count) return UVT_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        const uvt_problem &p = probs[i], &q = b->probs[i];
        if (p.n != q.n || p.m != q.m || p.prec != q.prec || p.route != q.route) return UVT_ERR_ARG;
        uvt_status st = check(p);
        if (st != UVT_OK) return st;
        // the 16-byte load path was fixed at capture time
        if (f64_desc(p).vec != b->f64_probs[b->f64_index[i]].vec) return UVT_ERR_ARG;
    }
    UVT_HIP(hipSetDevice(b->h->device));
    UVT_HIP(hipDeviceSynchronize());
    b->sp.clear();
    for (int i = 0; i < count; ++i) {
        const uvt_problem& p = probs[i];
        b->probs[i] = p;
        b->f64_probs[b->f64_index[i]] = f64_desc(p);
        if (b->int8[i]) b->sp.push_back(uvt_split_problem{p.A, p.lda, p.B, p.ldb, p.C, p.ldc, p.n, p.m});
    }
    if (!b->f64_probs.empty())
        UVT_HIP(hipMemcpy(b->d_probs, b->f64_probs.data(), b->f64_probs.size() * sizeof(BatchProb),
                          hipMemcpyHostToDevice));
    return upload_split(b);
}

extern "C" uvt_status uvt_batch_info(uvt_batch b, double* flops, double* int8_flops,
                                     int64_t* fp64_tiles, int32_t* graph_nodes)
{
    if (!b) return UVT_ERR_ARG;
    if (flops) *flops = b->flops;
    if (int8_flops) *int8_flops = b->int8_flops;
    if (fp64_tiles) {
        int64_t t = 0;
        for (auto& v : b->tiles) t += (int64_t)v.size();
        *fp64_tiles = t;
    }
    if (graph_nodes) *graph_nodes = (int32_t)b->nodes;
    return UVT_OK;
}

extern "C" uvt_status uvt_batch_errors(uvt_batch b, double* bound, double* estimate, int32_t count)
{
    if (!b || count != b->count) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(b->h->device));
    UVT_HIP(hipDeviceSynchronize());
    int k = 0;
    for (int i = 0; i < count; ++i) {
        double e[2] = {-1.0, -1.0};  // the problem ran in FP64
        if (b->int8[i]) UVT_HIP(hipMemcpy(e, b->split_err[k++], sizeof e, hipMemcpyDeviceToHost));
        if (bound) bound[i] = e[0];
        if (estimate) estimate[i] = e[1];
    }
    return UVT_OK;
}

extern "C" uvt_status uvt_batch_destroy(uvt_batch b)
{
    if (!b) return UVT_ERR_ARG;
    batch_free(b);
    return UVT_OK;
}

extern "C" uvt_status uvt_tune_choice(const char* tune_table, int64_t n, int64_t m, int32_t* cfg)
{
    if (!cfg || n <= 0 || m <= 0) return UVT_ERR_ARG;
    if (!tune_table) {
        *cfg = uvt_default_cfg(n, m);
        return UVT_OK;
    }
    std::vector<uvt_tune_entry> table;
    if (!uvt_tune_load(tune_table, table)) return UVT_ERR_ARG;
    *cfg = uvt_tune_pick(table, n, m);
    return UVT_OK;
}

// ---- autotuner -----------------------------------------------------------------

// For every (n, m) bucket, a batch of copies of one problem at the bucket's
// upper edge (enough to fill the GPU, as a mixed batch does) is timed with
// each tile shape (persistent kernel, best of three runs); the fastest shape
// is the bucket's entry.
extern "C" uvt_status uvt_autotune_f64(uvt_handle h, const char* out_path)
{
    if (!h || !out_path) return UVT_ERR_ARG;
    UVT_HIP(hipSetDevice(h->device));
    const int64_t ns[] = {64, 128, 256, 512, 1024, 2048, 4096};
    const int64_t ms[] = {16, 32, 64, 128, 256, 512, 1024};
    const int64_t nmax = 4096, mmax = 1024;
    double *A = nullptr, *B = nullptr, *C = nullptr;
    int* counter = nullptr;
    BatchProb* dprobs = nullptr;
    BatchTile* dtiles = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // Enough copies of a bucket's problem to keep every CU busy the way a
    // mixed batch does (~1.2e11 flops), within a 1 GiB C buffer.
    const int max_copies = 4096;
    const size_t c_bytes = (size_t)1 << 30;
    uvt_status st = UVT_OK;
    std::string json = "{\n  \"version\": 1,\n  \"device\": \"gfx950\",\n  \"what\": \"FP64 tile shape per (n, m) bucket, chosen by timing\",\n  \"table\": [\n";
    auto fail = [&](uvt_status e) {
        st = e;
        return e;
    };
    if (hipMalloc(&A, (size_t)nmax * nmax * 8) != hipSuccess ||
        hipMalloc(&B, (size_t)nmax * mmax * 8) != hipSuccess ||
        hipMalloc(&C, c_bytes) != hipSuccess ||
        hipMalloc(&counter, 64) != hipSuccess || hipMalloc(&dprobs, max_copies * sizeof(BatchProb)) != hipSuccess ||
        hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        fail(UVT_ERR_NOMEM);
    }
    if (st == UVT_OK) st = uvt_fill_seeded(h, nmax, mmax, 7, 0, A, nmax, B, mmax, s);
    bool first = true;
    for (int64_t n : ns) {
        for (int64_t m : ms) {
            if (st != UVT_OK) break;
            int copies = (int)std::min<double>(max_copies, std::max(1.0, 1.2e11 / ((double)m * n * n)));
            while (copies > 1 && (size_t)copies * n * m * 8 > c_bytes) --copies;
            std::vector<BatchProb> hp(copies);
            for (int c = 0; c < copies; ++c) {
                uvt_problem p = {A, nmax, B, mmax, C + (size_t)c * n * m, m, n, m, UVT_PREC_F64, 0};
                hp[c] = f64_desc(p);
            }
            if (hipMemcpy(dprobs, hp.data(), copies * sizeof(BatchProb), hipMemcpyHostToDevice) != hipSuccess) {
                fail(UVT_ERR_HIP);
                break;
            }
            int best_cfg = uvt_default_cfg(n, m);
            const int grid = persistent_grid(h->cus);
            float best_ms = 1e30f;
            float per_cfg[UVT_NUM_CFGS] = {0, 0, 0};
            for (int cfg = 0; cfg < UVT_NUM_CFGS; ++cfg) {
                std::vector<BatchTile> tiles;
                for (int c = 0; c < copies; ++c) trapezoid_tiles(c, n, m, cfg, tiles);
                sort_longest_first(tiles);
                if (dtiles) (void)hipFree(dtiles);
                dtiles = nullptr;
                if (hipMalloc(&dtiles, tiles.size() * sizeof(BatchTile)) != hipSuccess ||
                    hipMemcpy(dtiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice) != hipSuccess) {
                    fail(UVT_ERR_HIP);
                    break;
                }
                float cfg_best = 1e30f;
                for (int rep = 0; rep < 4; ++rep) {
                    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, counter, 16);
                    (void)hipEventRecord(e0, s);
                    launch_persistent(true, grid, dtiles, (int)tiles.size(), dprobs, counter, s);
                    (void)hipEventRecord(e1, s);
                    float ms_ = 0.f;
                    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms_, e0, e1) != hipSuccess) {
                        fail(UVT_ERR_HIP);
                        break;
                    }
                    if (rep > 0) cfg_best = std::min(cfg_best, ms_);  // first run warms up
                }
                per_cfg[cfg] = cfg_best;
                if (cfg_best < best_ms) {
                    best_ms = cfg_best;
                    best_cfg = cfg;
                }
            }
            if (st != UVT_OK) break;
            char line[512];
            const double fl = (double)m * n * n * copies;
            snprintf(line, sizeof line,
                     "%s    {\"n_max\": %lld, \"m_max\": %lld, \"cfg\": \"%s\", \"copies\": %d, "
                     "\"ms\": {\"128x128\": %.4f, \"64x64\": %.4f, \"128x32\": %.4f}, \"tflops\": %.3f}",
                     first ? "" : ",\n", (long long)n, (long long)m, cfg_name(best_cfg), copies,
                     per_cfg[UVT_CFG_128x128], per_cfg[UVT_CFG_64x64], per_cfg[UVT_CFG_128x32],
                     fl / (best_ms * 1e-3) * 1e-12);
            json += line;
            first = false;
        }
    }
    json += "\n  ]\n}\n";
    if (st == UVT_OK) {
        FILE* f = fopen(out_path, "wb");
        if (!f || fwrite(json.data(), 1, json.size(), f) != json.size()) st = UVT_ERR_ARG;
        if (f) fclose(f);
    }
    if (dtiles) (void)hipFree(dtiles);
    if (dprobs) (void)hipFree(dprobs);
    if (counter) (void)hipFree(counter);
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (C) (void)hipFree(C);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return st;
}
