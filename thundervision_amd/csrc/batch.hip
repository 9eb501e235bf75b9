// Batched execution: many stereo pairs of uneven size and disparity range
// (BT + SGM-8 + WTA each), captured as one hipGraph.
//
// Work items, one wave each:
//   BT    rows of a pair's term planes
//   LINES a group of 4 (or 8) consecutive lines of one SGM direction
//   SUM   a run of pixels of a pair: plane sum + lowest-index WTA
// Pairs are grouped by m (words per lane, scan_m(D)), the register shape
// of the SGM code, and by the scans' cost source (BT recomputed from the
// term planes, or relative cost bytes, chosen per pair as the per-pair path
// chooses; the byte form runs at fusion depth 1), so each class's kernels
// are compiled for their own shape and keep their occupancy; the classes
// run one after another (depth 1) or as parallel branches of the graph
// (persistent queues). A byte class runs one more phase
// before its line scans: the relative cost bytes of its pairs.
//
// Fusion depth (how many of the three phases share one launch per class):
//   1  one launch per phase (BT of all pairs; per class the line scans, then
//      the sums), kernel boundaries ordering the phases;
//   2  BT of all pairs, then per class one persistent launch that runs the
//      line groups and the sums from a device work queue;
//   3  per class one persistent launch running BT, line groups and sums.
// Persistent launches: each wave takes the next item of its class's queue
// with one global atomic add; items are queued so that an item's
// dependencies come earlier (BT of every pair, then the line groups pair by
// pair, longest first within a pair, with a pair's sums queued after the
// line groups of the next `sum_lag` pairs, or after all of them). A LINES item waits for its pair's BT counter, a
// SUM item for its pair's LINES counter; every wait is on items already
// taken by running waves, so it ends (a spin limit bounds it anyway and sets
// an error flag that tdvh_batch_wait reports). A wave publishes finished
// items of one pair and kind together (one device-scope release and one
// atomic add when it moves to another pair or kind, or before it waits) and
// acquires once per pair it waits on. The arithmetic is the per-pair path's
// (the same device functions), so every pair's disparities equal the plain
// call's.
#include <algorithm>
#include <vector>

#include "sgm_device.hpp"

namespace tdvh {

struct BPair {
    SgmArgs a;
    const uint8_t *L, *R;
    int pitch, m;
    int16_t *disp;
};

enum { IT_BT = 0, IT_LINES = 1, IT_SUM = 3 };

struct BItem {
    int pair, kind, a, b; // BT: first row, rows; LINES: direction, first line;
                          // SUM: first wave unit, wave units
};

constexpr int BT_ROWS = 4;             // rows per BT item
constexpr int SUM_UNITS_PERSIST = 64;  // wave units per queued SUM item
constexpr uint32_t SPIN_LIMIT = 1u << 24;
// control words: a queue head per class, the error flag, then per pair the
// BT and LINES completion counters
constexpr int CTL_ERR = 9, CTL_PAIRS = 16;
// classes: m + NM * (relative cost bytes)
constexpr int NM = 9, NCLS = 2 * NM;

// ---- per-item work (shared by the phase kernels and the persistent one) ----

__device__ __forceinline__ void run_bt_item(const BPair &P, const BItem &it)
{
    const int lane = threadIdx.x & 63, Wr = P.a.Wr;
    for (int y = it.a; y < it.a + it.b; y++)
        for (int i = lane; i < Wr; i += 64)
            bt_planes_at(P.L, P.R, P.a.W, P.a.H, P.pitch, P.m, (uint32_t *)P.a.LT,
                         (uint32_t *)P.a.RT, Wr, i, y);
}

// Wide plane sum of wave units [u0, u0 + n): unit u covers pixels
// u * (64 / G) .. + 64 / G - 1, G lanes each.
template <int M>
__device__ __forceinline__ void run_sum_units(const BPair &P, int u0, int n)
{
    constexpr int G = sum_group<M>();
    const int lane = threadIdx.x & 63;
    const long npix = (long)P.a.W * P.a.H;
    for (int u = u0; u < u0 + n; u++) {
        const long p = (long)u * (64 / G) + lane / G;
        if (p < npix) sgm_sum_pixel_wide<M, true>(P.a, 8, nullptr, P.disp, P.a.W, p, lane % G);
    }
}

// Relative cost bytes of wave units [u0, u0 + n) (the units of the sums).
template <int M>
__device__ __forceinline__ void run_rel_units(const BPair &P, int u0, int n)
{
    constexpr int G = sum_group<M>();
    const int lane = threadIdx.x & 63;
    const long npix = (long)P.a.W * P.a.H;
    for (int u = u0; u < u0 + n; u++) {
        const long p = (long)u * (64 / G) + lane / G;
        if (p < npix) sgm_rel_pixel<M, true>(P.a, p, lane % G);
    }
}

// ---- fusion depth 1: one launch per phase ------------------------------------

__global__ __launch_bounds__(256) void batch_bt_kernel(const BPair *__restrict__ pairs,
                                                       const BItem *__restrict__ items, int n)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const BItem it = items[w];
    run_bt_item(pairs[it.pair], it);
}

template <int M, int PF>
__global__ __launch_bounds__(256) void batch_lines_kernel(const BPair *__restrict__ pairs,
                                                          const BItem *__restrict__ items, int n,
                                                          int prio)
{
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= n) return;
    const BItem it = items[w];
    sgm_line_group<M, true, true, PF>(pairs[it.pair].a, it.a, it.b, it.a, prio);
}

// Buffer-addressed line groups of one class (LPL lanes per line, cost
// source SRC); waves_per_group waves per workgroup, two waves per SIMD.
template <int M, int SRC, bool PAD, int PF, int LPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
batch_lines_buf_kernel(const BPair *__restrict__ pairs, const BItem *__restrict__ items, int n,
                       int prio)
{
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                 (threadIdx.x >> 6));
    if (w >= n) return;
    const BItem it = items[w];
    sgm_line_group_any<M, SRC, PAD, PF, LPL>(pairs[it.pair].a, it.a, it.b, it.a, prio);
}

// Plane sum + WTA (REL false) or relative cost bytes (REL true) over the
// class's SUM items {pair, SUM, first wave, waves}, one per pair in
// increasing first wave; a wave finds its pair by binary search
// (wave-uniform, scalar loads).
template <int M, bool REL>
__global__ __launch_bounds__(256) void batch_sum_wide_kernel(const BPair *__restrict__ pairs,
                                                             const BItem *__restrict__ items,
                                                             int n, int waves)
{
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= waves) return;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].a <= w) lo = mid;
        else hi = mid - 1;
    }
    const BItem it = items[lo];
    if constexpr (REL) run_rel_units<M>(pairs[it.pair], w - it.a, 1);
    else run_sum_units<M>(pairs[it.pair], w - it.a, 1);
}

// ---- fusion depth 2 and 3: one persistent launch per class ----------------------

template <int M, bool PAD, int PF, int LPL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
batch_queue_kernel(const BPair *__restrict__ pairs, const BItem *__restrict__ items, int nitems,
                   const int *__restrict__ need, uint32_t *ctl, int npairs, int mc, int wait_bt,
                   int prio)
{
    const int lane = threadIdx.x & 63;
    uint32_t *head = ctl + mc, *err = ctl + CTL_ERR;
    uint32_t *bt_done = ctl + CTL_PAIRS, *lines_done = bt_done + npairs;
    int cur_pair = -1, cur_kind = -1;
    uint32_t pending = 0;
    int acq_bt = -1, acq_lines = -1; // pairs whose counters this wave has acquired
    // publishes this wave's finished items of (cur_pair, cur_kind)
    auto flush = [&]() {
        if (pending) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            uint32_t *c = (cur_kind == IT_BT ? bt_done : lines_done) + cur_pair;
            if (lane == 0) __hip_atomic_fetch_add(c, pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pending = 0;
        }
    };
    auto wait_for = [&](const uint32_t *c, uint32_t want) -> bool {
        uint32_t spins = 0;
        while ((uint32_t)__builtin_amdgcn_readfirstlane(
                   (int)__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < want) {
            if (++spins > SPIN_LIMIT) return false;
            __builtin_amdgcn_s_sleep(8);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        return true;
    };
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = __hip_atomic_fetch_add(head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)idx);
        if (idx >= (uint32_t)nitems) break;
        const BItem it = items[idx];
        if (pending && (it.pair != cur_pair || it.kind != cur_kind)) flush();
        bool ok = true;
        if (it.kind == IT_LINES && wait_bt && acq_bt != it.pair) {
            flush();
            ok = wait_for(bt_done + it.pair, (uint32_t)need[it.pair]);
            acq_bt = it.pair;
        } else if (it.kind == IT_SUM && acq_lines != it.pair) {
            flush();
            ok = wait_for(lines_done + it.pair, (uint32_t)need[npairs + it.pair]);
            acq_lines = it.pair;
        }
        if (!ok) {
            if (lane == 0) __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        const BPair &P = pairs[it.pair];
        if (it.kind == IT_BT) {
            run_bt_item(P, it);
        } else if (it.kind == IT_LINES) {
            sgm_line_group_any<M, SRC_TERMS, PAD, PF, LPL>(P.a, it.a, it.b, it.a, prio);
            __builtin_amdgcn_s_setprio(0);
        } else {
            run_sum_units<M>(P, it.a, it.b);
        }
        if (it.kind != IT_SUM) {
            cur_pair = it.pair;
            cur_kind = it.kind;
            pending++;
        }
    }
    flush();
}

__global__ void batch_reset_kernel(uint32_t *ctl, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        ctl[i] = 0;
}

} // namespace tdvh

using namespace tdvh;

struct tdvh_batch {
    tdvh_ctx *ctx = nullptr;
    int n = 0;
    std::vector<tdvh_pair> host_pairs;
    std::vector<BPair> host_bp; // the device descriptors, host copy
    tdvh_sgm_params params{};
    void *mem = nullptr; // pair buffers, descriptors, control words
    size_t mem_bytes = 0;
    BPair *d_pairs = nullptr;
    uint32_t *d_ctl = nullptr;
    int nctl = 0;
    // item plan (rebuilt when lines per wave or the fusion depth change)
    int *d_need = nullptr;    // per pair: BT items, then LINES items (own allocation)
    BItem *d_items = nullptr; // the items, after the need counts
    size_t items_cap = 0;
    struct Class {
        bool used = false, pad = false, buf = true;
        int bt_begin = 0, bt_count = 0;     // depth 3: BT items of the class
        int line_begin = 0, line_count = 0; // depth 1
        int sum_begin = 0, sum_count = 0, sum_waves = 0;
        int queue_begin = 0, queue_count = 0; // depth 2, 3
    } cls[NCLS];
    int bt_begin = 0, bt_count = 0; // depth 1, 2: BT items of every pair
    int64_t cells = 0;
    // knobs
    int lines_per_wave = 4, waves_per_group = 1, prefetch = 8, depth = 1, prio = 1;
    int sum_lag = -1; // persistent queues: pairs between a pair's lines and its sums
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    hipStream_t branch[NCLS] = {};
    hipEvent_t fork = nullptr, join[NCLS] = {};
};

static void drop_graph(tdvh_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Line groups of one pair, longest first.
static void pair_line_items(const tdvh_pair &p, int i, int lpw, std::vector<BItem> &out)
{
    std::vector<std::pair<int, BItem>> lines;
    for (int dir = 0; dir < 8; dir++) {
        const int nl = dir_lines(dir, p.width, p.height);
        for (int l0 = 0; l0 < nl; l0 += lpw) {
            int longest = 0;
            for (int l = l0; l < std::min(l0 + lpw, nl); l++) {
                int sx, sy, len;
                line_geom(dir, l, p.width, p.height, sx, sy, len);
                longest = std::max(longest, len);
            }
            lines.push_back({longest, BItem{i, IT_LINES, dir, l0}});
        }
    }
    std::stable_sort(lines.begin(), lines.end(),
                     [](const std::pair<int, BItem> &x, const std::pair<int, BItem> &y) {
                         return x.first > y.first;
                     });
    for (auto &l : lines) out.push_back(l.second);
}

static int pair_sum_units(const tdvh_pair &p, int mc)
{
    const long npix = (long)p.width * p.height;
    const int G = mc == 6 ? 16 : 2 * mc;
    return (int)((npix * G + 63) / 64);
}

// Class of pair i under the current fusion depth.
static int pair_class(const tdvh_batch *b, int i)
{
    const BPair &q = b->host_bp[i];
    return q.m + (b->depth == 1 && q.a.CR && buf_lines_fit(q.a) ? NM : 0);
}

// Lines per wave usable for class ci: 8 needs 8-lane lines (2m words per
// lane, at most 4, which keeps the prefetch ring within two waves per SIMD)
// and the buffer kernel.
static int class_lpw(const tdvh_batch *b, int ci)
{
    return (b->lines_per_wave == 8 && ci % NM <= 2 && b->cls[ci].buf) ? 8 : 4;
}

// Builds the item lists for the current knobs and uploads them (the batch
// is idle: callers synchronise first).
static int plan_items(tdvh_batch *b)
{
    const int n = b->n;
    std::vector<BItem> items;
    std::vector<int> need(2 * n, 0);
    auto bt_items = [&](int i) {
        const int H = b->host_pairs[i].height;
        for (int y = 0; y < H; y += BT_ROWS) {
            items.push_back({i, IT_BT, y, std::min(BT_ROWS, H - y)});
            need[i]++;
        }
    };
    b->bt_begin = (int)items.size();
    if (b->depth < 3)
        for (int i = 0; i < n; i++) bt_items(i);
    b->bt_count = (int)items.size() - b->bt_begin;
    // class flags under the current depth
    for (tdvh_batch::Class &c : b->cls) c = tdvh_batch::Class{};
    for (int i = 0; i < n; i++) {
        const BPair &q = b->host_bp[i];
        tdvh_batch::Class &c = b->cls[pair_class(b, i)];
        c.used = true;
        if (q.a.D < 32 * q.m) c.pad = true;
        if (!buf_lines_fit(q.a)) c.buf = false;
    }
    for (int ci = 1; ci < NCLS; ci++) {
        tdvh_batch::Class &c = b->cls[ci];
        if (!c.used) continue;
        const int mc = ci % NM;
        std::vector<int> members;
        for (int i = 0; i < n; i++)
            if (pair_class(b, i) == ci) members.push_back(i);
        const int lpw = class_lpw(b, ci);
        const bool persistent = b->depth >= 2 && ci < NM && c.buf;
        if (!persistent) {
            // each pair's line groups in the per-pair path's chiplet-banded,
            // longest-first order for this launch's waves per workgroup
            c.line_begin = (int)items.size();
            std::vector<uint32_t> tab;
            for (int i : members) {
                const tdvh_pair &p = b->host_pairs[i];
                const size_t before = tab.size();
                line_order_append(p.width, p.height, 0xFF, lpw, b->waves_per_group, tab);
                for (size_t k = before; k < tab.size(); k++)
                    items.push_back({i, IT_LINES, (int)((tab[k] >> 24) & 0xF),
                                     (int)(tab[k] & 0xFFFFFF)});
                need[n + i] = (int)(tab.size() - before);
            }
            c.line_count = (int)items.size() - c.line_begin;
            c.sum_begin = (int)items.size();
            int wave = 0;
            for (int i : members) {
                const int nw = pair_sum_units(b->host_pairs[i], mc);
                items.push_back({i, IT_SUM, wave, nw});
                wave += nw;
            }
            c.sum_count = (int)items.size() - c.sum_begin;
            c.sum_waves = wave;
            c.queue_count = 0;
            continue;
        }
        // one queue: [BT of the class's pairs (depth 3)], then pair by pair
        // the line groups, each pair's sums after the next pair's lines
        c.queue_begin = (int)items.size();
        if (b->depth == 3)
            for (int i : members) bt_items(i);
        auto sums = [&](int i) {
            const int units = pair_sum_units(b->host_pairs[i], mc);
            for (int u = 0; u < units; u += SUM_UNITS_PERSIST)
                items.push_back({i, IT_SUM, u, std::min(SUM_UNITS_PERSIST, units - u)});
        };
        // sum_lag: a pair's sums follow the line groups of that many later
        // pairs (-1: all sums after all line groups of the class)
        const int lag = b->sum_lag;
        for (size_t k = 0; k < members.size(); k++) {
            const int i = members[k];
            const size_t before = items.size();
            pair_line_items(b->host_pairs[i], i, lpw, items);
            need[n + i] = (int)(items.size() - before);
            if (lag > 0 && (int)k >= lag) sums(members[k - lag]);
        }
        for (size_t k = lag > 0 ? std::max<long>(0, (long)members.size() - lag) : 0;
             k < members.size(); k++)
            sums(members[k]);
        c.queue_count = (int)items.size() - c.queue_begin;
    }
    // device copies: the need counts sit in front of the items (one
    // allocation, grown when a plan needs more)
    const size_t need_bytes = (need.size() * sizeof(int) + 255) & ~(size_t)255;
    const size_t bytes = need_bytes + items.size() * sizeof(BItem);
    if (bytes > b->items_cap) {
        if (b->d_need) TDVH_CHECK_HIP(hipFree(b->d_need));
        b->d_need = nullptr;
        b->d_items = nullptr;
        b->items_cap = 0;
        void *p = nullptr;
        TDVH_CHECK_HIP(hipMalloc(&p, bytes));
        b->d_need = (int *)p;
        b->items_cap = bytes;
    }
    b->d_items = (BItem *)((uint8_t *)b->d_need + need_bytes);
    TDVH_CHECK_HIP(hipMemcpy(b->d_need, need.data(), need.size() * sizeof(int),
                             hipMemcpyHostToDevice));
    TDVH_CHECK_HIP(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BItem),
                             hipMemcpyHostToDevice));
    return TDVH_OK;
}

extern "C" int tdvh_batch_create(tdvh_ctx *ctx, const tdvh_pair *pairs, int n,
                                 const tdvh_sgm_params *params, tdvh_batch **out)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(pairs && out && n > 0 && params, "bad arguments");
    TDVH_REQUIRE(params->dir_mask == 0xFF && params->flags == 0,
                 "batches run the standard 8-direction geometry");
    TDVH_REQUIRE(params->p1 >= 0 && params->p1 <= 1024,
                 "P1 %d outside [0, 1024] (the per-pair limit)", params->p1);
    TDVH_REQUIRE(params->p2 >= 0 && params->p2 <= 255,
                 "batches need 0 <= P2 <= 255 (byte path-cost planes)");
    std::vector<BPair> bp(n);
    std::vector<size_t> off_terms(n), off_planes(n), off_rel(n);
    size_t total = 0;
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    int64_t cells = 0;
    for (int i = 0; i < n; i++) {
        const tdvh_pair &p = pairs[i];
        TDVH_REQUIRE(p.left && p.right && p.disp, "pair %d: null pointer", i);
        TDVH_REQUIRE(p.width > 0 && p.height > 0 && p.pitch >= p.width, "pair %d: bad image", i);
        TDVH_REQUIRE(p.depth > 0 && p.depth <= MAX_DEPTH, "pair %d: depth %d outside [1, %d]", i,
                     p.depth, MAX_DEPTH);
        const int m = scan_m(p.depth), Wr = p.width + 16 * m;
        const size_t W = p.width, H = p.height;
        off_terms[i] = total;
        total = align(total + (W * H + 3 * (size_t)Wr * H) * 4);
        off_planes[i] = total;
        total = align(total + 8 * W * H * 32 * m);
        off_rel[i] = 0;
        if (sgm_rel_auto(p.width, p.height, m, params->p2, false)) {
            off_rel[i] = total;
            total = align(total + W * H * 32 * m);
        }
        cells += (int64_t)W * H * p.depth;
    }
    const size_t trash_off = total;
    total = align(total + 64 * 64);
    const size_t pairs_off = total;
    total = align(total + sizeof(BPair) * n);
    const size_t ctl_off = total;
    const int nctl = CTL_PAIRS + 2 * n;
    total = align(total + 4 * (size_t)nctl);

    tdvh_batch *b = new tdvh_batch();
    hipError_t e = hipMalloc(&b->mem, total);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        set_error("hipMalloc(%zu) for the batch failed: %s", total, hipGetErrorString(e));
        return TDVH_ERR_OUT_OF_MEMORY;
    }
    uint8_t *base = (uint8_t *)b->mem;
    for (int i = 0; i < n; i++) {
        const tdvh_pair &p = pairs[i];
        BPair &q = bp[i];
        const int m = scan_m(p.depth);
        q.a = SgmArgs{};
        q.a.W = p.width;
        q.a.H = p.height;
        q.a.D = p.depth;
        q.a.P1 = params->p1;
        q.a.P2 = params->p2;
        q.a.C = nullptr;
        q.a.m = m;
        q.a.Wr = p.width + 16 * m;
        q.a.LT = (const uint32_t *)(base + off_terms[i]);
        q.a.RT = q.a.LT + (size_t)p.width * p.height;
        q.a.planes = base + off_planes[i];
        q.a.plane_bytes = (size_t)p.width * p.height * 32 * m;
        q.a.trash = base + trash_off;
        q.a.CR = off_rel[i] ? base + off_rel[i] : nullptr;
        q.L = p.left;
        q.R = p.right;
        q.pitch = p.pitch;
        q.m = m;
        q.disp = p.disp;
    }
    b->ctx = ctx;
    b->n = n;
    b->host_pairs.assign(pairs, pairs + n);
    b->params = *params;
    b->mem_bytes = total;
    b->d_pairs = (BPair *)(base + pairs_off);
    b->d_ctl = (uint32_t *)(base + ctl_off);
    b->nctl = nctl;
    b->host_bp = bp;
    b->cells = cells;
    int rc = TDVH_OK;
    if (hipMemcpy(b->d_pairs, bp.data(), sizeof(BPair) * n, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("uploading the batch descriptors failed");
        rc = TDVH_ERR_DEVICE;
    }
    if (rc == TDVH_OK) rc = plan_items(b);
    if (rc != TDVH_OK) {
        if (b->d_need) (void)hipFree(b->d_need);
        (void)hipFree(b->mem);
        delete b;
        return rc;
    }
    *out = b;
    return TDVH_OK;
}

extern "C" int tdvh_batch_destroy(tdvh_batch *b)
{
    if (!b) return TDVH_OK;
    clear_error();
    TDVH_TRY(use_device(b->ctx));
    TDVH_CHECK_HIP(hipDeviceSynchronize());
    drop_graph(b);
    if (b->graph_stream) (void)hipStreamDestroy(b->graph_stream);
    for (int ci = 0; ci < NCLS; ci++) {
        if (b->branch[ci]) (void)hipStreamDestroy(b->branch[ci]);
        if (b->join[ci]) (void)hipEventDestroy(b->join[ci]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->d_need) (void)hipFree(b->d_need);
    (void)hipFree(b->mem);
    delete b;
    return TDVH_OK;
}

extern "C" int tdvh_batch_config(tdvh_batch *b, int lines_per_wave, int waves_per_group,
                                 int prefetch, int fusion_depth, int sum_lag)
{
    clear_error();
    TDVH_REQUIRE(b, "null batch");
    TDVH_TRY(use_device(b->ctx));
    TDVH_REQUIRE(lines_per_wave == 0 || lines_per_wave == 4 || lines_per_wave == 8,
                 "lines per wave must be 4 or 8");
    TDVH_REQUIRE(waves_per_group == 0 || waves_per_group == 1 || waves_per_group == 2 ||
                     waves_per_group == 4,
                 "waves per group must be 1, 2 or 4");
    TDVH_REQUIRE(prefetch == 0 || prefetch == 4 || prefetch == 8, "prefetch must be 4 or 8");
    TDVH_REQUIRE(fusion_depth >= 0 && fusion_depth <= 3, "fusion depth must be 1, 2 or 3");
    TDVH_REQUIRE(sum_lag >= -1, "sum lag must be -1 or at least 0");
    const int lpw = lines_per_wave ? lines_per_wave : b->lines_per_wave;
    const int wpg = waves_per_group ? waves_per_group : b->waves_per_group;
    const int pf = prefetch ? prefetch : b->prefetch;
    const int dep = fusion_depth ? fusion_depth : b->depth;
    const int lag = sum_lag ? sum_lag : b->sum_lag;
    const bool replan = lpw != b->lines_per_wave || dep != b->depth || lag != b->sum_lag ||
                        wpg != b->waves_per_group;
    if (replan || wpg != b->waves_per_group || pf != b->prefetch) drop_graph(b);
    b->lines_per_wave = lpw;
    b->waves_per_group = wpg;
    b->prefetch = pf;
    b->depth = dep;
    b->sum_lag = lag;
    if (replan) {
        // queued runs may still read the previous item lists
        TDVH_CHECK_HIP(hipDeviceSynchronize());
        TDVH_TRY(plan_items(b));
    }
    return TDVH_OK;
}

template <int M, int SRC, bool PAD, int PF>
static void launch_lines_buf(tdvh_batch *b, int lpl, const BItem *li, int nl, int blocks,
                             int threads, hipStream_t s)
{
    if constexpr (M <= 2) {
        if (lpl == 8) {
            batch_lines_buf_kernel<M, SRC, PAD, PF, 8><<<blocks, threads, 0, s>>>(b->d_pairs, li,
                                                                                  nl, b->prio);
            return;
        }
    }
    batch_lines_buf_kernel<M, SRC, PAD, PF, 16><<<blocks, threads, 0, s>>>(b->d_pairs, li, nl,
                                                                           b->prio);
}

template <int M>
static void launch_lines_class(tdvh_batch *b, const tdvh_batch::Class &c, int ci, hipStream_t s)
{
    const int threads = 64 * b->waves_per_group, wpb = b->waves_per_group;
    const int nl = c.line_count;
    if (!nl) return;
    const BItem *li = b->d_items + c.line_begin;
    const int blocks = (nl + wpb - 1) / wpb;
    const int lpl = 64 / class_lpw(b, ci);
    if constexpr (rel_ok<M>()) {
        if (ci >= NM) {
            // relative cost bytes of the class's pairs, then the byte scans
            // (16 steps in flight, as the per-pair path)
            prof_begin(s);
            batch_sum_wide_kernel<M, true><<<(c.sum_waves + 3) / 4, 256, 0, s>>>(
                b->d_pairs, b->d_items + c.sum_begin, c.sum_count, c.sum_waves);
            (void)check_launch("batch_rel_kernel", s);
            prof_begin(s);
            launch_lines_buf<M, SRC_REL, false, 16>(b, lpl, li, nl, blocks, threads, s);
            (void)check_launch("batch_lines_kernel", s);
            return;
        }
    }
    prof_begin(s);
    if (c.buf) {
        // "prefetch 8": the deepest ring that keeps two waves per SIMD
        constexpr int PF8 = buf_lines_pf<M, true>();
        if (c.pad) {
            if (b->prefetch == 8)
                launch_lines_buf<M, SRC_TERMS, true, PF8>(b, lpl, li, nl, blocks, threads, s);
            else launch_lines_buf<M, SRC_TERMS, true, 4>(b, lpl, li, nl, blocks, threads, s);
        } else {
            if (b->prefetch == 8)
                launch_lines_buf<M, SRC_TERMS, false, PF8>(b, lpl, li, nl, blocks, threads, s);
            else launch_lines_buf<M, SRC_TERMS, false, 4>(b, lpl, li, nl, blocks, threads, s);
        }
    } else {
        if (b->prefetch == 8)
            batch_lines_kernel<M, 8><<<blocks, threads, 0, s>>>(b->d_pairs, li, nl, b->prio);
        else
            batch_lines_kernel<M, 4><<<blocks, threads, 0, s>>>(b->d_pairs, li, nl, b->prio);
    }
    (void)check_launch("batch_lines_kernel", s);
}

template <int M>
static void launch_sum_class(tdvh_batch *b, const tdvh_batch::Class &c, hipStream_t s)
{
    if (!c.sum_count) return;
    const BItem *si = b->d_items + c.sum_begin;
    prof_begin(s);
    batch_sum_wide_kernel<M, false><<<(c.sum_waves + 3) / 4, 256, 0, s>>>(b->d_pairs, si,
                                                                           c.sum_count, c.sum_waves);
    (void)check_launch("batch_sum_kernel", s);
}

template <int M, bool PAD, int PF>
static void launch_queue_pf(tdvh_batch *b, int lpl, const tdvh_batch::Class &c, int mc, int grid,
                            int threads, hipStream_t s)
{
    const BItem *qi = b->d_items + c.queue_begin;
    const int wait_bt = b->depth == 3;
    if constexpr (M <= 2) {
        if (lpl == 8) {
            batch_queue_kernel<M, PAD, PF, 8><<<grid, threads, 0, s>>>(
                b->d_pairs, qi, c.queue_count, b->d_need, b->d_ctl, b->n, mc, wait_bt, b->prio);
            return;
        }
    }
    batch_queue_kernel<M, PAD, PF, 16><<<grid, threads, 0, s>>>(
        b->d_pairs, qi, c.queue_count, b->d_need, b->d_ctl, b->n, mc, wait_bt, b->prio);
}

// One persistent launch for the class: as many waves as stay resident at two
// per SIMD (fewer for short queues).
template <int M>
static void launch_queue_class(tdvh_batch *b, const tdvh_batch::Class &c, int ci, hipStream_t s)
{
    if (!c.queue_count) return;
    const int wpb = b->waves_per_group, threads = 64 * wpb;
    const int resident = 2 * 4 * std::max(b->ctx->num_cus, 1);
    const int waves = std::min(c.queue_count, resident);
    const int grid = (waves + wpb - 1) / wpb;
    const int lpl = 64 / class_lpw(b, ci);
    constexpr int PF8 = buf_lines_pf<M, true>();
    prof_begin(s);
    if (c.pad) {
        if (b->prefetch == 8) launch_queue_pf<M, true, PF8>(b, lpl, c, ci, grid, threads, s);
        else launch_queue_pf<M, true, 4>(b, lpl, c, ci, grid, threads, s);
    } else {
        if (b->prefetch == 8) launch_queue_pf<M, false, PF8>(b, lpl, c, ci, grid, threads, s);
        else launch_queue_pf<M, false, 4>(b, lpl, c, ci, grid, threads, s);
    }
    (void)check_launch("batch_queue_kernel", s);
}

// Phase `what` (0 lines, 1 sums, 2 persistent queue) of class ci on `s`.
static int launch_class(tdvh_batch *b, int ci, int what, hipStream_t s)
{
    const tdvh_batch::Class &c = b->cls[ci];
    switch (ci % NM) {
#define CASE(m)                                                                                    \
    case m:                                                                                        \
        if (what == 0) launch_lines_class<m>(b, c, ci, s);                                         \
        else if (what == 1) launch_sum_class<m>(b, c, s);                                          \
        else launch_queue_class<m>(b, c, ci, s);                                                   \
        break;
        CASE(1) CASE(2) CASE(4) CASE(6) CASE(8)
#undef CASE
    default:
        set_error("no batch kernels for %d words per lane", ci % NM);
        return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("batch_class", s);
}

// Persistent queues: BT-scan classes at fusion depth 2 and 3 (the classes
// of relative cost bytes exist at depth 1 only).
static bool class_persistent(const tdvh_batch *b, int ci)
{
    return b->depth >= 2 && ci < NM && b->cls[ci].buf;
}

// Enqueues the batch on `s`. With fork each class runs on a branch stream
// forked from and joined back into `s`.
static int enqueue_batch(tdvh_batch *b, hipStream_t s, bool fork)
{
    if (b->depth >= 2) {
        prof_begin(s);
        batch_reset_kernel<<<1, 256, 0, s>>>(b->d_ctl, b->nctl);
        TDVH_TRY(check_launch("batch_reset_kernel", s));
    }
    if (b->bt_count) {
        prof_begin(s);
        batch_bt_kernel<<<(b->bt_count + 3) / 4, 256, 0, s>>>(b->d_pairs, b->d_items + b->bt_begin,
                                                              b->bt_count);
        TDVH_TRY(check_launch("batch_bt_kernel", s));
    }
    if (!fork) {
        for (int ci = 1; ci < NCLS; ci++) {
            if (!b->cls[ci].used) continue;
            if (class_persistent(b, ci)) {
                TDVH_TRY(launch_class(b, ci, 2, s));
            } else {
                TDVH_TRY(launch_class(b, ci, 0, s));
                TDVH_TRY(launch_class(b, ci, 1, s));
            }
        }
        return TDVH_OK;
    }
    if (!b->fork) TDVH_CHECK_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    TDVH_CHECK_HIP(hipEventRecord(b->fork, s));
    for (int ci = 1; ci < NCLS; ci++) {
        if (!b->cls[ci].used) continue;
        if (!b->branch[ci])
            TDVH_CHECK_HIP(hipStreamCreateWithFlags(&b->branch[ci], hipStreamNonBlocking));
        if (!b->join[ci])
            TDVH_CHECK_HIP(hipEventCreateWithFlags(&b->join[ci], hipEventDisableTiming));
        hipStream_t ls = b->branch[ci];
        TDVH_CHECK_HIP(hipStreamWaitEvent(ls, b->fork, 0));
        if (class_persistent(b, ci)) {
            TDVH_TRY(launch_class(b, ci, 2, ls));
        } else {
            TDVH_TRY(launch_class(b, ci, 0, ls));
            TDVH_TRY(launch_class(b, ci, 1, ls));
        }
        TDVH_CHECK_HIP(hipEventRecord(b->join[ci], ls));
        TDVH_CHECK_HIP(hipStreamWaitEvent(s, b->join[ci], 0));
    }
    return TDVH_OK;
}

extern "C" int tdvh_batch_run(tdvh_batch *b, int use_graph, void *stream)
{
    clear_error();
    TDVH_REQUIRE(b, "null batch");
    TDVH_TRY(use_device(b->ctx));
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) return enqueue_batch(b, s, false);
    if (!b->exec) {
        // capture on a private stream (the caller's may be the legacy null
        // stream, which cannot be captured); the graph then launches on `s`
        if (!b->graph_stream)
            TDVH_CHECK_HIP(hipStreamCreateWithFlags(&b->graph_stream, hipStreamNonBlocking));
        hipStream_t cs = b->graph_stream;
        TDVH_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        // depth 1: the classes one after another on the capture stream
        // (measured on the 64-pair mix: 24.7 ms against 24.9-27.0 ms as
        // parallel branches, whose launches only interleave on a device each
        // already fills); persistent queues run as parallel branches
        const int rc = enqueue_batch(b, cs, b->depth >= 2);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &g);
        TDVH_TRY(rc);
        TDVH_CHECK_HIP(ec);
        b->graph = g;
        TDVH_CHECK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    TDVH_CHECK_HIP(hipGraphLaunch(b->exec, s));
    return TDVH_OK;
}

extern "C" int tdvh_batch_wait(tdvh_batch *b, void *stream)
{
    clear_error();
    TDVH_REQUIRE(b, "null batch");
    TDVH_TRY(use_device(b->ctx));
    TDVH_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (b->depth >= 2) {
        uint32_t err = 0;
        TDVH_CHECK_HIP(hipMemcpy(&err, b->d_ctl + CTL_ERR, 4, hipMemcpyDeviceToHost));
        TDVH_REQUIRE(err == 0, "a batch work queue stalled (a wait hit its spin limit)");
    }
    return TDVH_OK;
}

extern "C" int64_t tdvh_batch_cells(const tdvh_batch *b) { return b ? b->cells : 0; }
extern "C" size_t tdvh_batch_scratch_bytes(const tdvh_batch *b) { return b ? b->mem_bytes : 0; }
