// Batched execution: many stereo pairs of uneven size and disparity range
// run as one hipGraph.
//
// The batch is planned once into device work-item lists (one wave per item):
//   BT term rows of every pair                     (one launch for all pairs)
//   4-line groups of one SGM direction of a pair   (one launch per depth class)
//   chunks of pixels: plane sum + WTA              (one launch per depth class)
// Pairs are grouped by m = ceil(D/32), the register shape of the SGM code, so
// each class's kernel is compiled for its own m and keeps its occupancy; the
// line items of a class are ordered longest line first across all its pairs,
// so the longest recurrences start first. Kernel boundaries order the phases;
// the depth classes run as parallel branches of the graph. The arithmetic is
// the per-pair path's (the same device functions), so every pair's
// disparities equal the plain call's.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "sgm_device.hpp"

namespace tdvh {

struct BPair {
    SgmArgs a;
    const uint8_t *L, *R;
    int pitch, m;
    int16_t *disp;
};

struct BItem {
    int pair, kind, a, b;
};

constexpr int SUM_CHUNK = 64; // pixels per sum item

__global__ __launch_bounds__(256) void batch_bt_kernel(const BPair *__restrict__ pairs,
                                                       const BItem *__restrict__ items, int n)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= n) return;
    const BItem it = items[w];
    const BPair &P = pairs[it.pair];
    const int Wr = P.a.Wr;
    for (int y = it.a; y < it.a + it.b; y++)
        for (int i = lane; i < Wr; i += 64)
            bt_planes_at(P.L, P.R, P.a.W, P.a.H, P.pitch, P.m, (uint32_t *)P.a.LT,
                         (uint32_t *)P.a.RT, Wr, i, y);
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

// Buffer-addressed line groups of one depth class (all 8 planes per pair);
// waves_per_group waves per workgroup, two waves per SIMD.
template <int M, bool PAD, int PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
batch_lines_buf_kernel(const BPair *__restrict__ pairs, const BItem *__restrict__ items, int n,
                       int prio)
{
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) +
                                                 (threadIdx.x >> 6));
    if (w >= n) return;
    const BItem it = items[w];
    sgm_line_group_buf<M, true, PAD, PF>(pairs[it.pair].a, it.a, it.b, it.a, prio);
}

// Plane sum + WTA with 16-byte lanes (sum_group lanes per pixel, one
// pass), for word counts 1, 2, 4, 6 and 8. items: one per pair of the
// class, {pair, 3, first wave, waves}, in increasing first wave; a wave finds
// its pair by binary search (wave-uniform, scalar loads).
template <int M>
__global__ __launch_bounds__(256) void batch_sum_wide_kernel(const BPair *__restrict__ pairs,
                                                             const BItem *__restrict__ items,
                                                             int n, int waves)
{
    constexpr int G = sum_group<M>();
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= waves) return;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].a <= w) lo = mid;
        else hi = mid - 1;
    }
    const BItem it = items[lo];
    const BPair &P = pairs[it.pair];
    const int lane = threadIdx.x & 63;
    const long p = (long)(w - it.a) * (64 / G) + lane / G;
    if (p < (long)P.a.W * P.a.H)
        sgm_sum_pixel_wide<M, true>(P.a, 8, nullptr, P.disp, P.a.W, p, lane % G);
}

template <int M>
__global__ __launch_bounds__(256) void batch_sum_kernel(const BPair *__restrict__ pairs,
                                                        const BItem *__restrict__ items, int n)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const BItem it = items[w];
    const BPair &P = pairs[it.pair];
    const int row = (threadIdx.x & 63) >> 4;
    const long npix = (long)P.a.W * P.a.H;
    for (int q = row; q < it.b; q += 4) {
        const long p = (long)it.a + q;
        if (p < npix) sgm_sum_pixel<M, true, true>(P.a, 8, nullptr, P.disp, P.a.W, p);
    }
}

} // namespace tdvh

using namespace tdvh;

struct tdvh_batch {
    tdvh_ctx *ctx = nullptr;
    int n = 0;
    tdvh_sgm_params params{};
    void *mem = nullptr; // all device memory of the batch
    size_t mem_bytes = 0;
    BPair *d_pairs = nullptr;
    BItem *d_items = nullptr;
    // item list ranges: BT rows, then per depth class m (1..8) lines and sums
    int bt_begin = 0, bt_count = 0;
    // segments: per depth class, consecutive groups of its pairs; a
    // segment's plane sums overlap the next segment's line scans
    struct Seg {
        int mc;
        int line_begin, line_count; // line-group items, longest first
        int sum_begin, sum_count;   // sum items
        int sum_waves;              // waves of the wide sum kernel
    };
    std::vector<Seg> segs;
    bool class_pad[9] = {}; // some pair of the class has padding cells (D < 32m)
    bool class_buf[9] = {}; // every pair of the class fits the buffer line kernel
    int64_t cells = 0;
    // knobs
    int waves_per_group = 1, prefetch = 8, prio = 1;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    hipStream_t branch[9] = {}, sum_branch[9] = {};
    hipEvent_t fork = nullptr, join[9] = {}, sum_join[9] = {};
    std::vector<hipEvent_t> seg_done; // per segment: its line scans finished
};

// Segments per depth class (TDVH_BATCH_SEGMENTS, default 1). Measured on the
// 64-pair mix: overlapping a segment's plane sums with the next segment's
// line scans (4 segments) ran ~10 % slower than one segment per class, as
// both phases then compete for the same memory pipeline and HBM.
static int batch_segments()
{
    static const int n = [] {
        const char *e = getenv("TDVH_BATCH_SEGMENTS");
        const int v = e ? atoi(e) : 1;
        return v < 1 ? 1 : (v > 16 ? 16 : v);
    }();
    return n;
}

static void drop_graph(tdvh_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
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
    std::vector<size_t> off_terms(n), off_planes(n);
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
        cells += (int64_t)W * H * p.depth;
    }
    const size_t trash_off = total;
    total = align(total + 64 * 64);
    // items: BT rows of all pairs, then per depth class the line groups
    // (longest line first) and the pixel chunks
    std::vector<BItem> items;
    const int rows_per_item = 4;
    for (int i = 0; i < n; i++)
        for (int y = 0; y < pairs[i].height; y += rows_per_item)
            items.push_back({i, 0, y, std::min(rows_per_item, pairs[i].height - y)});
    tdvh_batch plan;
    plan.bt_begin = 0;
    plan.bt_count = (int)items.size();
    for (int mc = 1; mc <= 8; mc++) {
        std::vector<int> cls;
        int64_t cls_cells = 0;
        for (int i = 0; i < n; i++)
            if (scan_m(pairs[i].depth) == mc) {
                cls.push_back(i);
                cls_cells += (int64_t)pairs[i].width * pairs[i].height * pairs[i].depth;
            }
        if (cls.empty()) continue;
        // K consecutive groups of the class's pairs with about equal cells:
        // group k ends where the running total reaches k+1 K-ths, leaving at
        // least one pair to each later group
        const int K = std::min((int)cls.size(), batch_segments());
        size_t c0 = 0;
        int64_t acc = 0;
        for (int k = 0; k < K; k++) {
            size_t c1 = c0;
            const int64_t goal = cls_cells * (k + 1) / K;
            const size_t max_end = cls.size() - (size_t)(K - k - 1);
            do {
                const tdvh_pair &q = pairs[cls[c1]];
                acc += (int64_t)q.width * q.height * q.depth;
                c1++;
            } while (c1 < max_end && acc < goal);
            tdvh_batch::Seg sg{mc, 0, 0, 0, 0, 0};
            std::vector<std::pair<int, BItem>> lines;
            for (size_t c = c0; c < c1; c++) {
                const int i = cls[c];
                const int W = pairs[i].width, H = pairs[i].height;
                for (int dir = 0; dir < 8; dir++) {
                    const int nl = dir_lines(dir, W, H);
                    for (int l0 = 0; l0 < nl; l0 += 4) {
                        int longest = 0;
                        for (int l = l0; l < std::min(l0 + 4, nl); l++) {
                            int sx, sy, len;
                            line_geom(dir, l, W, H, sx, sy, len);
                            longest = std::max(longest, len);
                        }
                        lines.push_back({longest, BItem{i, 1, dir, l0}});
                    }
                }
            }
            // pair-major, longest line first within a pair: the next pair's
            // long lines start while this pair's short ones finish, and only
            // about one pair's term planes are live in the caches at a time
            std::stable_sort(lines.begin(), lines.end(),
                             [](const std::pair<int, BItem> &x, const std::pair<int, BItem> &y) {
                                 if (x.second.pair != y.second.pair)
                                     return x.second.pair < y.second.pair;
                                 return x.first > y.first;
                             });
            sg.line_begin = (int)items.size();
            for (auto &l : lines) items.push_back(l.second);
            sg.line_count = (int)items.size() - sg.line_begin;
            sg.sum_begin = (int)items.size();
            if (mc == 1 || mc == 2 || mc == 4 || mc == 6 || mc == 8) {
                // one item per pair: {pair, 3, first wave, waves} (wide sum)
                int wave = 0;
                for (size_t c = c0; c < c1; c++) {
                    const int i = cls[c];
                    const long npix = (long)pairs[i].width * pairs[i].height;
                    const int nw = (int)((npix * (mc == 6 ? 16 : 2 * mc) + 63) / 64);
                    items.push_back({i, 3, wave, nw});
                    wave += nw;
                }
                sg.sum_waves = wave;
            } else {
                for (size_t c = c0; c < c1; c++) {
                    const int i = cls[c];
                    const long npix = (long)pairs[i].width * pairs[i].height;
                    for (long p = 0; p < npix; p += SUM_CHUNK)
                        items.push_back({i, 2, (int)p, (int)std::min<long>(SUM_CHUNK, npix - p)});
                }
            }
            sg.sum_count = (int)items.size() - sg.sum_begin;
            plan.segs.push_back(sg);
            c0 = c1;
        }
    }
    const size_t pairs_off = total;
    total = align(total + sizeof(BPair) * n);
    const size_t items_off = total;
    total = align(total + sizeof(BItem) * items.size());

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
        q.L = p.left;
        q.R = p.right;
        q.pitch = p.pitch;
        q.m = m;
        q.disp = p.disp;
    }
    b->ctx = ctx;
    b->n = n;
    b->params = *params;
    b->mem_bytes = total;
    b->d_pairs = (BPair *)(base + pairs_off);
    b->d_items = (BItem *)(base + items_off);
    b->bt_begin = plan.bt_begin;
    b->bt_count = plan.bt_count;
    b->segs = plan.segs;
    for (int mc = 1; mc <= 8; mc++) b->class_buf[mc] = true;
    for (int i = 0; i < n; i++) {
        const int mc = bp[i].m;
        if (bp[i].a.D < 32 * mc) b->class_pad[mc] = true;
        if (!buf_lines_fit(bp[i].a)) b->class_buf[mc] = false;
    }
    b->cells = cells;
    if (hipMemcpy(b->d_pairs, bp.data(), sizeof(BPair) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_items, items.data(), sizeof(BItem) * items.size(),
                  hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(b->mem);
        delete b;
        set_error("uploading the batch plan failed");
        return TDVH_ERR_DEVICE;
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
    for (int mc = 0; mc <= 8; mc++) {
        if (b->branch[mc]) (void)hipStreamDestroy(b->branch[mc]);
        if (b->sum_branch[mc]) (void)hipStreamDestroy(b->sum_branch[mc]);
        if (b->join[mc]) (void)hipEventDestroy(b->join[mc]);
        if (b->sum_join[mc]) (void)hipEventDestroy(b->sum_join[mc]);
    }
    for (hipEvent_t e : b->seg_done)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    (void)hipFree(b->mem);
    delete b;
    return TDVH_OK;
}

extern "C" int tdvh_batch_config(tdvh_batch *b, int lines_per_wave, int waves_per_group,
                                 int prefetch)
{
    clear_error();
    TDVH_REQUIRE(b, "null batch");
    TDVH_REQUIRE(lines_per_wave == 0 || lines_per_wave == 4, "lines per wave is 4");
    TDVH_REQUIRE(waves_per_group == 0 || waves_per_group == 1 || waves_per_group == 2 ||
                     waves_per_group == 4,
                 "waves per group must be 1, 2 or 4");
    TDVH_REQUIRE(prefetch == 0 || prefetch == 4 || prefetch == 8, "prefetch must be 4 or 8");
    const int wpg = waves_per_group ? waves_per_group : b->waves_per_group;
    const int pf = prefetch ? prefetch : b->prefetch;
    if (wpg != b->waves_per_group || pf != b->prefetch) drop_graph(b);
    b->waves_per_group = wpg;
    b->prefetch = pf;
    return TDVH_OK;
}

template <int M>
static void launch_lines_seg(tdvh_batch *b, const tdvh_batch::Seg &sg, hipStream_t s)
{
    const int threads = 64 * b->waves_per_group, wpb = b->waves_per_group;
    const int nl = sg.line_count;
    if (!nl) return;
    const BItem *li = b->d_items + sg.line_begin;
    const int blocks = (nl + wpb - 1) / wpb;
    prof_begin(s);
    if (b->class_buf[M]) {
        // "prefetch 8": the deepest ring that keeps two waves per SIMD
        constexpr int PF8 = buf_lines_pf<M, true>();
        if (b->class_pad[M]) {
            if (b->prefetch == 8)
                batch_lines_buf_kernel<M, true, PF8><<<blocks, threads, 0, s>>>(b->d_pairs, li, nl,
                                                                                b->prio);
            else
                batch_lines_buf_kernel<M, true, 4><<<blocks, threads, 0, s>>>(b->d_pairs, li, nl,
                                                                              b->prio);
        } else {
            if (b->prefetch == 8)
                batch_lines_buf_kernel<M, false, PF8><<<blocks, threads, 0, s>>>(b->d_pairs, li,
                                                                                 nl, b->prio);
            else
                batch_lines_buf_kernel<M, false, 4><<<blocks, threads, 0, s>>>(b->d_pairs, li, nl,
                                                                               b->prio);
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
static void launch_sum_seg(tdvh_batch *b, const tdvh_batch::Seg &sg, hipStream_t s)
{
    const int ns = sg.sum_count;
    if (!ns) return;
    const BItem *si = b->d_items + sg.sum_begin;
    prof_begin(s);
    if constexpr (sum_wide_ok<M>())
        batch_sum_wide_kernel<M><<<(sg.sum_waves + 3) / 4, 256, 0, s>>>(b->d_pairs, si, ns,
                                                                        sg.sum_waves);
    else
        batch_sum_kernel<M><<<(ns + 3) / 4, 256, 0, s>>>(b->d_pairs, si, ns);
    (void)check_launch("batch_sum_kernel", s);
}

static int launch_seg(tdvh_batch *b, const tdvh_batch::Seg &sg, bool sum, hipStream_t s)
{
    switch (sg.mc) {
#define CASE(m)                                                                                    \
    case m:                                                                                        \
        if (sum) launch_sum_seg<m>(b, sg, s);                                                      \
        else launch_lines_seg<m>(b, sg, s);                                                        \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: break;
    }
    return check_launch(sum ? "batch_sum" : "batch_lines", s);
}

// Enqueues the batch on `s`. With fork (graph capture) every depth class
// runs on two branch streams forked from and joined back into `s`: one scans
// the lines of its segments in order, the other sums each segment as soon as
// its lines are done, overlapping the next segment's scans (VALU-bound) with
// this segment's plane reads (HBM-bound).
static int enqueue_batch(tdvh_batch *b, hipStream_t s, bool fork)
{
    prof_begin(s);
    batch_bt_kernel<<<(b->bt_count + 3) / 4, 256, 0, s>>>(b->d_pairs, b->d_items + b->bt_begin,
                                                          b->bt_count);
    TDVH_TRY(check_launch("batch_bt_kernel", s));
    if (!fork) {
        for (const auto &sg : b->segs) {
            TDVH_TRY(launch_seg(b, sg, false, s));
            TDVH_TRY(launch_seg(b, sg, true, s));
        }
        return TDVH_OK;
    }
    if (!b->fork) TDVH_CHECK_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    if (b->seg_done.size() < b->segs.size()) b->seg_done.resize(b->segs.size(), nullptr);
    TDVH_CHECK_HIP(hipEventRecord(b->fork, s));
    for (int mc = 1; mc <= 8; mc++) {
        bool any = false;
        for (const auto &sg : b->segs) any |= sg.mc == mc;
        if (!any) continue;
        if (!b->branch[mc])
            TDVH_CHECK_HIP(hipStreamCreateWithFlags(&b->branch[mc], hipStreamNonBlocking));
        if (!b->sum_branch[mc])
            TDVH_CHECK_HIP(hipStreamCreateWithFlags(&b->sum_branch[mc], hipStreamNonBlocking));
        if (!b->join[mc])
            TDVH_CHECK_HIP(hipEventCreateWithFlags(&b->join[mc], hipEventDisableTiming));
        if (!b->sum_join[mc])
            TDVH_CHECK_HIP(hipEventCreateWithFlags(&b->sum_join[mc], hipEventDisableTiming));
        hipStream_t ls = b->branch[mc], ss = b->sum_branch[mc];
        TDVH_CHECK_HIP(hipStreamWaitEvent(ls, b->fork, 0));
        TDVH_CHECK_HIP(hipStreamWaitEvent(ss, b->fork, 0));
        for (size_t k = 0; k < b->segs.size(); k++) {
            const auto &sg = b->segs[k];
            if (sg.mc != mc) continue;
            if (!b->seg_done[k])
                TDVH_CHECK_HIP(hipEventCreateWithFlags(&b->seg_done[k], hipEventDisableTiming));
            TDVH_TRY(launch_seg(b, sg, false, ls));
            TDVH_CHECK_HIP(hipEventRecord(b->seg_done[k], ls));
            TDVH_CHECK_HIP(hipStreamWaitEvent(ss, b->seg_done[k], 0));
            TDVH_TRY(launch_seg(b, sg, true, ss));
        }
        TDVH_CHECK_HIP(hipEventRecord(b->join[mc], ls));
        TDVH_CHECK_HIP(hipEventRecord(b->sum_join[mc], ss));
        TDVH_CHECK_HIP(hipStreamWaitEvent(s, b->join[mc], 0));
        TDVH_CHECK_HIP(hipStreamWaitEvent(s, b->sum_join[mc], 0));
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
        const int rc = enqueue_batch(b, cs, true);
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
    return TDVH_OK;
}

extern "C" int64_t tdvh_batch_cells(const tdvh_batch *b) { return b ? b->cells : 0; }
extern "C" size_t tdvh_batch_scratch_bytes(const tdvh_batch *b) { return b ? b->mem_bytes : 0; }
