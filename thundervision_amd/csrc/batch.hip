// Batched execution: many stereo pairs of uneven size and disparity range
// in one persistent launch fed by a device work queue.
//
// Work items (one wave each), queued in this order:
//   phase 0  BT term rows of a pair              (no dependency)
//   phase 1  4 lines of one SGM direction        (needs all phase-0 items of its pair)
//            ordered longest line first across all pairs, so the longest
//            recurrences start first and the launch has a short tail
//   phase 2  a chunk of pixels: plane sum + WTA  (needs all phase-1 items of its pair)
// A wave takes the next item with one atomic add on the queue head. Items
// are taken in queue order, so an item only ever waits for items already
// taken by running waves: the waits always end. Each wait is still bounded
// (a spin limit sets an error flag that tdvh_batch_wait reports).
// Completion counters are published with a device-scope release fence and
// read with acquire loads. The arithmetic is the per-pair path's (the same
// device functions), so every pair's disparities equal the plain call's.
#include <algorithm>
#include <vector>

#include "sgm_device.hpp"

namespace tdvh {

struct BPair {
    SgmArgs a;
    const uint8_t *L, *R;
    int pitch, m;
    int16_t *disp;
    int need[3]; // items per phase
};

struct BItem {
    int pair, kind, a, b;
};

constexpr int BCTL_HEADER = 16; // head, error, spare
constexpr int SUM_CHUNK = 64;   // pixels per phase-2 item
constexpr uint32_t SPIN_LIMIT = 1u << 26;

__device__ __forceinline__ uint32_t load_acquire(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int M, int PF>
__device__ __forceinline__ void run_item_m(const BPair &P, const BItem &it, int prio)
{
    if (it.kind == 1) {
        sgm_line_group<M, true, true, PF>(P.a, it.a, it.b, it.a, prio);
    } else {
        const int j = threadIdx.x & 15, row = (threadIdx.x & 63) >> 4;
        const long npix = (long)P.a.W * P.a.H;
        for (int q = row; q < it.b; q += 4) {
            const long p = (long)it.a + q;
            if (p < npix) sgm_sum_pixel<M, true, true>(P.a, 8, nullptr, P.disp, P.a.W, p);
        }
        (void)j;
    }
}

template <int PF>
__global__ __launch_bounds__(256) void batch_kernel(const BPair *__restrict__ pairs,
                                                    const BItem *__restrict__ items, int nitems,
                                                    int npairs, uint32_t *ctl, int prio)
{
    const int lane = threadIdx.x & 63;
    uint32_t *head = ctl, *err = ctl + 1, *done = ctl + BCTL_HEADER;
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(head, 1u);
        idx = __builtin_amdgcn_readfirstlane(idx);
        if (idx >= (uint32_t)nitems) break;
        const BItem it = items[idx];
        const BPair &P = pairs[it.pair];
        if (it.kind > 0) {
            const uint32_t *dep = done + (it.kind - 1) * npairs + it.pair;
            const uint32_t need = (uint32_t)P.need[it.kind - 1];
            uint32_t spins = 0;
            bool ok = true;
            while (__builtin_amdgcn_readfirstlane(load_acquire(dep)) < need) {
                if (++spins > SPIN_LIMIT) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            if (!ok) {
                if (lane == 0) atomicOr(err, 1u);
                return;
            }
        }
        if (it.kind == 0) {
            const int Wr = P.a.Wr;
            for (int y = it.a; y < it.a + it.b; y++)
                for (int i = lane; i < Wr; i += 64)
                    bt_planes_at(P.L, P.R, P.a.W, P.a.H, P.pitch, P.m, (uint32_t *)P.a.LT,
                                 (uint32_t *)P.a.RT, Wr, i, y);
        } else {
            switch (P.m) {
            case 1: run_item_m<1, PF>(P, it, prio); break;
            case 2: run_item_m<2, PF>(P, it, prio); break;
            case 3: run_item_m<3, PF>(P, it, prio); break;
            case 4: run_item_m<4, PF>(P, it, prio); break;
            case 5: run_item_m<5, PF>(P, it, prio); break;
            case 6: run_item_m<6, PF>(P, it, prio); break;
            case 7: run_item_m<7, PF>(P, it, prio); break;
            default: run_item_m<8, PF>(P, it, prio); break;
            }
        }
        __threadfence(); // publish this item's results (device-scope release)
        if (lane == 0) atomicAdd(done + it.kind * npairs + it.pair, 1u);
        __builtin_amdgcn_s_setprio(0);
    }
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
    tdvh_sgm_params params{};
    void *mem = nullptr;     // all device memory of the batch
    size_t mem_bytes = 0;
    BPair *d_pairs = nullptr;
    BItem *d_items = nullptr;
    uint32_t *d_ctl = nullptr;
    int nitems = 0, nctl = 0;
    int64_t cells = 0;
    // knobs
    int waves_per_group = 4, prefetch = 4, prio = 1;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

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
    TDVH_REQUIRE(params->p1 >= 0 && params->p2 >= 0 && params->p2 <= 255,
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
        const int m = split_m(p.depth), Wr = p.width + 16 * m;
        const size_t W = p.width, H = p.height;
        off_terms[i] = total;
        total = align(total + (3 * W * H + 3 * (size_t)Wr * H) * 4);
        off_planes[i] = total;
        total = align(total + 8 * W * H * 32 * m);
        cells += (int64_t)W * H * p.depth;
    }
    const size_t trash_off = total;
    total = align(total + 64 * 64);
    // items
    std::vector<BItem> items;
    std::vector<std::pair<int, BItem>> lines; // (length, item)
    std::vector<int> need0(n, 0), need1(n, 0), need2(n, 0);
    for (int i = 0; i < n; i++) {
        const int W = pairs[i].width, H = pairs[i].height;
        const int rows_per_item = 4;
        for (int y = 0; y < H; y += rows_per_item) {
            items.push_back({i, 0, y, std::min(rows_per_item, H - y)});
            need0[i]++;
        }
        for (int dir = 0; dir < 8; dir++) {
            const int nl = dir_lines(dir, W, H);
            for (int l0 = 0; l0 < nl; l0 += 4) {
                // longest line of the group
                int longest = 0;
                for (int l = l0; l < std::min(l0 + 4, nl); l++) {
                    int sx, sy, len;
                    line_geom(dir, l, W, H, sx, sy, len);
                    longest = std::max(longest, len);
                }
                lines.push_back({longest, BItem{i, 1, dir, l0}});
                need1[i]++;
            }
        }
    }
    std::stable_sort(lines.begin(), lines.end(),
                     [](const std::pair<int, BItem> &x, const std::pair<int, BItem> &y) {
                         return x.first > y.first;
                     });
    for (auto &l : lines) items.push_back(l.second);
    // sums: pairs with the shortest longest-line first (their lines end first)
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return std::max(pairs[x].width, pairs[x].height) < std::max(pairs[y].width, pairs[y].height);
    });
    for (int i : order) {
        const long npix = (long)pairs[i].width * pairs[i].height;
        for (long p = 0; p < npix; p += SUM_CHUNK) {
            items.push_back({i, 2, (int)p, (int)std::min<long>(SUM_CHUNK, npix - p)});
            need2[i]++;
        }
    }
    const size_t pairs_off = total;
    total = align(total + sizeof(BPair) * n);
    const size_t items_off = total;
    total = align(total + sizeof(BItem) * items.size());
    const size_t ctl_off = total;
    const int nctl = BCTL_HEADER + 3 * n;
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
        const int m = split_m(p.depth);
        q.a = SgmArgs{};
        q.a.W = p.width;
        q.a.H = p.height;
        q.a.D = p.depth;
        q.a.P1 = params->p1;
        q.a.P2 = params->p2;
        q.a.C = nullptr;
        q.a.Wr = p.width + 16 * m;
        q.a.LT = (const uint32_t *)(base + off_terms[i]);
        q.a.RT = q.a.LT + 3 * (size_t)p.width * p.height;
        q.a.planes = base + off_planes[i];
        q.a.plane_bytes = (size_t)p.width * p.height * 32 * m;
        q.a.trash = base + trash_off;
        q.L = p.left;
        q.R = p.right;
        q.pitch = p.pitch;
        q.m = m;
        q.disp = p.disp;
        q.need[0] = need0[i];
        q.need[1] = need1[i];
        q.need[2] = need2[i];
    }
    b->ctx = ctx;
    b->n = n;
    b->host_pairs.assign(pairs, pairs + n);
    b->params = *params;
    b->mem_bytes = total;
    b->d_pairs = (BPair *)(base + pairs_off);
    b->d_items = (BItem *)(base + items_off);
    b->d_ctl = (uint32_t *)(base + ctl_off);
    b->nitems = (int)items.size();
    b->nctl = nctl;
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

static int enqueue_batch(tdvh_batch *b, hipStream_t s)
{
    batch_reset_kernel<<<1, 256, 0, s>>>(b->d_ctl, b->nctl);
    TDVH_TRY(check_launch("batch_reset_kernel", s));
    const int threads = 64 * b->waves_per_group;
    int per_cu = 0;
    hipError_t e;
    if (b->prefetch == 8)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<8>, threads, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<4>, threads, 0);
    if (e != hipSuccess || per_cu <= 0) per_cu = 1;
    const int blocks = per_cu * b->ctx->num_cus;
    prof_begin(s);
    if (b->prefetch == 8)
        batch_kernel<8><<<blocks, threads, 0, s>>>(b->d_pairs, b->d_items, b->nitems, b->n,
                                                   b->d_ctl, b->prio);
    else
        batch_kernel<4><<<blocks, threads, 0, s>>>(b->d_pairs, b->d_items, b->nitems, b->n,
                                                   b->d_ctl, b->prio);
    return check_launch("batch_kernel", s);
}

extern "C" int tdvh_batch_run(tdvh_batch *b, int use_graph, void *stream)
{
    clear_error();
    TDVH_REQUIRE(b, "null batch");
    TDVH_TRY(use_device(b->ctx));
    hipStream_t s = (hipStream_t)stream;
    if (!use_graph) return enqueue_batch(b, s);
    if (!b->exec) {
        TDVH_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch(b, s);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(s, &g);
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
    uint32_t ctl[2] = {0, 0};
    TDVH_CHECK_HIP(hipMemcpy(ctl, b->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    if (ctl[1]) {
        set_error("batch work queue stalled (a dependency wait hit its spin limit)");
        return TDVH_ERR_DEVICE;
    }
    return TDVH_OK;
}

extern "C" int64_t tdvh_batch_cells(const tdvh_batch *b) { return b ? b->cells : 0; }
extern "C" size_t tdvh_batch_scratch_bytes(const tdvh_batch *b) { return b ? b->mem_bytes : 0; }
