// Semi-global matching, standard geometry, on gfx950.
//
// Recurrence (semiglobalcpu.cpp:87-125 with constant penalties):
//   L(p0, d) = C(p0, d)
//   L(p, d)  = C(p, d) + min(L'(d), L'(d-1) + P1, L'(d+1) + P1, m + P2) - m
// where L' is the previous pixel's vector on the path and m = min_d L'(d).
// Invalid costs (0xFFFF) give invalid L; invalid operands drop out of the
// minimum (saturating uint16 adds keep 0xFFFF + P1 at 0xFFFF).
//
// Mapping: one 16-lane DPP row owns one scan line and holds its disparity
// vector in m packed uint16x2 registers per lane (SPLIT16 order: word j*m+r
// of lane j is d = j*m+r in the low half and d = 16m + j*m+r in the high
// half). A wave scans 4 lines (8 lines of 8 lanes with 2m words each for
// m = 2 on large images); the d-1 / d+1 neighbours are register renames
// except at the lane boundary (one DPP move each way plus a byte-permute
// fix-up on the end lanes), and m is a DPP all-reduce over the line's lanes. Every
// direction's lines are independent, so all 8 directions run in one launch
// and each writes its own path-cost plane; a second kernel sums the planes
// and takes the lowest-index argmin (wtacpu.cpp:20-33).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sgm_device.hpp"

namespace tdvh {

static int env_int(const char *name, int dflt)
{
    const char *e = getenv(name);
    return e ? atoi(e) : dflt;
}

int popcount8(int mask)
{
    int n = 0;
    for (int i = 0; i < 8; i++) n += (mask >> i) & 1;
    return n;
}

// Prefetch depth (steps of cost data in flight per line): 4, 8 or 16.
static int lines_prefetch()
{
    static int pf = [] {
        const char *e = getenv("TDVH_SGM_PREFETCH");
        const int v = e ? atoi(e) : 4;
        return (v == 8 || v == 12 || v == 16) ? v : 4;
    }();
    return pf;
}

// Buffer-addressed line kernel (byte planes below 2^31 bytes) unless
// TDVH_SGM_LINES=flat selects the flat-address kernel.
static bool lines_buf_enabled()
{
    static const bool on = [] {
        const char *e = getenv("TDVH_SGM_LINES");
        return !(e && std::string(e) == "flat");
    }();
    return on;
}

template <int M, bool FUSED>
static void launch_lines_buf(const SgmArgs &a, const uint32_t *order, int groups, int prio,
                             bool lanes8, hipStream_t s)
{
    const bool pad = a.D < 32 * M;
    if constexpr (FUSED && M == 2) {
        if (lanes8) {
            if (pad) sgm_lines_buf_kernel<M, FUSED, true, 8><<<groups, 64, 0, s>>>(a, order, prio);
            else sgm_lines_buf_kernel<M, FUSED, false, 8><<<groups, 64, 0, s>>>(a, order, prio);
            return;
        }
    }
    if (pad) sgm_lines_buf_kernel<M, FUSED, true, 16><<<groups, 64, 0, s>>>(a, order, prio);
    else sgm_lines_buf_kernel<M, FUSED, false, 16><<<groups, 64, 0, s>>>(a, order, prio);
}

// Whether the scans read relative cost bytes (sgm_rel_pixel): automatic,
// or forced by TDVH_SGM_REL=1 (on where applicable) / 0 (off).
//
// Measured (BT + SGM-8 + WTA per pair, fused): the byte scans move 1 B per
// cell and direction from memory instead of recomputing BT on the vector
// ALUs. That pays where the BT step's fixed costs dominate (1 or 2 words
// per lane: D <= 64, 14-20 % faster at 1280x720 and 1920x1080) and where
// the bytes stay in the device's caches (640x480, D = 128..256: 10-15 %
// faster). At 118 MB of cost bytes (1280x720, D = 128) the result varies
// from device to device (-2 % to +7 %), and beyond (1280x720 D = 256,
// 1920x1080 D >= 128) the byte scans are 10-20 % slower.
// Fused (BT terms): the bytes replace the scans' BT work with a read per
// cell and direction, which pays for D <= 64 and up to 96 MiB of bytes.
// A given uint16 volume: the bytes halve the scans' cost reads (2 bytes per
// cell and direction to 1), measured faster at every shape (1280x720 D =
// 128: 0.715 vs 0.729 ms, 1920x1080 D = 256: 3.01 vs 3.50 ms per pair).
static bool rel_wanted(int W, int H, int m, bool volume)
{
    static const int mode = [] {
        const char *e = getenv("TDVH_SGM_REL");
        return e ? atoi(e) : -1;
    }();
    if (mode >= 0) return mode != 0;
    return volume || m <= 2 || (size_t)W * H * 32 * m <= ((size_t)96 << 20);
}

bool sgm_rel_auto(int W, int H, int m, int p2, bool volume)
{
    return p2 <= 63 && (m == 1 || m % 2 == 0) && rel_wanted(W, H, m, volume);
}

template <int M>
static void launch_lines_rel(const SgmArgs &a, const uint32_t *order, int groups, int prio,
                             bool lanes8, hipStream_t s)
{
    if constexpr (rel_ok<M>()) {
        if constexpr (M <= 4) {
            if (lanes8) {
                sgm_lines_rel_kernel<M, 8><<<groups, 64, 0, s>>>(a, order, prio);
                return;
            }
        }
        sgm_lines_rel_kernel<M, 16><<<groups, 64, 0, s>>>(a, order, prio);
    }
}

// 8-lane lines (8 per wave, 4 words per lane) for the fused scans at 2 words
// per lane on large images: measured 12 % faster at 1920x1080 (D = 64), but
// slower where the launch is short and its longest lines set the time
// (640x480: +3 %), and slower at 1 word per lane (1280x720, D = 32: +6 %).
static bool lines_8lane(const SgmArgs &a)
{
    return (a.C == nullptr || a.CR != nullptr) && a.m == 2 && (long)a.W * a.H >= 1536L * 1024;
}

template <int M, bool FUSED, bool EOUT>
static void launch_lines_k(const SgmArgs &a, const LineSet &ls, hipStream_t s)
{
    const int waves = ls.wave_begin[ls.ndirs];
    const int blocks = (waves + 3) / 4;
    switch (lines_prefetch()) {
    case 8: sgm_lines_kernel<M, FUSED, EOUT, 8><<<blocks, 256, 0, s>>>(a, ls); break;
    case 16: sgm_lines_kernel<M, FUSED, EOUT, 16><<<blocks, 256, 0, s>>>(a, ls); break;
    default: sgm_lines_kernel<M, FUSED, EOUT, 4><<<blocks, 256, 0, s>>>(a, ls); break;
    }
}

template <int M>
static void launch_lines_m(const SgmArgs &a, const LineSet &ls, bool eout, hipStream_t s)
{
    const bool fused = a.C == nullptr;
    if (fused && eout) launch_lines_k<M, true, true>(a, ls, s);
    else if (fused) launch_lines_k<M, true, false>(a, ls, s);
    else if (eout) launch_lines_k<M, false, true>(a, ls, s);
    else launch_lines_k<M, false, false>(a, ls, s);
}


static LineSet make_lineset(int W, int H, int dir_mask)
{
    LineSet ls{};
    static const int prio = env_int("TDVH_SGM_PRIO", 1);
    ls.prio = prio;
    ls.wave_begin[0] = 0;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        ls.dir[ls.ndirs] = dir;
        ls.wave_begin[ls.ndirs + 1] = ls.wave_begin[ls.ndirs] + (dir_lines(dir, W, H) + 3) / 4;
        ls.ndirs++;
    }
    return ls;
}

// Spatial coordinate of a line across its direction: lines with close keys
// read neighbouring rows/columns/diagonals of the term planes.
static int line_key(int dir, int sx, int sy)
{
    switch (dir) {
    case 0: case 1: return sy;            // horizontal: row
    case 2: case 3: return sx;            // vertical: column
    case 4: case 5: return sx - sy;       // (1,1) / (-1,-1) diagonals
    default: return sx + sy;              // (-1,1) / (1,-1) anti-diagonals
    }
}

// Chiplets (XCDs) of the device, each with its own L2; workgroup b of a
// launch runs on chiplet b % NXCD.
constexpr int NXCD = 8;

// Line-group table of one pair (lpw consecutive lines of one direction: the
// lines of one wave, 4, or 8 with 8-lane lines) over the enabled
// directions, appended to `out`. Entry: plane slot << 28 | direction << 24
// | first line. Entry k of the launch (k = out's index, counting the
// entries already there) runs on workgroup k / wpg, which the hardware
// places on chiplet (k / wpg) mod NXCD. Two goals:
//  * longest first: the dispatcher then runs a
//    longest-processing-time-first list schedule;
//  * chiplet locality: each direction's groups, in spatial order, are cut
//    into NXCD bands of equal work and band c goes to chiplet c, so the
//    waves sharing rows of the term planes share one L2 (round-robin
//    placement would make every chiplet fetch every row).
// Entry k is the next entry of its chiplet's longest-first queue (another
// chiplet's longest remaining one once that queue is empty).
void line_order_append(int W, int H, int dir_mask, int lpw, int wpg, std::vector<uint32_t> &out)
{
    struct Group {
        int len, key;
        uint32_t entry;
    };
    std::vector<std::pair<int, uint32_t>> q[NXCD]; // (length, entry) per chiplet
    size_t total = 0;
    int slot = 0;
    for (int dir = 0; dir < 8; dir++) {
        if (!(dir_mask & (1 << dir))) continue;
        const int n = dir_lines(dir, W, H);
        std::vector<Group> g;
        long work = 0;
        for (int l0 = 0; l0 < n; l0 += lpw) {
            int kmax = 0, key = 0;
            for (int l = l0; l < l0 + lpw && l < n; l++) {
                int sx, sy, len;
                line_geom(dir, l, W, H, sx, sy, len);
                kmax = len > kmax ? len : kmax;
                if (l == l0) key = line_key(dir, sx, sy);
            }
            g.push_back({kmax, key, (uint32_t)slot << 28 | (uint32_t)dir << 24 | (uint32_t)l0});
            work += kmax;
        }
        std::stable_sort(g.begin(), g.end(),
                         [](const Group &x, const Group &y) { return x.key < y.key; });
        long acc = 0;
        for (const Group &e : g) {
            const long mid = acc + e.len / 2; // band of the group's midpoint
            const int c = (int)std::min<long>(NXCD - 1, mid * NXCD / std::max(work, 1L));
            q[c].push_back({e.len, e.entry});
            acc += e.len;
        }
        total += g.size();
        slot++;
    }
    for (auto &v : q)
        std::stable_sort(v.begin(), v.end(),
                         [](const std::pair<int, uint32_t> &x, const std::pair<int, uint32_t> &y) {
                             return x.first > y.first;
                         });
    size_t head[NXCD] = {};
    for (size_t b = 0; b < total; b++) {
        int c = (int)((out.size() / std::max(wpg, 1)) % NXCD);
        if (head[c] == q[c].size()) {
            // this chiplet's queue is drained: take the longest remaining head
            int best = -1;
            for (int o = 0; o < NXCD; o++)
                if (head[o] < q[o].size() &&
                    (best < 0 || q[o][head[o]].first > q[best][head[best]].first))
                    best = o;
            c = best;
        }
        out.push_back(q[c][head[c]++].second);
    }
}

// The table of line_order_append for one pair and one wave per workgroup,
// built and uploaded once per (W, H, dir_mask, lpw) and context.
static int line_order(tdvh_ctx *ctx, int W, int H, int dir_mask, int lpw, const uint32_t **out,
                      int *groups, hipStream_t s)
{
    TDVH_REQUIRE(W + H < (1 << 24), "image %dx%d too large for the line table", W, H);
    // one table per shape, mask and lines per wave: a context serving several
    // shapes keeps them all, each in its own buffer that no other shape
    // reuses, so the upload is ordered on the launch stream (no device-wide
    // synchronisation) and steady-state calls never rebuild or upload
    char name[64];
    snprintf(name, sizeof name, "sgm_line_order:%dx%d:%02x:%d", W, H, dir_mask, lpw);
    const uint64_t tag = ((uint64_t)W << 40) | ((uint64_t)H << 16) | ((uint64_t)lpw << 8) |
                         (uint64_t)dir_mask;
    std::vector<uint32_t> &host = ctx->host_tables[name];
    tdvh_ctx::Slot &sl0 = ctx->slot(name);
    if (sl0.tag == tag && sl0.ptr) {
        *groups = (int)host.size();
        *out = (const uint32_t *)sl0.ptr;
        return TDVH_OK;
    }
    std::vector<uint32_t> h;
    line_order_append(W, H, dir_mask, lpw, 1, h);
    void *buf = nullptr;
    TDVH_TRY(ctx->scratch(name, h.size() * 4, &buf));
    tdvh_ctx::Slot &sl = ctx->slot(name);
    // the host copy lives in the context for as long as the buffer, so the
    // asynchronous copy's source stays valid
    host = std::move(h);
    TDVH_CHECK_HIP(hipMemcpyAsync(buf, host.data(), host.size() * 4, hipMemcpyHostToDevice, s));
    sl.tag = tag;
    *groups = (int)host.size();
    *out = (const uint32_t *)buf;
    return TDVH_OK;
}

int launch_sgm_lines(tdvh_ctx *ctx, const SgmArgs &a, int dir_mask, bool eout, hipStream_t s)
{
    const LineSet ls = make_lineset(a.W, a.H, dir_mask);
    if (ls.ndirs == 0) return TDVH_OK;
    if (eout && lines_buf_enabled() && buf_lines_fit(a)) {
        const uint32_t *order = nullptr;
        int groups = 0;
        const bool fused = a.C == nullptr;
        const bool rel = a.CR != nullptr;
        const bool lanes8 = lines_8lane(a);
        const int lpw = lanes8 ? 8 : 4; // lines per wave
        TDVH_TRY(line_order(ctx, a.W, a.H, dir_mask, lpw, &order, &groups, s));
        prof_begin(s);
        switch (a.m) {
#define CASE(m)                                                                                    \
    case m:                                                                                        \
        if (rel) launch_lines_rel<m>(a, order, groups, ls.prio, lanes8, s);                        \
        else if (fused) launch_lines_buf<m, true>(a, order, groups, ls.prio, lanes8, s);           \
        else launch_lines_buf<m, false>(a, order, groups, ls.prio, false, s);                      \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default: set_error("depth %d above %d", a.D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
        }
        return check_launch("sgm_lines_kernel", s);
    }
    prof_begin(s);
    switch (a.m) {
#define CASE(m)                                                                                    \
    case m: launch_lines_m<m>(a, ls, eout, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", a.D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_lines_kernel", s);
}

__global__ __launch_bounds__(256) void bt_planes_kernel(const uint8_t *__restrict__ L,
                                                        const uint8_t *__restrict__ R, int W,
                                                        int H, int pitch, int m,
                                                        uint32_t *__restrict__ LT,
                                                        uint32_t *__restrict__ RT, int Wr)
{
    bt_planes_at(L, R, W, H, pitch, m, LT, RT, Wr, blockIdx.x * blockDim.x + threadIdx.x,
                 blockIdx.y);
}

int launch_bt_planes(const uint8_t *L, const uint8_t *R, int W, int H, int pitch, int m,
                     uint32_t *LT, uint32_t *RT, hipStream_t s)
{
    const int Wr = W + 16 * m;
    dim3 grid((Wr + 255) / 256, H);
    prof_begin(s);
    bt_planes_kernel<<<grid, 256, 0, s>>>(L, R, W, H, pitch, m, LT, RT, Wr);
    return check_launch("bt_planes_kernel", s);
}

// Relative cost bytes a.CR of every pixel from the term planes (fused) or
// the SPLIT16 volume a.C.
int launch_sgm_rel(const SgmArgs &a, hipStream_t s)
{
    const bool fused = a.C == nullptr;
    prof_begin(s);
    switch (a.m) {
#define CASE(m)                                                                                    \
    case m: {                                                                                      \
        const long threads = (long)a.W * a.H * sum_group<m>();                                    \
        const int blocks = (int)((threads + 255) / 256);                                           \
        if (fused) sgm_rel_kernel<m, true><<<blocks, 256, 0, s>>>(a);                              \
        else sgm_rel_kernel<m, false><<<blocks, 256, 0, s>>>(a);                                   \
        break;                                                                                     \
    }
        CASE(1) CASE(2) CASE(4) CASE(6) CASE(8)
#undef CASE
    default: set_error("no relative cost bytes at %d words per lane", a.m); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_rel_kernel", s);
}

// Wide-lane plane sum (16 bytes per lane and plane) for power-of-two word
// counts unless TDVH_SGM_SUM=narrow selects one 16-lane row per pixel.
static bool sum_wide_enabled()
{
    static const bool on = [] {
        const char *e = getenv("TDVH_SGM_SUM");
        return !(e && std::string(e) == "narrow");
    }();
    return on;
}

template <int M>
static void launch_sum_m(const SgmArgs &a, int nplanes, bool eout, uint32_t *S, int16_t *disp,
                         int pitch, int blocks, hipStream_t s)
{
    const bool fused = a.C == nullptr;
    if constexpr (sum_wide_ok<M>()) {
        if (eout && sum_wide_enabled()) {
            const long threads = (long)a.W * a.H * sum_group<M>();
            const int wb = (int)((threads + 255) / 256);
            if (fused)
                sgm_sum_wide_kernel<M, true><<<wb, 256, 0, s>>>(a, nplanes, S, disp, pitch);
            else
                sgm_sum_wide_kernel<M, false><<<wb, 256, 0, s>>>(a, nplanes, S, disp, pitch);
            return;
        }
    }
    if (eout && fused)
        sgm_sum_wta_kernel<M, true, true><<<blocks, 256, 0, s>>>(a, nplanes, S, disp, pitch);
    else if (eout)
        sgm_sum_wta_kernel<M, false, true><<<blocks, 256, 0, s>>>(a, nplanes, S, disp, pitch);
    else
        sgm_sum_wta_kernel<M, false, false><<<blocks, 256, 0, s>>>(a, nplanes, S, disp, pitch);
}

int launch_sgm_sum_wta(const SgmArgs &a, int nplanes, bool eout, uint32_t *S_split, int16_t *disp,
                       int disp_pitch, hipStream_t s)
{
    const long pixels = (long)a.W * a.H;
    const int blocks = (int)((pixels + 15) / 16);
    prof_begin(s);
    switch (a.m) {
#define CASE(m)                                                                                    \
    case m: launch_sum_m<m>(a, nplanes, eout, S_split, disp, disp_pitch, blocks, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", a.D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_sum_wta_kernel", s);
}

} // namespace tdvh

using namespace tdvh;

// Penalties of the cost volume's dtype and the direction mask.
static int validate_sgm_params(const tdvh_sgm_params *p, int dtype = TDVH_U16)
{
    TDVH_REQUIRE(p != nullptr, "params is null");
    if (dtype == TDVH_F32) {
        TDVH_REQUIRE(p->fp1 >= 0.0f && p->fp2 >= 0.0f && p->fp1 < __builtin_inff() &&
                         p->fp2 < __builtin_inff(),
                     "float penalties out of range (P1=%g, P2=%g)", (double)p->fp1,
                     (double)p->fp2);
    } else {
        const int lim = dtype == TDVH_U32 ? (1 << 30) : 1024;
        TDVH_REQUIRE(p->p1 >= 0 && p->p2 >= 0 && p->p1 <= lim && p->p2 <= lim,
                     "penalties out of range (P1=%d, P2=%d, limit %d)", p->p1, p->p2, lim);
    }
    TDVH_REQUIRE((p->dir_mask & 0xFF) != 0 && (p->dir_mask & ~0xFF) == 0, "bad dir_mask 0x%x",
                 p->dir_mask);
    return TDVH_OK;
}

// SGM + WTA over a strided uint32 or float volume (sgm_wide.hip).
static int sgm_wide(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *params,
                    tdvh_dsi *S_out, int16_t *disp, int disp_pitch, void *stream)
{
    const int W = C->width, H = C->height, D = C->depth;
    TDVH_REQUIRE(C->layout == TDVH_LAYOUT_STRIDED, "uint32 and float volumes must be strided");
    TDVH_REQUIRE(params->flags == 0, "reference geometry needs a uint16 cost volume");
    if (!S_out && sgm_wide_one_launch(C, params))
        return run_sgm_wide(ctx, C, params, nullptr, disp, disp_pitch, (hipStream_t)stream);
    tdvh_dsi S;
    if (S_out) {
        TDVH_REQUIRE(S_out->ptr && S_out->width == W && S_out->height == H &&
                         S_out->depth == D && S_out->dtype == C->dtype &&
                         S_out->layout == TDVH_LAYOUT_STRIDED,
                     "S_out must be a strided %dx%dx%d volume of the cost dtype", W, H, D);
        S = *S_out;
    } else {
        void *sb = nullptr;
        TDVH_TRY(ctx->scratch("sgm_wide_s", (size_t)W * H * D * 4, &sb));
        TDVH_TRY(tdvh_dsi_init(&S, sb, W, H, D, C->dtype, TDVH_LAYOUT_STRIDED));
    }
    TDVH_TRY(run_sgm_wide(ctx, C, params, &S, nullptr, 0, (hipStream_t)stream));
    return tdvh_wta(ctx, &S, TDVH_WTA_LOWEST, disp, disp_pitch, stream);
}

// Standard-geometry pipeline from a SPLIT16 cost volume (Csplit) or from
// the images (Csplit null: BT recomputed from term planes, no C volume).
static int sgm_std(tdvh_ctx *ctx, const uint32_t *Csplit, const uint8_t *left,
                   const uint8_t *right, int pitch, int W, int H, int D,
                   const tdvh_sgm_params *params, tdvh_dsi *S_out, int16_t *disp, int disp_pitch,
                   hipStream_t s)
{
    const int nplanes = popcount8(params->dir_mask);
    // path-cost planes: L - C in bytes when P2 fits a byte, else L words
    const bool eout = params->p2 <= 255;
    // a SPLIT16 volume (C in, S out, or the staged fallback) fixes the
    // register shape; the fused byte path picks its own
    const int m = (!Csplit && eout && !S_out) ? scan_m(D) : split_m(D);
    SgmArgs a{};
    a.W = W;
    a.H = H;
    a.D = D;
    a.P1 = params->p1;
    a.P2 = params->p2;
    a.C = Csplit;
    a.m = m;
    a.Wr = W + 16 * m;
    a.plane_bytes = (size_t)W * H * 32 * m * (eout ? 1 : 2);
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("sgm_planes", a.plane_bytes * nplanes, &pb));
    a.planes = (uint8_t *)pb;
    void *trash = nullptr;
    TDVH_TRY(ctx->scratch("sgm_trash", 64 * 64, &trash));
    a.trash = (uint8_t *)trash;
    if (!Csplit) {
        void *tb = nullptr;
        const size_t words = (size_t)W * H + 3 * (size_t)a.Wr * H;
        TDVH_TRY(ctx->scratch("bt_planes", words * 4, &tb));
        a.LT = (const uint32_t *)tb;
        a.RT = (const uint32_t *)tb + (size_t)W * H;
        TDVH_TRY(launch_bt_planes(left, right, W, H, pitch, m, (uint32_t *)a.LT,
                                  (uint32_t *)a.RT, s));
        if (!eout) {
            // L words need C in the sum: fall back to a staged volume
            void *cb = nullptr;
            TDVH_TRY(ctx->scratch("cost_c", tdvh_dsi_bytes(W, H, D, TDVH_U16, TDVH_LAYOUT_SPLIT16),
                                  &cb));
            void *tl = nullptr;
            TDVH_TRY(ctx->scratch("bt_terms", (size_t)W * H * 8, &tl));
            uint32_t *TL = (uint32_t *)tl, *TR = TL + (size_t)W * H;
            TDVH_TRY(launch_bt_terms(left, right, W, H, pitch, TL, TR, s));
            TDVH_TRY(launch_bt_split(TL, TR, W, H, D, (uint32_t *)cb, s));
            a.C = (const uint32_t *)cb;
        }
    }
    // relative cost bytes (byte planes, 2 P2 below 128, an even word count,
    // buffer-addressed scans)
    if (eout && sgm_rel_auto(W, H, m, params->p2, Csplit != nullptr) && lines_buf_enabled()) {
        SgmArgs r = a;
        r.CR = reinterpret_cast<uint8_t *>(1); // sized below; checked by buf_lines_fit
        if (buf_lines_fit(r)) {
            void *cr = nullptr;
            TDVH_TRY(ctx->scratch("sgm_rel", (size_t)W * H * 32 * m, &cr));
            a.CR = (uint8_t *)cr;
            TDVH_TRY(launch_sgm_rel(a, s));
        }
    }
    TDVH_TRY(launch_sgm_lines(ctx, a, params->dir_mask, eout, s));
    uint32_t *Ssplit = nullptr;
    if (S_out) {
        TDVH_REQUIRE(S_out->width == W && S_out->height == H && S_out->depth == D &&
                         S_out->dtype == TDVH_U16 && S_out->ptr,
                     "S_out must be a uint16 %dx%dx%d volume", W, H, D);
        if (S_out->layout == TDVH_LAYOUT_SPLIT16)
            Ssplit = (uint32_t *)S_out->ptr;
        else {
            void *sb = nullptr;
            TDVH_TRY(ctx->scratch("sgm_s", (size_t)W * H * 64 * m, &sb));
            Ssplit = (uint32_t *)sb;
        }
    }
    TDVH_TRY(launch_sgm_sum_wta(a, nplanes, eout, Ssplit, disp, disp_pitch, s));
    if (S_out && S_out->layout != TDVH_LAYOUT_SPLIT16)
        TDVH_TRY(launch_convert((uint16_t *)S_out->ptr, S_out->stride_x, S_out->stride_y, Ssplit,
                                W, H, D, false, s));
    return TDVH_OK;
}

extern "C" int tdvh_sgm(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *params,
                        tdvh_dsi *S_out, int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(C && C->ptr && disp, "null pointer argument");
    TDVH_REQUIRE(C->dtype == TDVH_U16 || C->dtype == TDVH_U32 || C->dtype == TDVH_F32,
                 "SGM needs a uint16, uint32 or float cost volume (got dtype %d)", C->dtype);
    TDVH_TRY(validate_sgm_params(params, C->dtype));
    const int W = C->width, H = C->height, D = C->depth;
    TDVH_REQUIRE(W > 0 && H > 0 && D > 0 && D <= MAX_DEPTH, "bad volume %dx%dx%d (D <= %d)", W, H,
                 D, MAX_DEPTH);
    TDVH_REQUIRE(disp_pitch >= W, "disp_pitch %d < width %d", disp_pitch, W);
    if (C->dtype != TDVH_U16) return sgm_wide(ctx, C, params, S_out, disp, disp_pitch, stream);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t *Csplit = (const uint32_t *)C->ptr;
    if (C->layout != TDVH_LAYOUT_SPLIT16) {
        void *cb = nullptr;
        TDVH_TRY(ctx->scratch("sgm_c", tdvh_dsi_bytes(W, H, D, TDVH_U16, TDVH_LAYOUT_SPLIT16), &cb));
        TDVH_TRY(launch_convert((uint16_t *)C->ptr, C->stride_x, C->stride_y, (uint32_t *)cb, W, H,
                                D, true, s));
        Csplit = (const uint32_t *)cb;
    }
    if (params->flags & TDVH_SGM_REF_GEOMETRY) {
        TDVH_REQUIRE(D >= 2, "reference geometry needs D >= 2");
        int32_t *Sref = nullptr;
        if (S_out) {
            TDVH_REQUIRE(S_out->width == W && S_out->height == H && S_out->depth == D &&
                             S_out->dtype == TDVH_I32 && S_out->layout == TDVH_LAYOUT_STRIDED &&
                             S_out->stride_x == D && S_out->stride_y == (int64_t)W * D &&
                             S_out->ptr,
                         "S_out must be a row-major int32 %dx%dx%d volume", W, H, D);
            Sref = (int32_t *)S_out->ptr;
        } else {
            void *sb = nullptr;
            TDVH_TRY(ctx->scratch("sgm_sref", (size_t)W * H * D * 4, &sb));
            Sref = (int32_t *)sb;
        }
        TDVH_TRY(run_sgm_ref(ctx, Csplit, W, H, D, params->p1, params->p2, params->flags, Sref, s));
        tdvh_dsi sd;
        TDVH_TRY(tdvh_dsi_init(&sd, Sref, W, H, D, TDVH_I32, TDVH_LAYOUT_STRIDED));
        return tdvh_wta(ctx, &sd, TDVH_WTA_LOWEST, disp, disp_pitch, stream);
    }
    return sgm_std(ctx, Csplit, nullptr, nullptr, 0, W, H, D, params, S_out, disp, disp_pitch, s);
}

extern "C" int tdvh_stereo_bt_sgm(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                                  int width, int height, int pitch, int depth,
                                  const tdvh_sgm_params *params, int fusion, int16_t *disp,
                                  int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_TRY(validate_sgm_params(params));
    TDVH_REQUIRE(left && right && disp, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && pitch >= width, "bad image %dx%d pitch %d", width,
                 height, pitch);
    TDVH_REQUIRE(depth > 0 && depth <= MAX_DEPTH, "depth %d outside [1, %d]", depth, MAX_DEPTH);
    TDVH_REQUIRE(disp_pitch >= width, "disp_pitch %d < width %d", disp_pitch, width);
    TDVH_REQUIRE(!(params->flags & TDVH_SGM_REF_GEOMETRY) || fusion == TDVH_FUSE_STAGED,
                 "reference geometry runs staged only");
    hipStream_t s = (hipStream_t)stream;
    if (fusion == TDVH_FUSE_COST)
        return sgm_std(ctx, nullptr, left, right, pitch, width, height, depth, params, nullptr,
                       disp, disp_pitch, s);
    TDVH_REQUIRE(fusion == TDVH_FUSE_STAGED, "unknown fusion mode %d", fusion);
    void *tl = nullptr;
    TDVH_TRY(ctx->scratch("bt_terms", (size_t)width * height * 8, &tl));
    uint32_t *TL = (uint32_t *)tl, *TR = TL + (size_t)width * height;
    TDVH_TRY(launch_bt_terms(left, right, width, height, pitch, TL, TR, s));
    void *cb = nullptr;
    TDVH_TRY(ctx->scratch("cost_c", tdvh_dsi_bytes(width, height, depth, TDVH_U16,
                                                   TDVH_LAYOUT_SPLIT16),
                          &cb));
    TDVH_TRY(launch_bt_split(TL, TR, width, height, depth, (uint32_t *)cb, s));
    tdvh_dsi cd;
    TDVH_TRY(tdvh_dsi_init(&cd, cb, width, height, depth, TDVH_U16, TDVH_LAYOUT_SPLIT16));
    return tdvh_sgm(ctx, &cd, params, nullptr, disp, disp_pitch, stream);
}
