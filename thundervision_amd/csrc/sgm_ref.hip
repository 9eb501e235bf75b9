// Semi-global matching in the reference's own path geometry (fidelity mode).
//
// The reference CPU optimiser walks the path table of semiglobal.cu:57-230
// sequentially, each path forward from `start` then backward from `end`
// (semiglobalcpu.cpp:210-222). Its table's end points are off the true line
// ends when W != H, so some pixels are revisited many times per direction and
// some never; at every traversal start it assigns aggreg = C
// (semiglobalcpu.cpp:81), wiping earlier sums there; and its minimum skips
// the last disparity (semiglobalcpu.cpp:95-97), which lets path costs go
// negative. Everything else is the standard recurrence.
//
// Order-free restatement used here: every traversal's path costs depend only
// on C, so all traversals run in parallel (one 16-lane DPP row each, int32
// values, K = ceil(D/16) per lane). With start-assign, pixel p keeps only
// the contributions of traversals t >= t*(p), the last traversal (in table
// order) that starts at p; t* is a per-pixel map built on the host. Integer
// atomic adds make the sum independent of execution order.
#include <vector>

#include "kernels.hpp"

namespace tdvh {

struct RefTables {
    int W = 0, H = 0;
    tdvh_sgpath *paths = nullptr; // device
    int32_t *tstar = nullptr;     // device, W*H, -1 where no traversal starts
    int npaths = 0;
};

static std::unordered_map<const tdvh_ctx *, RefTables> g_ref_tables;
static std::mutex g_ref_mutex;

__device__ __forceinline__ int32_t c_at(const uint32_t *C, size_t p, int d, int m)
{
    const int h = d >= 16 * m;
    const int dd = h ? d - 16 * m : d;
    const uint32_t w = C[p * 16 * m + dd];
    const uint32_t v = h ? (w >> 16) : (w & 0xFFFFu);
    return v == INV16 ? INF_I32 : (int32_t)v;
}

template <int K>
__global__ __launch_bounds__(256) void sgm_ref_kernel(const uint32_t *__restrict__ C,
                                                      const tdvh_sgpath *__restrict__ paths,
                                                      const int32_t *__restrict__ tstar, int ntrav,
                                                      int W, int D, int m, int P1, int P2,
                                                      int flags, int32_t *__restrict__ S)
{
    const int lane = threadIdx.x & 63;
    const int t = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 4 + (lane >> 4);
    const int j = lane & 15;
    if (t >= ntrav) return;
    const tdvh_sgpath ph = paths[t >> 1];
    const bool back = t & 1;
    int x = back ? ph.ex : ph.sx, y = back ? ph.ey : ph.sy;
    const int dx = back ? -ph.dx : ph.dx, dy = back ? -ph.dy : ph.dy;
    const bool assign = flags & TDVH_SGM_START_ASSIGN;
    const int mlimit = (flags & TDVH_SGM_MIN_SKIP_LAST) ? D - 1 : D;

    int32_t Lp[K];
    for (int k = 0; k < ph.size; k++) {
        const size_t p = (size_t)y * W + x;
        const bool take = !assign || t >= tstar[p];
        int32_t Ln[K];
        if (k == 0) {
#pragma unroll
            for (int r = 0; r < K; r++) {
                const int d = j * K + r;
                Ln[r] = d < D ? c_at(C, p, d, m) : INF_I32;
            }
        } else {
            // minimum of the previous vector over d < mlimit
            int32_t mn = INF_I32;
#pragma unroll
            for (int r = 0; r < K; r++)
                if (j * K + r < mlimit) mn = min(mn, Lp[r]);
            mn = row_min_i32(mn);
            // neighbours across the lane boundary (INF outside [0, D))
            int32_t left = (int32_t)__builtin_amdgcn_update_dpp(INF_I32, Lp[K - 1], 0x111, 0xF,
                                                                0xF, false); // row_shr:1
            int32_t right = (int32_t)__builtin_amdgcn_update_dpp(INF_I32, Lp[0], 0x101, 0xF,
                                                                 0xF, false); // row_shl:1
#pragma unroll
            for (int r = 0; r < K; r++) {
                const int d = j * K + r;
                const int32_t c = d < D ? c_at(C, p, d, m) : INF_I32;
                if (c == INF_I32) {
                    Ln[r] = INF_I32;
                    continue;
                }
                const int32_t lm1 = r == 0 ? left : Lp[r - 1];
                const int32_t lp1 = r == K - 1 ? right : Lp[r + 1];
                int64_t best = (int64_t)mn + P2;
                if (Lp[r] != INF_I32) best = min(best, (int64_t)Lp[r]);
                if (lm1 != INF_I32) best = min(best, (int64_t)lm1 + P1);
                if (lp1 != INF_I32) best = min(best, (int64_t)lp1 + P1);
                Ln[r] = (int32_t)((int64_t)c + best - mn);
            }
        }
        if (take) {
#pragma unroll
            for (int r = 0; r < K; r++) {
                const int d = j * K + r;
                if (d < D && Ln[r] != INF_I32) atomicAdd(&S[p * D + d], Ln[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < K; r++) Lp[r] = Ln[r];
        x += dx;
        y += dy;
    }
}

__global__ __launch_bounds__(256) void sgm_ref_init_kernel(const uint32_t *__restrict__ C, int W,
                                                           int H, int D, int m,
                                                           int32_t *__restrict__ S)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)W * H * D) return;
    const int d = (int)(i % D);
    const size_t p = (size_t)(i / D);
    S[i] = c_at(C, p, d, m) == INF_I32 ? INF_I32 : 0;
}

static int ref_tables(tdvh_ctx *ctx, int W, int H, RefTables **out)
{
    std::lock_guard<std::mutex> lock(g_ref_mutex);
    RefTables &t = g_ref_tables[ctx];
    if (t.W == W && t.H == H && t.paths) {
        *out = &t;
        return TDVH_OK;
    }
    if (t.paths) {
        TDVH_CHECK_HIP(hipDeviceSynchronize());
        TDVH_CHECK_HIP(hipFree(t.paths));
        TDVH_CHECK_HIP(hipFree(t.tstar));
        t.paths = nullptr;
        t.tstar = nullptr;
    }
    const size_t np = tdvh_sgm_paths(W, H, nullptr);
    std::vector<tdvh_sgpath> hp(np);
    tdvh_sgm_paths(W, H, hp.data());
    std::vector<int32_t> ts((size_t)W * H, -1);
    for (size_t i = 0; i < np; i++) {
        // traversal 2i starts at `start`, 2i+1 at `end`; keep the last one
        ts[(size_t)hp[i].sy * W + hp[i].sx] = (int32_t)(2 * i);
        ts[(size_t)hp[i].ey * W + hp[i].ex] = (int32_t)(2 * i + 1);
    }
    TDVH_CHECK_HIP(hipMalloc(&t.paths, np * sizeof(tdvh_sgpath)));
    TDVH_CHECK_HIP(hipMalloc(&t.tstar, ts.size() * sizeof(int32_t)));
    TDVH_CHECK_HIP(hipMemcpy(t.paths, hp.data(), np * sizeof(tdvh_sgpath), hipMemcpyHostToDevice));
    TDVH_CHECK_HIP(hipMemcpy(t.tstar, ts.data(), ts.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice));
    t.W = W;
    t.H = H;
    t.npaths = (int)np;
    *out = &t;
    return TDVH_OK;
}

void release_ref_tables(const tdvh_ctx *ctx)
{
    std::lock_guard<std::mutex> lock(g_ref_mutex);
    auto it = g_ref_tables.find(ctx);
    if (it == g_ref_tables.end()) return;
    (void)hipFree(it->second.paths);
    (void)hipFree(it->second.tstar);
    g_ref_tables.erase(it);
}

int run_sgm_ref(tdvh_ctx *ctx, const uint32_t *C, int W, int H, int D, int P1, int P2, int flags,
                int32_t *S, hipStream_t s)
{
    RefTables *t = nullptr;
    TDVH_TRY(ref_tables(ctx, W, H, &t));
    const int m = split_m(D);
    const int64_t cells = (int64_t)W * H * D;
    prof_begin(s);
    sgm_ref_init_kernel<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(C, W, H, D, m, S);
    TDVH_TRY(check_launch("sgm_ref_init_kernel", s));
    const int ntrav = 2 * t->npaths;
    const int blocks = (ntrav + 15) / 16;
    const int K = (D + 15) / 16;
    prof_begin(s);
    switch (K) {
#define CASE(k)                                                                                    \
    case k:                                                                                        \
        sgm_ref_kernel<k><<<blocks, 256, 0, s>>>(C, t->paths, t->tstar, ntrav, W, D, m, P1, P2,    \
                                                 flags, S);                                        \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
        CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: set_error("depth %d above %d", D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("sgm_ref_kernel", s);
}

} // namespace tdvh
