// Shared host/device helpers of the tdvh library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "tdvh.h"

namespace tdvh {

// ---- errors -------------------------------------------------------------
void set_error(const char *fmt, ...);
void clear_error();

#define TDVH_CHECK_HIP(expr)                                                   \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::tdvh::set_error("%s failed: %s (%s:%d)", #expr,                  \
                              hipGetErrorString(e_), __FILE__, __LINE__);      \
            return TDVH_ERR_DEVICE;                                            \
        }                                                                      \
    } while (0)

#define TDVH_REQUIRE(cond, ...)                                                \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::tdvh::set_error(__VA_ARGS__);                                    \
            return TDVH_ERR_INVALID_ARG;                                       \
        }                                                                      \
    } while (0)

#define TDVH_TRY(expr)                                                         \
    do {                                                                       \
        int rc_ = (expr);                                                      \
        if (rc_ != TDVH_OK) return rc_;                                        \
    } while (0)

// Profiling hook around a kernel launch: prof_begin before it, check_launch
// (which also records the end event) after it.
void prof_begin(hipStream_t s);
int check_launch(const char *what, hipStream_t s);

// ---- constants ----------------------------------------------------------
constexpr uint16_t INV16 = 0xFFFFu;
constexpr uint32_t INV32 = 0xFFFFFFFFu;
constexpr int32_t INF_I32 = 0x3FFFFFFF;
constexpr int MAX_DEPTH = 256;     // SGM/WTA depth limit of the split layout
constexpr int SPLIT_LANES = 16;    // lanes per disparity vector (one DPP row)

// words per pixel of the SPLIT16 layout: 16 * m, m = ceil(D / 32)
inline int split_m(int depth) { return (depth + 31) / 32; }
inline int split_words(int depth) { return 16 * split_m(depth); }
// Words per lane of the fused line scans and plane sums, where no SPLIT16
// volume fixes the shape: odd counts above 1 run one word wider (padding
// cells d >= D), which the scans and sums handle faster than the odd shape
// (measured on 640x480 and 1280x720 pairs).
inline int scan_m(int depth)
{
    const int m = split_m(depth);
    return (m == 3 || m == 5 || m == 7) ? m + 1 : m;
}

// Direction index -> (dx, dy): 0 (1,0) 1 (-1,0) 2 (0,1) 3 (0,-1) 4 (1,1)
// 5 (-1,-1) 6 (-1,1) 7 (1,-1), the order of the oracle and of tdvh_sgm_params.
__host__ __device__ inline int dir_dx(int dir) { return ((0x62 >> dir) & 1) ? -1 : ((0x91 >> dir) & 1); }
__host__ __device__ inline int dir_dy(int dir) { return ((0xA8 >> dir) & 1) ? -1 : ((0x54 >> dir) & 1); }

// Lines of direction `dir` in standard geometry.
__host__ __device__ inline int dir_lines(int dir, int W, int H)
{
    return dir < 2 ? H : (dir < 4 ? W : W + H - 1);
}

size_t dtype_size(int dtype);

} // namespace tdvh

// ---- context ---------------------------------------------------------------
struct tdvh_ctx {
    int device = 0;
    int num_cus = 0;
    // Grow-only device scratch slots, keyed by purpose.
    struct Slot {
        void *ptr = nullptr;
        size_t bytes = 0;
        uint64_t tag = 0; // what the contents were built for (0: nothing)
    };
    std::unordered_map<std::string, Slot> slots;
    // host copies of tables uploaded asynchronously into slots (kept as
    // long as the context, so an in-flight copy never loses its source)
    std::unordered_map<std::string, std::vector<uint32_t>> host_tables;

    // stage timing
    struct ProfMark {
        const char *name;
        hipEvent_t start, stop;
    };
    bool prof_on = false;
    hipEvent_t prof_pending = nullptr; // start event of the launch in flight
    std::vector<ProfMark> prof_marks;  // recorded, not yet resolved
    std::vector<hipEvent_t> prof_pool; // free events
    std::vector<std::pair<std::string, std::pair<double, int64_t>>> prof_table;
    int prof_resolve();

    // Returns a device buffer of at least `bytes` for `name` (reused).
    int scratch(const char *name, size_t bytes, void **out);
    // The slot itself (after scratch()), to read or set its tag.
    Slot &slot(const char *name) { return slots[name]; }
    int release_all();
    size_t held() const;
};

namespace tdvh {
// Makes the context's device current for the calling thread.
int use_device(tdvh_ctx *ctx);
// Entry points without a context: the launch hooks record nothing (no
// profiler events are created).
void use_no_context();
} // namespace tdvh

// ---- device helpers ----------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
typedef unsigned short tdvh_u16x2 __attribute__((ext_vector_type(2)));

namespace tdvh {

__device__ __forceinline__ tdvh_u16x2 v2(uint32_t v) { return __builtin_bit_cast(tdvh_u16x2, v); }
__device__ __forceinline__ uint32_t u32(tdvh_u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// packed uint16 min / saturating add / wrapping subtract
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b)
{
    return u32(__builtin_elementwise_min(v2(a), v2(b)));
}
__device__ __forceinline__ uint32_t pk_adds(uint32_t a, uint32_t b)
{
    return u32(__builtin_elementwise_add_sat(v2(a), v2(b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) { return u32(v2(a) + v2(b)); }
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) { return u32(v2(a) - v2(b)); }
__device__ __forceinline__ uint32_t pk_splat(uint32_t v) { return (v & 0xFFFFu) * 0x00010001u; }

// Minimum over the 16 lanes of a DPP row; every lane receives it.
__device__ __forceinline__ uint32_t row_min_u32(uint32_t v)
{
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false)); // ror 8
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false)); // ror 4
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false)); // ror 2
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false)); // ror 1
    return v;
}

// Signed minimum over the 16 lanes of a DPP row (biased to unsigned order).
__device__ __forceinline__ int32_t row_min_i32(int32_t v)
{
    return (int32_t)(row_min_u32((uint32_t)v ^ 0x80000000u) ^ 0x80000000u);
}

// Raw buffer resources: 32-bit byte offsets from a base, range-checked
// against a byte count (loads at or past it return 0, stores there are
// dropped).
typedef __amdgpu_buffer_rsrc_t buf_t;
typedef unsigned int tdvh_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int tdvh_u32x3 __attribute__((ext_vector_type(3)));
typedef unsigned int tdvh_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ buf_t buf_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             0x00020000);
}

// N consecutive dwords at byte offset off (0 outside the buffer); POL is
// the cache-policy operand (2: non-temporal, for data read once).
template <int N, int I = 0, int POL = 0>
__device__ __forceinline__ void buf_load(buf_t r, uint32_t off, uint32_t *dst)
{
    if constexpr (N - I >= 4) {
        const tdvh_u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 4 * I, 0, POL);
        dst[I] = v.x, dst[I + 1] = v.y, dst[I + 2] = v.z, dst[I + 3] = v.w;
        buf_load<N, I + 4, POL>(r, off, dst);
    } else if constexpr (N - I == 3) {
        const tdvh_u32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, off + 4 * I, 0, POL);
        dst[I] = v.x, dst[I + 1] = v.y, dst[I + 2] = v.z;
    } else if constexpr (N - I == 2) {
        const tdvh_u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off + 4 * I, 0, POL);
        dst[I] = v.x, dst[I + 1] = v.y;
    } else if constexpr (N - I == 1) {
        dst[I] = __builtin_amdgcn_raw_buffer_load_b32(r, off + 4 * I, 0, POL);
    }
}

// Sum over the 16 lanes of a DPP row; every lane receives it.
__device__ __forceinline__ uint32_t row_sum_u32(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false);
    return v;
}

// Birchfield-Tomasi terms of one pixel, packed: bits 0-9 twice the value,
// 10-19 min and 20-29 max over the value and its two half-way interpolants
// (all in 2x units); bit 31 marks "no pixel" (x < 0).
__device__ __forceinline__ uint32_t bt_pack(int c, int left, int right)
{
    const int v2 = 2 * c, a = c + left, b = c + right;
    const int mn = min(v2, min(a, b)), mx = max(v2, max(a, b));
    return (uint32_t)v2 | ((uint32_t)mn << 10) | ((uint32_t)mx << 20);
}
constexpr uint32_t BT_NONE = 0x80000000u;

// BT2 cost of a left term word and a right term word (uint16 result).
__device__ __forceinline__ uint32_t bt_cost(uint32_t lt, uint32_t rt)
{
    if (rt & BT_NONE) return INV16;
    const int l2 = lt & 0x3FF, lmin = (lt >> 10) & 0x3FF, lmax = (lt >> 20) & 0x3FF;
    const int r2 = rt & 0x3FF, rmin = (rt >> 10) & 0x3FF, rmax = (rt >> 20) & 0x3FF;
    const int a = max(0, max(l2 - rmax, rmin - l2));
    const int b = max(0, max(r2 - lmax, lmin - r2));
    return (uint32_t)min(a, b);
}

} // namespace tdvh
#endif
