// Context, error reporting, DSI descriptors and the reference path table.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "kernels.hpp"

namespace tdvh {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

// Context of the entry point running on this thread (set by use_device);
// the launch hooks find the profiler through it.
static thread_local tdvh_ctx *g_cur_ctx = nullptr;

static hipEvent_t prof_event(tdvh_ctx *c)
{
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(hipStream_t s)
{
    tdvh_ctx *c = g_cur_ctx;
    if (!c || !c->prof_on) return;
    c->prof_pending = prof_event(c);
    if (c->prof_pending) (void)hipEventRecord(c->prof_pending, s);
}

int check_launch(const char *what, hipStream_t s)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return TDVH_ERR_DEVICE;
    }
    tdvh_ctx *c = g_cur_ctx;
    if (c && c->prof_on && c->prof_pending) {
        hipEvent_t stop = prof_event(c);
        if (stop && hipEventRecord(stop, s) == hipSuccess)
            c->prof_marks.push_back({what, c->prof_pending, stop});
        c->prof_pending = nullptr;
    }
    return TDVH_OK;
}

size_t dtype_size(int dtype)
{
    switch (dtype) {
    case TDVH_U16: return 2;
    case TDVH_U32: return 4;
    case TDVH_F32: return 4;
    case TDVH_I32: return 4;
    default: return 0;
    }
}

int use_device(tdvh_ctx *ctx)
{
    TDVH_REQUIRE(ctx != nullptr, "null context");
    TDVH_CHECK_HIP(hipSetDevice(ctx->device));
    g_cur_ctx = ctx;
    return TDVH_OK;
}

void use_no_context() { g_cur_ctx = nullptr; }

} // namespace tdvh

int tdvh_ctx::scratch(const char *name, size_t bytes, void **out)
{
    Slot &s = slots[name];
    if (s.bytes < bytes) {
        if (s.ptr) {
            // buffers may still be in use by queued work on any stream
            TDVH_CHECK_HIP(hipDeviceSynchronize());
            TDVH_CHECK_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        s.tag = 0;
        hipError_t e = hipMalloc(&s.ptr, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            s.ptr = nullptr;
            tdvh::set_error("hipMalloc(%zu) for scratch '%s' failed: %s", bytes, name,
                            hipGetErrorString(e));
            return TDVH_ERR_OUT_OF_MEMORY;
        }
        s.bytes = bytes;
    }
    *out = s.ptr;
    return TDVH_OK;
}

int tdvh_ctx::release_all()
{
    if (!slots.empty()) TDVH_CHECK_HIP(hipDeviceSynchronize());
    for (auto &kv : slots)
        if (kv.second.ptr) TDVH_CHECK_HIP(hipFree(kv.second.ptr));
    slots.clear();
    host_tables.clear(); // their copies completed at the synchronisation above
    return TDVH_OK;
}

int tdvh_ctx::prof_resolve()
{
    for (auto &m : prof_marks) {
        TDVH_CHECK_HIP(hipEventSynchronize(m.stop));
        float ms = 0.f;
        TDVH_CHECK_HIP(hipEventElapsedTime(&ms, m.start, m.stop));
        bool found = false;
        for (auto &row : prof_table)
            if (row.first == m.name) {
                row.second.first += ms;
                row.second.second += 1;
                found = true;
            }
        if (!found) prof_table.push_back({m.name, {(double)ms, 1}});
        prof_pool.push_back(m.start);
        prof_pool.push_back(m.stop);
    }
    prof_marks.clear();
    return TDVH_OK;
}

size_t tdvh_ctx::held() const
{
    size_t n = 0;
    for (auto &kv : slots) n += kv.second.bytes;
    return n;
}

extern "C" {

const char *tdvh_last_error(void) { return tdvh::g_last_error.c_str(); }

int tdvh_ctx_create(int device, tdvh_ctx **out)
{
    tdvh::clear_error();
    TDVH_REQUIRE(out != nullptr, "out is null");
    int count = 0;
    TDVH_CHECK_HIP(hipGetDeviceCount(&count));
    TDVH_REQUIRE(device >= 0 && device < count, "device %d out of range (%d devices)", device,
                 count);
    TDVH_CHECK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    TDVH_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    tdvh_ctx *c = new tdvh_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    *out = c;
    return TDVH_OK;
}

int tdvh_ctx_destroy(tdvh_ctx *ctx)
{
    if (!ctx) return TDVH_OK;
    int rc = TDVH_OK;
    if (hipSetDevice(ctx->device) == hipSuccess) {
        rc = ctx->release_all();
        tdvh::release_ref_tables(ctx);
        (void)ctx->prof_resolve();
        for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
    }
    if (tdvh::g_cur_ctx == ctx) tdvh::g_cur_ctx = nullptr;
    delete ctx;
    return rc;
}

int tdvh_profile_enable(tdvh_ctx *ctx, int on)
{
    tdvh::clear_error();
    TDVH_TRY(tdvh::use_device(ctx));
    ctx->prof_on = on != 0;
    return TDVH_OK;
}

int tdvh_profile_reset(tdvh_ctx *ctx)
{
    tdvh::clear_error();
    TDVH_TRY(tdvh::use_device(ctx));
    TDVH_TRY(ctx->prof_resolve());
    ctx->prof_table.clear();
    return TDVH_OK;
}

int tdvh_profile_read(tdvh_ctx *ctx, int index, char *name, size_t name_len, double *total_ms,
                      int64_t *launches)
{
    tdvh::clear_error();
    TDVH_TRY(tdvh::use_device(ctx));
    TDVH_TRY(ctx->prof_resolve());
    TDVH_REQUIRE(index >= 0 && index < (int)ctx->prof_table.size(), "no profile entry %d", index);
    const auto &row = ctx->prof_table[index];
    if (name && name_len) {
        snprintf(name, name_len, "%s", row.first.c_str());
    }
    if (total_ms) *total_ms = row.second.first;
    if (launches) *launches = row.second.second;
    return TDVH_OK;
}

int tdvh_ctx_trim(tdvh_ctx *ctx)
{
    TDVH_TRY(tdvh::use_device(ctx));
    return ctx->release_all();
}

size_t tdvh_ctx_scratch_bytes(const tdvh_ctx *ctx) { return ctx ? ctx->held() : 0; }

size_t tdvh_dsi_bytes(int width, int height, int depth, int dtype, int layout)
{
    if (width <= 0 || height <= 0 || depth <= 0) return 0;
    if (layout == TDVH_LAYOUT_SPLIT16)
        return (size_t)width * height * tdvh::split_words(depth) * 4;
    return (size_t)width * height * depth * tdvh::dtype_size(dtype);
}

int tdvh_dsi_init(tdvh_dsi *out, void *ptr, int width, int height, int depth, int dtype,
                  int layout)
{
    tdvh::clear_error();
    TDVH_REQUIRE(out != nullptr, "out is null");
    TDVH_REQUIRE(width > 0 && height > 0 && depth > 0, "bad shape %dx%dx%d", width, height,
                 depth);
    TDVH_REQUIRE(tdvh::dtype_size(dtype) != 0, "bad dtype %d", dtype);
    TDVH_REQUIRE(layout == TDVH_LAYOUT_STRIDED || (layout == TDVH_LAYOUT_SPLIT16 && dtype == TDVH_U16),
                 "layout %d not available for dtype %d", layout, dtype);
    out->ptr = ptr;
    out->width = width;
    out->height = height;
    out->depth = depth;
    out->dtype = dtype;
    out->layout = layout;
    out->stride_x = depth;
    out->stride_y = (int64_t)width * depth;
    return TDVH_OK;
}

size_t tdvh_sgm_paths(int width, int height, tdvh_sgpath *paths)
{
    if (width <= 0 || height <= 0) return 0;
    const size_t count = (size_t)width + height + 2 * ((size_t)width + height - 1);
    if (!paths) return count;
    const int lastX = width - 1, lastY = height - 1;
    size_t k = 0;
    auto put = [&](int sx, int sy, int ex, int ey, int dx, int dy, int size) {
        tdvh_sgpath &p = paths[k++];
        p.sx = (uint16_t)sx;
        p.sy = (uint16_t)sy;
        p.ex = (uint16_t)ex;
        p.ey = (uint16_t)ey;
        p.dx = (int16_t)dx;
        p.dy = (int16_t)dy;
        p.size = (uint16_t)size;
        p.pad = 0;
    };
    // rows then columns (semiglobal.cu:116-146)
    for (int i = 0; i < height; i++) put(0, i, lastX, i, 1, 0, width);
    for (int i = 0; i < width; i++) put(i, 0, i, lastY, 0, 1, height);
    // (1,1): top row then left column, with the table's own end points
    // (semiglobal.cu:154-185)
    for (int i = 0; i < width; i++) {
        const int ey = i < lastY ? i : lastY;
        put(lastX - i, 0, lastX, ey, 1, 1, ey + 1);
    }
    for (int i = 0; i < height - 1; i++) {
        const int a = lastY - (i + 1);
        const int ex = a < lastX ? a : lastX;
        put(0, i + 1, ex, lastY, 1, 1, ex + 1);
    }
    // (-1,1): top row then right column (semiglobal.cu:191-230)
    for (int i = 0; i < width; i++) {
        const int size = (i + 1) < height ? (i + 1) : height;
        put(i, 0, 0, size - 1, -1, 1, size);
    }
    for (int i = 0; i < height - 1; i++) {
        const int sy = i + 1;
        const int size = (height - sy) < width ? (height - sy) : width;
        put(lastX, sy, lastX - size + 1, sy + size - 1, -1, 1, size);
    }
    return k;
}

} // extern "C"
