// Device-resident stereo pipeline for one frame shape, replayed as a hipGraph:
//   [Gaussian prefilter of both inputs] -> BT + SGM-8 + WTA (fused)
//   -> [erode/dilate postfilter]
// The reference runs these as WorkUnits on one CUDAProcess thread, each stage
// blocking on its own kernels (cudaprocess.cpp:9-30, devstereomatcher.cpp:14-49,
// reconstruction.cpp:30-31). Here the stages are nodes of one graph on the
// caller's stream: the host enqueues a frame as one graph launch and returns.
// Device-resident images and disparity buffers are read and written in place
// by a graph captured for that set of pointers (the last set is kept; a new
// set is captured on its first run); host buffers go through the pipeline's
// own device buffers with one copy each way around a second graph. The
// pipeline owns a private tdvh context so that the graphs' fixed kernel
// arguments (scratch pointers) cannot be invalidated by other calls; the
// first run executes eagerly (allocating scratch and building per-shape
// tables) and then captures.
#include "kernels.hpp"

// The buffers one graph reads and writes.
struct PipeIO {
    const uint8_t *L = nullptr, *R = nullptr;
    int pitch = 0;
    int16_t *disp = nullptr;
    int disp_pitch = 0;
    bool operator==(const PipeIO &o) const
    {
        return L == o.L && R == o.R && pitch == o.pitch && disp == o.disp &&
               disp_pitch == o.disp_pitch;
    }
};

struct PipeGraph {
    PipeIO io;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct tdvh_pipeline {
    tdvh_ctx *ctx = nullptr; // private context
    int W = 0, H = 0, D = 0, flags = 0;
    tdvh_sgm_params params{};
    uint8_t *L = nullptr, *R = nullptr, *Lf = nullptr, *Rf = nullptr;
    int16_t *disp = nullptr;
    void *mem = nullptr;
    hipStream_t cap = nullptr;
    bool warm = false;   // an eager run has allocated scratch and tables
    PipeGraph staged;    // over the pipeline's own buffers (host images)
    PipeGraph direct;    // over the caller's device buffers (last set seen)
};

namespace {

// The stages of one frame on stream s over the buffers io.
int pipeline_body(tdvh_pipeline *p, const PipeIO &io, hipStream_t s)
{
    const uint8_t *l = io.L, *r = io.R;
    int pitch = io.pitch;
    if (p->flags & TDVH_PIPE_PREFILTER) {
        TDVH_TRY(tdvh_prefilter_gauss5(p->ctx, io.L, io.pitch, p->W, p->H, p->Lf, p->W, s));
        TDVH_TRY(tdvh_prefilter_gauss5(p->ctx, io.R, io.pitch, p->W, p->H, p->Rf, p->W, s));
        l = p->Lf;
        r = p->Rf;
        pitch = p->W;
    }
    TDVH_TRY(tdvh_stereo_bt_sgm(p->ctx, l, r, p->W, p->H, pitch, p->D, &p->params,
                                TDVH_FUSE_COST, io.disp, io.disp_pitch, s));
    if (p->flags & TDVH_PIPE_POSTFILTER)
        TDVH_TRY(tdvh_postfilter_open(p->ctx, io.disp, io.disp_pitch, p->W, p->H, 3, io.disp,
                                      io.disp_pitch, s));
    return TDVH_OK;
}

void drop_graph(PipeGraph &g)
{
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = PipeGraph{};
}

// Runs the frame over io on s from graph g. If g holds another buffer set it
// is captured anew first; the pipeline's very first run executes eagerly (to
// allocate scratch and tables) and captures the graph for the next run.
int run_graph(tdvh_pipeline *p, PipeGraph &g, const PipeIO &io, hipStream_t s)
{
    bool launched = false;
    if (!(g.exec && g.io == io)) {
        drop_graph(g);
        if (!p->warm) {
            TDVH_TRY(pipeline_body(p, io, s));
            p->warm = true;
            launched = true;
        }
        if (!p->cap) TDVH_CHECK_HIP(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
        TDVH_CHECK_HIP(hipStreamSynchronize(s));
        TDVH_CHECK_HIP(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
        const int rc = pipeline_body(p, io, p->cap);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(p->cap, &gr);
        TDVH_TRY(rc);
        TDVH_CHECK_HIP(ec);
        g.graph = gr;
        TDVH_CHECK_HIP(hipGraphInstantiate(&g.exec, gr, nullptr, nullptr, 0));
        g.io = io;
    }
    if (!launched) TDVH_CHECK_HIP(hipGraphLaunch(g.exec, s));
    return TDVH_OK;
}

// Whether ptr is device memory of the current device's address space.
bool on_device(const void *ptr)
{
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, ptr) != hipSuccess) {
        (void)hipGetLastError(); // plain host memory
        return false;
    }
    return at.type == hipMemoryTypeDevice;
}

void drop(tdvh_pipeline *p)
{
    drop_graph(p->staged);
    drop_graph(p->direct);
    if (p->cap) (void)hipStreamDestroy(p->cap);
    if (p->mem) (void)hipFree(p->mem);
    if (p->ctx) (void)tdvh_ctx_destroy(p->ctx);
    delete p;
}

} // namespace

using namespace tdvh;

extern "C" int tdvh_pipeline_create(tdvh_ctx *ctx, int width, int height, int depth,
                                    const tdvh_sgm_params *params, int flags,
                                    tdvh_pipeline **out)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(out && params, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0, "bad image %dx%d", width, height);
    TDVH_REQUIRE(depth > 0 && depth <= MAX_DEPTH, "depth %d outside [1, %d]", depth, MAX_DEPTH);
    TDVH_REQUIRE(!(params->flags & TDVH_SGM_REF_GEOMETRY), "the pipeline runs the standard geometry");
    TDVH_REQUIRE((flags & ~(TDVH_PIPE_PREFILTER | TDVH_PIPE_POSTFILTER)) == 0, "unknown flags 0x%x",
                 flags);
    tdvh_pipeline *p = new tdvh_pipeline();
    p->W = width;
    p->H = height;
    p->D = depth;
    p->flags = flags;
    p->params = *params;
    const size_t n = (size_t)width * height, a = (n + 255) & ~(size_t)255;
    int rc = tdvh_ctx_create(ctx->device, &p->ctx);
    if (rc != TDVH_OK) {
        drop(p);
        return rc;
    }
    hipError_t e = hipMalloc(&p->mem, 4 * a + 2 * a);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        drop(p);
        set_error("hipMalloc for the pipeline failed: %s", hipGetErrorString(e));
        return TDVH_ERR_OUT_OF_MEMORY;
    }
    uint8_t *b = (uint8_t *)p->mem;
    p->L = b;
    p->R = b + a;
    p->Lf = b + 2 * a;
    p->Rf = b + 3 * a;
    p->disp = (int16_t *)(b + 4 * a);
    *out = p;
    return TDVH_OK;
}

extern "C" int tdvh_pipeline_run(tdvh_pipeline *p, const uint8_t *left, const uint8_t *right,
                                 int pitch, int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_REQUIRE(p && left && right && disp, "null pointer argument");
    TDVH_REQUIRE(pitch >= p->W && disp_pitch >= p->W, "pitch %d / disp_pitch %d < width %d",
                 pitch, disp_pitch, p->W);
    TDVH_TRY(use_device(p->ctx));
    hipStream_t s = (hipStream_t)stream;
    if (on_device(left) && on_device(right) && on_device(disp)) {
        PipeIO io;
        io.L = left;
        io.R = right;
        io.pitch = pitch;
        io.disp = disp;
        io.disp_pitch = disp_pitch;
        return run_graph(p, p->direct, io, s);
    }
    TDVH_CHECK_HIP(hipMemcpy2DAsync(p->L, p->W, left, pitch, p->W, p->H, hipMemcpyDefault, s));
    TDVH_CHECK_HIP(hipMemcpy2DAsync(p->R, p->W, right, pitch, p->W, p->H, hipMemcpyDefault, s));
    PipeIO io;
    io.L = p->L;
    io.R = p->R;
    io.pitch = p->W;
    io.disp = p->disp;
    io.disp_pitch = p->W;
    TDVH_TRY(run_graph(p, p->staged, io, s));
    TDVH_CHECK_HIP(hipMemcpy2DAsync(disp, (size_t)disp_pitch * 2, p->disp, (size_t)p->W * 2,
                                    (size_t)p->W * 2, p->H, hipMemcpyDefault, s));
    return TDVH_OK;
}

extern "C" int tdvh_pipeline_destroy(tdvh_pipeline *p)
{
    if (!p) return TDVH_OK;
    clear_error();
    TDVH_TRY(use_device(p->ctx));
    TDVH_CHECK_HIP(hipDeviceSynchronize());
    drop(p);
    return TDVH_OK;
}
