// FloatImage, DSIMem and the event timer of the C++ stage layer.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <vector>

#include "tdvision/images.hpp"

TDV_NAMESPACE_BEGIN

namespace {
struct CtxHolder {
    std::map<int, tdvh_ctx *> ctxs;
    ~CtxHolder()
    {
        for (auto &kv : ctxs) tdvh_ctx_destroy(kv.second);
    }
};
thread_local CtxHolder g_ctxs;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    void *get(size_t n)
    {
        if (bytes < n) {
            if (p) checkHip(hipFree(p), "hipFree");
            p = nullptr;
            checkHip(hipMalloc(&p, n), "hipMalloc");
            bytes = n;
        }
        return p;
    }
};
} // namespace

tdvh_ctx *deviceContext(int device)
{
    auto it = g_ctxs.ctxs.find(device);
    if (it != g_ctxs.ctxs.end()) return it->second;
    tdvh_ctx *c = nullptr;
    checkTdvh(tdvh_ctx_create(device, &c), "tdvh_ctx_create");
    g_ctxs.ctxs[device] = c;
    return c;
}

// ---- FloatImage ----------------------------------------------------------------

struct FloatImage::Impl {
    Dim dim;
    std::vector<float> host;        // float view on the host
    std::vector<uint8_t> host8;     // uint8 gray on the host
    DevBuf dev, dev8, disp;         // float, uint8, int16 on the device
    bool hostValid = false, host8Valid = false, devValid = false, dev8Valid = false;
    int depth = 0;
};

FloatImage::FloatImage() {}

FloatImage FloatImage::CreateCPU(const Dim &dim)
{
    FloatImage f;
    f.m_impl = std::make_shared<Impl>();
    f.m_impl->dim = Dim(dim.width(), dim.height());
    f.m_impl->host.assign(dim.width() * dim.height(), 0.f);
    f.m_impl->hostValid = true;
    return f;
}

FloatImage FloatImage::CreateDev(const Dim &dim)
{
    FloatImage f;
    f.m_impl = std::make_shared<Impl>();
    f.m_impl->dim = Dim(dim.width(), dim.height());
    f.m_impl->dev.get(dim.width() * dim.height() * sizeof(float));
    f.m_impl->devValid = true;
    return f;
}

FloatImage FloatImage::FromGray8(const uint8_t *pixels, size_t width, size_t height)
{
    FloatImage f;
    f.m_impl = std::make_shared<Impl>();
    f.m_impl->dim = Dim(width, height);
    f.m_impl->host8.assign(pixels, pixels + width * height);
    f.m_impl->host8Valid = true;
    return f;
}

const Dim &FloatImage::dim() const
{
    static const Dim zero(0, 0, 0);
    return m_impl ? m_impl->dim : zero;
}

bool FloatImage::empty() const { return !m_impl; }

float *FloatImage::devMem()
{
    Impl &m = *m_impl;
    const size_t n = m.dim.width() * m.dim.height();
    float *d = (float *)m.dev.get(n * sizeof(float));
    if (!m.devValid) {
        if (!m.hostValid) cpuMem();
        checkHip(hipMemcpy(d, m.host.data(), n * sizeof(float), hipMemcpyHostToDevice),
                 "hipMemcpy H2D");
        m.devValid = true;
    }
    return d;
}

float *FloatImage::cpuMem()
{
    Impl &m = *m_impl;
    const size_t n = m.dim.width() * m.dim.height();
    if (!m.hostValid) {
        m.host.resize(n);
        if (m.devValid) {
            checkHip(hipMemcpy(m.host.data(), m.dev.p, n * sizeof(float), hipMemcpyDeviceToHost),
                     "hipMemcpy D2H");
        } else if (m.host8Valid) {
            for (size_t i = 0; i < n; i++) m.host[i] = m.host8[i] * (1.0f / 255.0f);
        } else if (m.dev8Valid) {
            m.host8.resize(n);
            checkHip(hipMemcpy(m.host8.data(), m.dev8.p, n, hipMemcpyDeviceToHost),
                     "hipMemcpy D2H");
            m.host8Valid = true;
            for (size_t i = 0; i < n; i++) m.host[i] = m.host8[i] * (1.0f / 255.0f);
        }
        m.hostValid = true;
    }
    return m.host.data();
}

const uint8_t *FloatImage::devMemU8()
{
    Impl &m = *m_impl;
    const size_t W = m.dim.width(), H = m.dim.height();
    uint8_t *d = (uint8_t *)m.dev8.get(W * H);
    if (!m.dev8Valid) {
        if (m.host8Valid) {
            checkHip(hipMemcpy(d, m.host8.data(), W * H, hipMemcpyHostToDevice), "hipMemcpy H2D");
        } else {
            float *f = devMem();
            checkTdvh(tdvh_unit_float_to_u8(deviceContext(), f, (int)W, (int)W, (int)H, d, (int)W,
                                            nullptr),
                      "tdvh_unit_float_to_u8");
        }
        m.dev8Valid = true;
    }
    return d;
}

uint8_t *FloatImage::writableDevU8()
{
    Impl &m = *m_impl;
    uint8_t *d = (uint8_t *)m.dev8.get(m.dim.width() * m.dim.height());
    m.dev8Valid = true;
    m.devValid = m.hostValid = m.host8Valid = false;
    return d;
}

int16_t *FloatImage::disparityDev()
{
    const size_t n = m_impl->dim.width() * m_impl->dim.height();
    return (int16_t *)m_impl->disp.get(n * sizeof(int16_t));
}

void FloatImage::disparityDepth(int depth) { m_impl->depth = depth; }
int FloatImage::disparityDepth() const { return m_impl ? m_impl->depth : 0; }

void FloatImage::syncFromDisparity()
{
    Impl &m = *m_impl;
    const int W = (int)m.dim.width(), H = (int)m.dim.height();
    float *f = (float *)m.dev.get((size_t)W * H * sizeof(float));
    checkTdvh(tdvh_disp_to_unit_float(deviceContext(), disparityDev(), W, W, H,
                                      m.depth > 0 ? m.depth : 1, f, W, nullptr),
              "tdvh_disp_to_unit_float");
    m.devValid = true;
    m.hostValid = false;
    m.host8Valid = false;
    m.dev8Valid = false;
}

// ---- DSIMem -----------------------------------------------------------------------

struct DSIMem::Impl {
    Dim dim;
    int dtype = TDVH_F32, layout = TDVH_LAYOUT_STRIDED;
    DevBuf buf;
    FloatImage leftOrigin;
    uint32_t costBound = 0;
};

DSIMem::DSIMem() {}

DSIMem DSIMem::Create(const Dim &dim, FloatImage lorigin, int dtype, int layout)
{
    DSIMem d;
    d.m_impl = std::make_shared<Impl>();
    d.m_impl->dim = dim;
    d.m_impl->dtype = dtype;
    d.m_impl->layout = layout;
    const size_t bytes = tdvh_dsi_bytes((int)dim.width(), (int)dim.height(), (int)dim.depth(),
                                        dtype, layout);
    if (!bytes) throw Exception("DSIMem::Create: bad shape or dtype");
    d.m_impl->buf.get(bytes);
    d.m_impl->leftOrigin = lorigin;
    return d;
}

tdvh_dsi DSIMem::mem() const
{
    tdvh_dsi d{};
    if (!m_impl) return d;
    const Dim &dm = m_impl->dim;
    checkTdvh(tdvh_dsi_init(&d, m_impl->buf.p, (int)dm.width(), (int)dm.height(), (int)dm.depth(),
                            m_impl->dtype, m_impl->layout),
              "tdvh_dsi_init");
    if (m_impl->layout == TDVH_LAYOUT_STRIDED) { // reference order [x][y][d]
        d.stride_x = (int64_t)dm.height() * dm.depth();
        d.stride_y = (int64_t)dm.depth();
    }
    return d;
}

const Dim &DSIMem::dim() const
{
    static const Dim zero(0, 0, 0);
    return m_impl ? m_impl->dim : zero;
}

int DSIMem::dtype() const { return m_impl ? m_impl->dtype : 0; }

void *DSIMem::toCpuMem() const
{
    if (!m_impl) return nullptr;
    const Dim &dm = m_impl->dim;
    const size_t esize = m_impl->dtype == TDVH_U16 ? 2 : 4;
    const size_t bytes = dm.size() * esize;
    unsigned char *host = new unsigned char[bytes];
    try {
        if (m_impl->layout == TDVH_LAYOUT_STRIDED) {
            checkHip(hipMemcpy(host, m_impl->buf.p, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        } else {
            DevBuf tmp;
            tdvh_dsi src = mem(), dst{};
            checkTdvh(tdvh_dsi_init(&dst, tmp.get(bytes), (int)dm.width(), (int)dm.height(),
                                    (int)dm.depth(), m_impl->dtype, TDVH_LAYOUT_STRIDED),
                      "tdvh_dsi_init");
            dst.stride_x = (int64_t)dm.height() * dm.depth();
            dst.stride_y = (int64_t)dm.depth();
            checkTdvh(tdvh_dsi_convert(deviceContext(), &src, &dst, nullptr), "tdvh_dsi_convert");
            checkHip(hipMemcpy(host, tmp.p, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
        }
    } catch (...) {
        delete[] host;
        throw;
    }
    return host;
}

FloatImage DSIMem::leftOrigin() const { return m_impl ? m_impl->leftOrigin : FloatImage(); }
uint32_t DSIMem::costBound() const { return m_impl ? m_impl->costBound : 0; }
void DSIMem::costBound(uint32_t bound)
{
    if (m_impl) m_impl->costBound = bound;
}
void DSIMem::leftOrigin(FloatImage img)
{
    if (m_impl) m_impl->leftOrigin = img;
}

// ---- DeviceBenchmarker ----------------------------------------------------------------

DeviceBenchmarker::DeviceBenchmarker() : m_start(nullptr), m_stop(nullptr)
{
    checkHip(hipEventCreate((hipEvent_t *)&m_start), "hipEventCreate");
    checkHip(hipEventCreate((hipEvent_t *)&m_stop), "hipEventCreate");
}

DeviceBenchmarker::~DeviceBenchmarker()
{
    (void)hipEventDestroy((hipEvent_t)m_start);
    (void)hipEventDestroy((hipEvent_t)m_stop);
}

void DeviceBenchmarker::begin() { checkHip(hipEventRecord((hipEvent_t)m_start, nullptr), "hipEventRecord"); }

void DeviceBenchmarker::end()
{
    checkHip(hipEventRecord((hipEvent_t)m_stop, nullptr), "hipEventRecord");
    checkHip(hipEventSynchronize((hipEvent_t)m_stop), "hipEventSynchronize");
    float ms = 0.f;
    checkHip(hipEventElapsedTime(&ms, (hipEvent_t)m_start, (hipEvent_t)m_stop),
             "hipEventElapsedTime");
    m_elapsed = Benchmark();
    m_elapsed.addProbeSec(ms / 1000.0);
}

TDV_NAMESPACE_END
