// Image and disparity-space-image handles (floatimage.hpp:15-140,
// dsimem.hpp:12-136, benchmark.hpp:18-89 of the reference), backed by HIP
// device memory instead of OpenCV/CUDA buffers.
#ifndef TDV_IMAGES_HPP
#define TDV_IMAGES_HPP

#include <cstdint>
#include <memory>

#include "common.hpp"
#include "tdvh.h"

TDV_NAMESPACE_BEGIN

// Per-thread tdvh context of a device (contexts are single-thread-affine).
tdvh_ctx *deviceContext(int device = 0);

// Shared handle to a W x H image that lives on the host and/or the device
// and is converted lazily: float [0,1] (the reference's value convention)
// and uint8 gray (the integer path's input, round(v * 255)). Optimizer
// outputs also carry the int16 disparity plane they were built from.
class FloatImage {
public:
    FloatImage();
    static FloatImage CreateCPU(const Dim &dim);
    static FloatImage CreateDev(const Dim &dim);
    // Host uint8 gray image (row-major, pitch = width).
    static FloatImage FromGray8(const uint8_t *pixels, size_t width, size_t height);

    const Dim &dim() const;
    bool empty() const;
    float *devMem();              // float copy on the device
    float *cpuMem();              // float copy on the host
    const uint8_t *devMemU8();    // uint8 gray copy on the device
    // The uint8 plane on the device, to be written by a producer; the other
    // views are refreshed from it on demand.
    uint8_t *writableDevU8();
    int16_t *disparityDev();      // int16 disparity plane (allocated on first use)
    void disparityDepth(int depth); // the D of d/D for the float view
    int disparityDepth() const;
    // Refreshes the float view from the disparity plane (d / D).
    void syncFromDisparity();

    struct Impl;

private:
    std::shared_ptr<Impl> m_impl;
};

// Shared handle to a cost volume on the device. Strided volumes use the
// reference's [x][y][d] layout (dsimem.cpp:34-35); uint16 BT volumes use the
// SGM kernels' SPLIT16 layout.
class DSIMem {
public:
    DSIMem();
    static DSIMem Create(const Dim &dim, FloatImage lorigin, int dtype = TDVH_F32,
                         int layout = TDVH_LAYOUT_STRIDED);
    tdvh_dsi mem() const;
    const Dim &dim() const;
    int dtype() const;
    // New[]'d host copy in [x][y][d] order (element size of dtype); the
    // caller deletes it. One bulk transfer (dsimem.cpp:55-92 used W*H).
    void *toCpuMem() const;
    FloatImage leftOrigin() const;
    void leftOrigin(FloatImage img);
    // Largest valid cost the producing stage writes, with INVALID exactly on
    // the cells with d > x (0 = unknown): lets a consumer choose narrower
    // exact arithmetic (DynamicProgDev's 32-bit sums).
    uint32_t costBound() const;
    void costBound(uint32_t bound);

    struct Impl;

private:
    std::shared_ptr<Impl> m_impl;
};

// Running mean of stage times (benchmark.hpp:18-46).
class Benchmark {
public:
    Benchmark() : m_secs(0.0), m_count(0) {}
    void addProbeSec(double s)
    {
        m_secs += s;
        m_count += 1;
    }
    void addProbe(const Benchmark &b) { addProbeSec(b.secs()); }
    double secs() const { return m_count ? m_secs / double(m_count) : 0.0; }
    double millisecs() const { return secs() * 1000.0; }
    size_t probes() const { return m_count; }

private:
    double m_secs;
    size_t m_count;
};

// hipEvent timer on the null stream (the CudaBenchmarker role,
// benchmark.cpp:7-31); events are created once and reused.
class DeviceBenchmarker {
public:
    DeviceBenchmarker();
    ~DeviceBenchmarker();
    void begin();
    void end();
    Benchmark elapsedTime() const { return m_elapsed; }

private:
    void *m_start, *m_stop;
    Benchmark m_elapsed;
};
typedef DeviceBenchmarker CudaBenchmarker;

TDV_NAMESPACE_END

#endif
