// Stereo matchers and their factory (stereomatcher.hpp:12-25,
// devstereomatcher.cpp:14-49, fastwtamatcher.cpp:32-52,
// semigloballabmatcher.cpp:40-68, commonstereomatcherfactory.cpp:24-107).
#ifndef TDV_MATCHERS_HPP
#define TDV_MATCHERS_HPP

#include <memory>
#include <string>

#include "stages.hpp"

TDV_NAMESPACE_BEGIN

class StereoMatcher : public ProcessGroup {
public:
    virtual void inputs(ReadPipe<FloatImage> *leftInput, ReadPipe<FloatImage> *rightInput) = 0;
    virtual ReadPipe<FloatImage> *output() = 0;
    virtual std::string name() const = 0;
    virtual Benchmark matchcostBenchmark() const = 0;
    virtual Benchmark optimizationBenchmark() const = 0;
};

// Copies optimizer outputs to the host (cpyimagetocpu.cpp:5-10).
class CpyImageToCPU : public WorkUnit {
public:
    CpyImageToCPU() { workName("Copy to CPU"); }
    void input(ReadPipe<FloatImage> *p) { m_rpipe = p; }
    ReadPipe<FloatImage> *output() { return &m_wpipe; }
    bool update();

private:
    ReadPipe<FloatImage> *m_rpipe = nullptr;
    ReadWritePipe<FloatImage> m_wpipe;
};

// Frame-pair gate in front of a matcher (CtrlWork, ctrlprocess.cpp:27-73):
// passes or drops left/right pairs according to the FlowCtrl mode.
class CtrlWork : public WorkUnit, public FlowCtrl {
public:
    CtrlWork() { workName("Flow control"); }
    void inputs(ReadPipe<FloatImage> *l, ReadPipe<FloatImage> *r)
    {
        m_lrpipe = l;
        m_rrpipe = r;
    }
    ReadPipe<FloatImage> *leftImgOutput() { return &m_lwpipe; }
    ReadPipe<FloatImage> *rightImgOutput() { return &m_rwpipe; }
    bool update();
    size_t passed() const { return m_passed; }
    size_t dropped() const { return m_dropped; }

private:
    ReadPipe<FloatImage> *m_lrpipe = nullptr, *m_rrpipe = nullptr;
    ReadWritePipe<FloatImage> m_lwpipe, m_rwpipe;
    std::atomic<size_t> m_passed{0}, m_dropped{0};
};
typedef WorkUnitProcess<CtrlWork> CtrlProcess;

// Gaussian 5x5 prefilter of an input image on the device (the reference's
// "median filter" stage: MedianFilterCPU, medianfiltercpu.cpp:7-19, which
// DevStereoMatcher uses; MedianFilterDev, medianfilterdev.cpp:11-27, is the
// device stage of the same name). Runs tdvh_prefilter_gauss5 on the uint8
// plane.
class MedianFilterDev : public WorkUnit {
public:
    MedianFilterDev() { workName("Median filter device"); }
    void input(ReadPipe<FloatImage> *p) { m_rpipe = p; }
    ReadPipe<FloatImage> *output() { return &m_wpipe; }
    bool update();

private:
    ReadPipe<FloatImage> *m_rpipe = nullptr;
    ReadWritePipe<FloatImage> m_wpipe;
};
typedef MedianFilterDev MedianFilterCPU;

// Erode x3 then dilate x3 (3x3 element) of a disparity map (dilate.cpp:12-22),
// on the device via tdvh_postfilter_open on the int16 plane; the float d/D
// view is refreshed from it.
class Dilate : public WorkUnit {
public:
    Dilate(int iterations = 3) : m_iterations(iterations) { workName("Dilate"); }
    void input(ReadPipe<FloatImage> *p) { m_rpipe = p; }
    ReadPipe<FloatImage> *output() { return &m_wpipe; }
    bool update();

private:
    int m_iterations;
    ReadPipe<FloatImage> *m_rpipe = nullptr;
    ReadWritePipe<FloatImage> m_wpipe;
};

// [prefilter ->] matching cost -> optimizer -> host copy on one device
// process. As in the reference (devstereomatcher.cpp:6-49) both inputs go
// through the Gaussian prefilter unless useMedianFilter(false) is set before
// inputs().
class DevStereoMatcher : public StereoMatcher {
public:
    DevStereoMatcher(int device = 0);
    void useMedianFilter(bool on) { m_useMedianfilter = on; }
    bool useMedianFilter() const { return m_useMedianfilter; }
    void inputs(ReadPipe<FloatImage> *leftInput, ReadPipe<FloatImage> *rightInput);
    Process **processes() { return m_procs; }
    void setMatchingCost(std::shared_ptr<MatchingCost> mc) { m_matchCost = mc; }
    void setOptimizer(std::shared_ptr<Optimizer> opt) { m_optimizer = opt; }
    ReadPipe<FloatImage> *output() { return m_cpy.output(); }
    std::string name() const;
    Benchmark matchcostBenchmark() const
    {
        return m_matchCost ? m_matchCost->benchmark() : Benchmark();
    }
    Benchmark optimizationBenchmark() const
    {
        return m_optimizer ? m_optimizer->benchmark() : Benchmark();
    }

private:
    std::shared_ptr<MatchingCost> m_matchCost;
    std::shared_ptr<Optimizer> m_optimizer;
    MedianFilterDev m_medianFilter[2];
    bool m_useMedianfilter = true;
    CpyImageToCPU m_cpy;
    DeviceProcess m_process;
    Process *m_procs[2];
};

// Base of the fused single-process matchers.
class FusedMatcher : public StereoMatcher, public Process {
public:
    FusedMatcher(int maxDisparity, int device);
    void inputs(ReadPipe<FloatImage> *l, ReadPipe<FloatImage> *r)
    {
        m_lrpipe = l;
        m_rrpipe = r;
    }
    Process **processes() { return m_procs; }
    ReadPipe<FloatImage> *output() { return &m_wpipe; }
    Benchmark matchcostBenchmark() const { return m_mark; }
    Benchmark optimizationBenchmark() const { return m_mark; }
    void process();
    bool update();

protected:
    virtual void run(FloatImage left, FloatImage right, FloatImage out) = 0;
    int m_maxDisparity;
    int m_device;

private:
    ReadPipe<FloatImage> *m_lrpipe = nullptr, *m_rrpipe = nullptr;
    ReadWritePipe<FloatImage> m_wpipe;
    Benchmark m_mark;
    Process *m_procs[2];
};

// DevStereoMatcher's BT + SGM chain (with its Gaussian prefilter and,
// optionally, the Dilate postfilter) as one hipGraph per frame
// (tdvh_pipeline): the stages run as graph nodes on the device instead of
// WorkUnits with blocking kernel sequences. The pipeline is built for the
// first frame's shape and rebuilt when the shape changes.
class GraphStereoMatcher : public FusedMatcher {
public:
    explicit GraphStereoMatcher(int maxDisparity, bool postfilter = false, int device = 0)
        : FusedMatcher(maxDisparity, device), m_postfilter(postfilter)
    {
    }
    ~GraphStereoMatcher();
    std::string name() const { return "GraphStereoMatcher"; }

protected:
    void run(FloatImage left, FloatImage right, FloatImage out);

private:
    bool m_postfilter;
    tdvh_pipeline *m_pipe = nullptr;
    int m_w = 0, m_h = 0;
};

// SSD 7x7 + lowest-index WTA without a cost volume.
class FastWTAMatcher : public FusedMatcher {
public:
    explicit FastWTAMatcher(int maxDisparity, int device = 0) : FusedMatcher(maxDisparity, device)
    {
    }
    std::string name() const { return "FastWTAMatcher"; }

protected:
    void run(FloatImage left, FloatImage right, FloatImage out);
};

// BT recomputed inside SGM, then WTA: no cost volume in memory.
class SemiglobalLabMatcher : public FusedMatcher {
public:
    explicit SemiglobalLabMatcher(int maxDisparity, int device = 0)
        : FusedMatcher(maxDisparity, device)
    {
    }
    std::string name() const { return "SemiglobalLabMatcher"; }

protected:
    void run(FloatImage left, FloatImage right, FloatImage out);
};

// commonstereomatcherfactory.hpp:9-55: any matching cost with any optimiser
// on the device (commonstereomatcherfactory.cpp:39-81). The reference's CPU
// optimiser modes (DynamicProgOnCPU, SemiGlobalOnCPU, WTAOnCPU) map to the
// device-backed aliases SemiGlobalCPU / WTACPU / DynamicProgCPU, and Global
// to SemiGlobal as there. The CPU compute mode (OpenCV matchers,
// commonstereomatcherfactory.cpp:83-115) is not part of this library.
class CommonStereoMatcherFactory {
public:
    enum ComputeDevMode { Device, CPU };
    enum MatchMode { SSD, CrossCorrelationNorm, BirchfieldTomasi };
    enum OptMode { WTA, DynamicProg, DynamicProgOnCPU, Global, SemiGlobal, SemiGlobalOnCPU, WTAOnCPU };
    // the names this library used before (same values)
    typedef ComputeDevMode MatchingMode;
    typedef MatchMode MatchingCostType;
    typedef OptMode OptimizationType;

    // the reference's defaults (commonstereomatcherfactory.cpp:16-22) except
    // the compute mode: Device, the only one here
    CommonStereoMatcherFactory()
        : m_compMode(Device), m_matchMode(SSD), m_optMode(WTA), m_maxDisparity(64)
    {
    }
    void computeDev(ComputeDevMode m) { m_compMode = m; }
    void matchingCost(MatchMode t) { m_matchMode = t; }
    void optimization(OptMode t) { m_optMode = t; }
    void maxDisparity(int d) { m_maxDisparity = d; }
    // Throws tdv::Exception for the CPU compute mode.
    StereoMatcher *createStereoMatcher();

private:
    ComputeDevMode m_compMode;
    MatchMode m_matchMode;
    OptMode m_optMode;
    int m_maxDisparity;
};

TDV_NAMESPACE_END

#endif
