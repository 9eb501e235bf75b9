// Stage API of the stereo path: MatchingCost / Optimizer work units and the
// concrete stages, with the reference's names and contracts
// (matchingcost.hpp:13-57, optimizer.hpp:13-55, birchfieldcostdev.hpp,
// ssddev.hpp, crosscorrelationdev.hpp, semiglobaldev.hpp, wtadev.hpp,
// dynamicprogdev.hpp). Every updateImpl calls the tdvh C ABI.
#ifndef TDV_STAGES_HPP
#define TDV_STAGES_HPP

#include "images.hpp"
#include "pipe.hpp"
#include "process.hpp"

TDV_NAMESPACE_BEGIN

class MatchingCost : public WorkUnit {
public:
    virtual void inputs(ReadPipe<FloatImage> *leftInput, ReadPipe<FloatImage> *rightInput) = 0;
    virtual ReadPipe<DSIMem> *output() = 0;
    virtual Benchmark benchmark() const = 0;
};

// Reads a pair, (re)allocates the DSI only when its size changes
// (matchingcost.cpp:27-31), times updateImpl and passes the DSI on.
class AbstractMatchingCost : public MatchingCost {
public:
    explicit AbstractMatchingCost(int disparityMax);
    void inputs(ReadPipe<FloatImage> *lpipe, ReadPipe<FloatImage> *rpipe)
    {
        m_lrpipe = lpipe;
        m_rrpipe = rpipe;
    }
    ReadPipe<DSIMem> *output() { return &m_wpipe; }
    bool update();
    Benchmark benchmark() const { return m_mark; }

protected:
    virtual void updateImpl(FloatImage left, FloatImage right, DSIMem dsi) = 0;
    // Element type and layout of the volume the stage writes: float strided
    // [x][y][d] by default, the reference's DSIMem (dsimem.cpp:22-42), so a
    // subclass that only overrides updateImpl keeps its contract.
    virtual int dsiDtype() const { return TDVH_F32; }
    virtual int dsiLayout() const { return TDVH_LAYOUT_STRIDED; }

private:
    ReadPipe<FloatImage> *m_lrpipe, *m_rrpipe;
    ReadWritePipe<DSIMem, DSIMem> m_wpipe;
    size_t m_maxDisparity;
    DSIMem m_dsi;
    Benchmark m_mark;
};

// Birchfield-Tomasi, uint16 in 2x units (birchfieldcostdev.cpp:9-17).
class BirchfieldCostDev : public AbstractMatchingCost {
public:
    explicit BirchfieldCostDev(int disparityMax) : AbstractMatchingCost(disparityMax)
    {
        workName("Birchfield Cost");
    }

protected:
    void updateImpl(FloatImage left, FloatImage right, DSIMem dsi);
    int dsiDtype() const { return TDVH_U16; }
    int dsiLayout() const { return TDVH_LAYOUT_SPLIT16; }
};

// Window SSD, uint32 (ssddev.cpp:9-17; 7x7 there).
class SSDDev : public AbstractMatchingCost {
public:
    explicit SSDDev(int disparityMax, int radius = 3)
        : AbstractMatchingCost(disparityMax), m_radius(radius)
    {
        workName("SSD");
    }

protected:
    void updateImpl(FloatImage left, FloatImage right, DSIMem dsi);
    int dsiDtype() const { return TDVH_U32; }

private:
    int m_radius;
};

// NCC cost 1 - rho, float (crosscorrelationdev.cpp:9-17; 3x3 plain there).
class CrossCorrelationDev : public AbstractMatchingCost {
public:
    explicit CrossCorrelationDev(int disparityMax, int radius = 1, bool zeroMean = false)
        : AbstractMatchingCost(disparityMax), m_radius(radius), m_zeroMean(zeroMean)
    {
        workName("Cross Correlation");
    }

protected:
    void updateImpl(FloatImage left, FloatImage right, DSIMem dsi);
    int dsiDtype() const { return TDVH_F32; }

private:
    int m_radius;
    bool m_zeroMean;
};

class Optimizer : public WorkUnit {
public:
    virtual void input(ReadPipe<DSIMem> *rpipe) = 0;
    virtual ReadPipe<FloatImage> *output() = 0;
    virtual Benchmark benchmark() const = 0;
};

// Reads a DSI, creates the output image (float d/D view + int16 disparity
// plane, optimizer.cpp:18-19), times updateImpl; calls finished() when the
// input ends.
class AbstractOptimizer : public Optimizer {
public:
    AbstractOptimizer() : m_rpipe(nullptr) {}
    void input(ReadPipe<DSIMem> *rpipe) { m_rpipe = rpipe; }
    ReadPipe<FloatImage> *output() { return &m_wpipe; }
    Benchmark benchmark() const { return m_mark; }
    bool update();

protected:
    virtual void updateImpl(DSIMem dsi, FloatImage outimg) = 0;
    virtual void finished() {}

private:
    ReadPipe<DSIMem> *m_rpipe;
    ReadWritePipe<FloatImage, FloatImage> m_wpipe;
    Benchmark m_mark;
};

// SGM + WTA (semiglobaldev.cpp:16-35) over any cost stage's volume: integer
// penalties p1/p2 for uint16 (BT) and uint32 (SSD) costs, float penalties
// for float (NCC) costs, by default the reference's 30/255 and 150/255
// (semiglobal.cu:254-255).
class SemiGlobalDev : public AbstractOptimizer {
public:
    SemiGlobalDev(int p1 = 8, int p2 = 32, int flags = 0)
    {
        workName("Semi-global device");
        m_params = tdvh_sgm_params{p1, p2, 0xFF, flags, 30.0f / 255.0f, 150.0f / 255.0f};
    }
    void params(const tdvh_sgm_params &p) { m_params = p; }
    void floatPenalties(float p1, float p2)
    {
        m_params.fp1 = p1;
        m_params.fp2 = p2;
    }

protected:
    void updateImpl(DSIMem dsi, FloatImage outimg);
    void finished();

private:
    tdvh_sgm_params m_params;
};

// Winner-take-all (wtadev.cpp:7-11); TDVH_WTA_TREE keeps the GPU tie rule.
class WTADev : public AbstractOptimizer {
public:
    explicit WTADev(int mode = TDVH_WTA_LOWEST) : m_mode(mode) { workName("WTA"); }

protected:
    void updateImpl(DSIMem dsi, FloatImage outimg);

private:
    int m_mode;
};

// Scanline DP with occlusion penalty (dynamicprogdev.cpp:11-28) over any
// cost stage's volume: occlusionPenalty for integer (BT, SSD) costs, a float
// penalty for float (NCC) costs, 0 by default as the reference's DP, which
// has none (dynamicprogcpu.cpp:29-59).
class DynamicProgDev : public AbstractOptimizer {
public:
    explicit DynamicProgDev(uint32_t occlusionPenalty = 20, int flags = 0)
        : m_pocc(occlusionPenalty), m_poccf(0.0f), m_flags(flags)
    {
        workName("DynamicProg");
    }
    void floatPenalty(float p) { m_poccf = p; }

protected:
    void updateImpl(DSIMem dsi, FloatImage outimg);
    void finished();

private:
    uint32_t m_pocc;
    float m_poccf;
    int m_flags;
};

// The reference's CPU optimisers (semiglobalcpu.cpp:173-230,
// wtacpu.cpp:14-41, dynamicprogcpu.cpp:94-112) under their names, run on the
// device: same inputs and outputs, the same arithmetic as the device stages
// (lowest-index WTA ties as wtacpu.cpp:28; DynamicProgCPU computes the DP its
// reference compiles out, dynamicprogcpu.cpp:96-112). Nothing here runs on
// the host.
typedef SemiGlobalDev SemiGlobalCPU;
typedef WTADev WTACPU;
typedef DynamicProgDev DynamicProgCPU;

TDV_NAMESPACE_END

#endif
