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
    virtual int dsiDtype() const = 0;
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

// SGM + WTA (semiglobaldev.cpp:16-35).
class SemiGlobalDev : public AbstractOptimizer {
public:
    SemiGlobalDev(int p1 = 8, int p2 = 32, int flags = 0)
    {
        workName("Semi-global device");
        m_params = tdvh_sgm_params{p1, p2, 0xFF, flags};
    }
    void params(const tdvh_sgm_params &p) { m_params = p; }

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

// Scanline DP with occlusion penalty (dynamicprogdev.cpp:11-28).
class DynamicProgDev : public AbstractOptimizer {
public:
    explicit DynamicProgDev(uint32_t occlusionPenalty = 20, int flags = 0)
        : m_pocc(occlusionPenalty), m_flags(flags)
    {
        workName("DynamicProg");
    }

protected:
    void updateImpl(DSIMem dsi, FloatImage outimg);
    void finished();

private:
    uint32_t m_pocc;
    int m_flags;
};

TDV_NAMESPACE_END

#endif
