// Stereo matchers and the factory.
#include <hip/hip_runtime.h>

#include "tdvision/matchers.hpp"

TDV_NAMESPACE_BEGIN

bool CpyImageToCPU::update()
{
    WriteGuard<ReadWritePipe<FloatImage>> guard(m_wpipe);
    FloatImage img;
    if (m_rpipe->read(&img)) {
        img.cpuMem();
        guard.write(img);
    }
    return guard.wasWrite();
}

DevStereoMatcher::DevStereoMatcher(int device) : m_process(device)
{
    m_procs[0] = &m_process;
    m_procs[1] = nullptr;
}

bool CtrlWork::update()
{
    FloatImage l, r;
    if (!(m_lrpipe->read(&l) && m_rrpipe->read(&r))) {
        m_lwpipe.finish();
        m_rwpipe.finish();
        return false;
    }
    if (testFlow() &&
        (mode() != Continuous || (!m_lwpipe.isFull() && !m_rwpipe.isFull()))) {
        m_lwpipe.write(l);
        m_rwpipe.write(r);
        m_passed++;
    } else {
        m_dropped++;
    }
    return true;
}

bool MedianFilterDev::update()
{
    WriteGuard<ReadWritePipe<FloatImage>> guard(m_wpipe);
    FloatImage in;
    if (m_rpipe->read(&in)) {
        const int W = (int)in.dim().width(), H = (int)in.dim().height();
        FloatImage out = FloatImage::CreateDev(in.dim());
        const uint8_t *src = in.devMemU8();
        checkTdvh(tdvh_prefilter_gauss5(deviceContext(), src, W, W, H, out.writableDevU8(), W,
                                        nullptr),
                  "tdvh_prefilter_gauss5");
        checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        guard.write(out);
    }
    return guard.wasWrite();
}

bool Dilate::update()
{
    WriteGuard<ReadWritePipe<FloatImage>> guard(m_wpipe);
    FloatImage img;
    if (m_rpipe->read(&img)) {
        const int W = (int)img.dim().width(), H = (int)img.dim().height();
        int16_t *d = img.disparityDev();
        checkTdvh(tdvh_postfilter_open(deviceContext(), d, W, W, H, m_iterations, d, W, nullptr),
                  "tdvh_postfilter_open");
        img.syncFromDisparity();
        checkHip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
        guard.write(img);
    }
    return guard.wasWrite();
}

void DevStereoMatcher::inputs(ReadPipe<FloatImage> *leftInput, ReadPipe<FloatImage> *rightInput)
{
    if (!m_matchCost || !m_optimizer) throw Exception("DevStereoMatcher: stages not set");
    if (m_useMedianfilter) {
        m_medianFilter[0].input(leftInput);
        m_medianFilter[1].input(rightInput);
        m_matchCost->inputs(m_medianFilter[0].output(), m_medianFilter[1].output());
        m_process.addWork(&m_medianFilter[0]);
        m_process.addWork(&m_medianFilter[1]);
    } else {
        m_matchCost->inputs(leftInput, rightInput);
    }
    m_optimizer->input(m_matchCost->output());
    m_cpy.input(m_optimizer->output());
    m_process.addWork(m_matchCost.get());
    m_process.addWork(m_optimizer.get());
    m_process.addWork(&m_cpy);
}

std::string DevStereoMatcher::name() const
{
    return (m_matchCost ? m_matchCost->workName() : std::string("")) + "+" +
           (m_optimizer ? m_optimizer->workName() : std::string(""));
}

FusedMatcher::FusedMatcher(int maxDisparity, int device)
    : m_maxDisparity(maxDisparity), m_device(device)
{
    m_procs[0] = this;
    m_procs[1] = nullptr;
}

void FusedMatcher::process()
{
    checkHip(hipSetDevice(m_device), "hipSetDevice");
    while (update()) {
    }
}

bool FusedMatcher::update()
{
    WriteGuard<ReadWritePipe<FloatImage>> guard(m_wpipe);
    FloatImage left, right;
    if (m_lrpipe->read(&left) && m_rrpipe->read(&right)) {
        if (left.dim() != right.dim()) throw Exception("left and right images differ in size");
        FloatImage out = FloatImage::CreateDev(left.dim());
        out.disparityDepth(m_maxDisparity);
        DeviceBenchmarker bm;
        bm.begin();
        run(left, right, out);
        out.syncFromDisparity();
        bm.end();
        m_mark.addProbe(bm.elapsedTime());
        out.cpuMem();
        guard.write(out);
    }
    return guard.wasWrite();
}

void FastWTAMatcher::run(FloatImage left, FloatImage right, FloatImage out)
{
    const int W = (int)left.dim().width(), H = (int)left.dim().height();
    checkTdvh(tdvh_fast_wta_ssd(deviceContext(m_device), left.devMemU8(), right.devMemU8(), W, H,
                                W, m_maxDisparity, 3, out.disparityDev(), W, nullptr),
              "tdvh_fast_wta_ssd");
}

void SemiglobalLabMatcher::run(FloatImage left, FloatImage right, FloatImage out)
{
    const int W = (int)left.dim().width(), H = (int)left.dim().height();
    const tdvh_sgm_params p{8, 32, 0xFF, 0};
    checkTdvh(tdvh_stereo_bt_sgm(deviceContext(m_device), left.devMemU8(), right.devMemU8(), W, H,
                                 W, m_maxDisparity, &p, TDVH_FUSE_COST, out.disparityDev(), W,
                                 nullptr),
              "tdvh_stereo_bt_sgm");
}

GraphStereoMatcher::~GraphStereoMatcher()
{
    if (m_pipe) (void)tdvh_pipeline_destroy(m_pipe);
}

void GraphStereoMatcher::run(FloatImage left, FloatImage right, FloatImage out)
{
    const int W = (int)left.dim().width(), H = (int)left.dim().height();
    if (!m_pipe || W != m_w || H != m_h) {
        if (m_pipe) checkTdvh(tdvh_pipeline_destroy(m_pipe), "tdvh_pipeline_destroy");
        m_pipe = nullptr;
        const tdvh_sgm_params p{8, 32, 0xFF, 0};
        const int flags = TDVH_PIPE_PREFILTER | (m_postfilter ? TDVH_PIPE_POSTFILTER : 0);
        checkTdvh(tdvh_pipeline_create(deviceContext(m_device), W, H, m_maxDisparity, &p, flags,
                                       &m_pipe),
                  "tdvh_pipeline_create");
        m_w = W;
        m_h = H;
    }
    checkTdvh(tdvh_pipeline_run(m_pipe, left.devMemU8(), right.devMemU8(), W, out.disparityDev(),
                                W, nullptr),
              "tdvh_pipeline_run");
}

StereoMatcher *CommonStereoMatcherFactory::createStereoMatcher()
{
    if (m_compMode == CPU)
        throw Exception("CPU (OpenCV) matchers are not part of this library; use Device");
    std::shared_ptr<MatchingCost> cost;
    switch (m_matchMode) {
    case SSD: cost = std::make_shared<SSDDev>(m_maxDisparity); break;
    case CrossCorrelationNorm: cost = std::make_shared<CrossCorrelationDev>(m_maxDisparity); break;
    case BirchfieldTomasi: cost = std::make_shared<BirchfieldCostDev>(m_maxDisparity); break;
    default: throw Exception("unknown matching cost mode");
    }
    std::shared_ptr<Optimizer> opt;
    switch (m_optMode) {
    case WTA: opt = std::make_shared<WTADev>(); break;
    case WTAOnCPU: opt = std::make_shared<WTACPU>(); break;
    case DynamicProg: opt = std::make_shared<DynamicProgDev>(); break;
    case DynamicProgOnCPU: opt = std::make_shared<DynamicProgCPU>(); break;
    case Global:
    case SemiGlobal: opt = std::make_shared<SemiGlobalDev>(); break;
    case SemiGlobalOnCPU: opt = std::make_shared<SemiGlobalCPU>(); break;
    default: throw Exception("unknown optimisation mode");
    }
    DevStereoMatcher *m = new DevStereoMatcher(0);
    m->setMatchingCost(cost);
    m->setOptimizer(opt);
    return m;
}

TDV_NAMESPACE_END
