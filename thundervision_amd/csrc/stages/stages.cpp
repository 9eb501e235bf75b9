// Concrete matching-cost and optimizer stages: each updateImpl is one tdvh
// C-ABI call on the null stream of the calling thread's device.
#include <algorithm>

#include "tdvision/stages.hpp"

TDV_NAMESPACE_BEGIN

AbstractMatchingCost::AbstractMatchingCost(int disparityMax)
    : m_lrpipe(nullptr), m_rrpipe(nullptr), m_maxDisparity((size_t)disparityMax)
{
    if (disparityMax <= 0) throw Exception("disparityMax must be positive");
}

bool AbstractMatchingCost::update()
{
    WriteGuard<ReadWritePipe<DSIMem, DSIMem>> wguard(m_wpipe);
    FloatImage left, right;
    if (m_lrpipe->read(&left) && m_rrpipe->read(&right)) {
        // The reference takes min(left, right) sizes but reads both images
        // with one pitch (matchingcost.cpp:81-85, birchfieldcost.cu:150-158);
        // here the images must agree.
        if (left.dim() != right.dim()) throw Exception("left and right images differ in size");
        const Dim pkt(left.dim().width(), left.dim().height(), m_maxDisparity);
        if (m_dsi.dim() != pkt) m_dsi = DSIMem::Create(pkt, left, dsiDtype(), dsiLayout());
        m_dsi.leftOrigin(left);
        DeviceBenchmarker bm;
        bm.begin();
        updateImpl(left, right, m_dsi);
        bm.end();
        m_mark.addProbe(bm.elapsedTime());
        wguard.write(m_dsi);
    }
    return wguard.wasWrite();
}

void BirchfieldCostDev::updateImpl(FloatImage left, FloatImage right, DSIMem dsi)
{
    tdvh_dsi d = dsi.mem();
    const int W = d.width;
    checkTdvh(tdvh_cost_bt(deviceContext(), left.devMemU8(), right.devMemU8(), W, d.height, W, &d,
                           nullptr),
              "tdvh_cost_bt");
}

void SSDDev::updateImpl(FloatImage left, FloatImage right, DSIMem dsi)
{
    tdvh_dsi d = dsi.mem();
    const int W = d.width;
    checkTdvh(tdvh_cost_ssd(deviceContext(), left.devMemU8(), right.devMemU8(), W, d.height, W,
                            m_radius, &d, nullptr),
              "tdvh_cost_ssd");
    // (2r + 1)^2 * 255^2 at most; INVALID exactly where d > x
    const uint64_t side = 2 * (uint64_t)m_radius + 1, bound = side * side * 255 * 255;
    dsi.costBound(bound <= 0xFFFFFFFEull ? (uint32_t)bound : 0);
}

void CrossCorrelationDev::updateImpl(FloatImage left, FloatImage right, DSIMem dsi)
{
    tdvh_dsi d = dsi.mem();
    const int W = d.width;
    checkTdvh(tdvh_cost_ncc(deviceContext(), left.devMemU8(), right.devMemU8(), W, d.height, W,
                            m_radius, m_zeroMean ? 1 : 0, &d, nullptr),
              "tdvh_cost_ncc");
}

bool AbstractOptimizer::update()
{
    WriteGuard<ReadWritePipe<FloatImage, FloatImage>> guard(m_wpipe);
    DSIMem dsi;
    if (m_rpipe->read(&dsi)) {
        FloatImage out = FloatImage::CreateDev(Dim(dsi.dim().width(), dsi.dim().height()));
        out.disparityDepth((int)dsi.dim().depth());
        DeviceBenchmarker bm;
        bm.begin();
        updateImpl(dsi, out);
        out.syncFromDisparity();
        bm.end();
        m_mark.addProbe(bm.elapsedTime());
        guard.write(out);
    }
    if (!guard.wasWrite()) finished();
    return guard.wasWrite();
}

void SemiGlobalDev::updateImpl(DSIMem dsi, FloatImage out)
{
    tdvh_dsi c = dsi.mem();
    checkTdvh(tdvh_sgm(deviceContext(), &c, &m_params, nullptr, out.disparityDev(), c.width,
                       nullptr),
              "tdvh_sgm");
}

void SemiGlobalDev::finished() { (void)tdvh_ctx_trim(deviceContext()); }

void WTADev::updateImpl(DSIMem dsi, FloatImage out)
{
    tdvh_dsi c = dsi.mem();
    checkTdvh(tdvh_wta(deviceContext(), &c, m_mode, out.disparityDev(), c.width, nullptr),
              "tdvh_wta");
}

void DynamicProgDev::updateImpl(DSIMem dsi, FloatImage out)
{
    tdvh_dsi c = dsi.mem();
    if (c.dtype == TDVH_F32) {
        checkTdvh(tdvh_dp_f32(deviceContext(), &c, m_poccf, m_flags, out.disparityDev(), c.width,
                              nullptr),
                  "tdvh_dp_f32");
        return;
    }
    int flags = m_flags;
    // costs below 2^22 from a known producer: 32-bit sums, same results
    const uint32_t bound = dsi.costBound();
    if (bound && bound < (1u << 22) && m_pocc < (1u << 20)) flags |= TDVH_DP_COST_BOUNDED;
    checkTdvh(tdvh_dp(deviceContext(), &c, m_pocc, flags, out.disparityDev(), c.width, nullptr),
              "tdvh_dp");
}

void DynamicProgDev::finished() { (void)tdvh_ctx_trim(deviceContext()); }

TDV_NAMESPACE_END
