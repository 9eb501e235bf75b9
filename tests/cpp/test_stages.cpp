// Tests of the C++ stage layer. `test_stages cpu` runs the runtime tests
// (pipes, runner) that need no device, following the reference's
// testpipeline.cpp (Connections, PipeAndFilter, BoundedBuffer);
// `test_stages gpu <dir>` runs stage chains on the GPU and compares them with
// expected results written into <dir> by tests/test_stages_cpp.py.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "tdvision/matchers.hpp"
// the reference's own header names (as its tests include them)
#include "tdvision/dynamicprogcpu.hpp"
#include "tdvision/semiglobalcpu.hpp"
#include "tdvision/ssddev.hpp"
#include "tdvision/wtacpu.hpp"

using namespace tdv;

static int g_fail = 0;
#define CHECK(c)                                                                                   \
    do {                                                                                           \
        if (!(c)) {                                                                                \
            fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c);                   \
            g_fail++;                                                                              \
        }                                                                                          \
    } while (0)

// ---- runtime (CPU) ------------------------------------------------------------

static void test_pipe_connections()
{
    ReadWritePipe<int> p(4);
    p.write(1);
    p.write(2);
    int v = 0;
    CHECK(p.read(&v) && v == 1);
    CHECK(p.read(&v) && v == 2);
    p.finish();
    CHECK(!p.read(&v));
    {
        ReadWritePipe<int> q;
        {
            WriteGuard<ReadWritePipe<int>> g(q);
        } // nothing written: the guard finishes the pipe
        CHECK(!q.read(&v));
    }
}

static void test_bounded_buffer()
{
    ReadWritePipe<int> p(10);
    const int N = 1000;
    std::thread prod([&] {
        for (int i = 0; i < N; i++) p.write(i);
        p.finish();
    });
    int v, expect = 0;
    while (p.read(&v)) {
        CHECK(v == expect);
        expect++;
    }
    prod.join();
    CHECK(expect == N);
}

namespace {
// Doubles its input; ends when the input ends.
class Doubler : public WorkUnit, public Process {
public:
    ReadPipe<int> *in = nullptr;
    ReadWritePipe<int> out{2};
    bool update()
    {
        WriteGuard<ReadWritePipe<int>> g(out);
        int v;
        if (in->read(&v)) g.write(2 * v);
        return g.wasWrite();
    }
    void process()
    {
        while (update()) {
        }
    }
};
class Source : public Process {
public:
    ReadWritePipe<int> out{2};
    int n = 50;
    void process()
    {
        for (int i = 0; i < n; i++) out.write(i);
        out.finish();
    }
};
class Thrower : public Process {
public:
    void process() { throw Exception("planned failure"); }
};
class Group : public ProcessGroup {
public:
    std::vector<Process *> v;
    Process **processes()
    {
        if (v.empty() || v.back()) v.push_back(nullptr);
        return v.data();
    }
};
} // namespace

static void test_pipe_and_filter()
{
    Source src;
    Doubler dbl;
    dbl.in = &src.out;
    Group g;
    g.v = {&src, &dbl};
    ProcessRunner runner(g);
    runner.run();
    int v, i = 0;
    while (dbl.out.read(&v)) {
        CHECK(v == 2 * i);
        i++;
    }
    runner.join();
    CHECK(i == 50);
    CHECK(!runner.errorOcurred());
}

static void test_runner_reports_errors()
{
    Thrower t;
    Group g;
    g.v = {&t};
    ProcessRunner runner(g);
    runner.run();
    runner.join();
    CHECK(runner.errorOcurred());
    CHECK(runner.lastError().find("planned failure") != std::string::npos);
}

// ---- GPU stage chains -------------------------------------------------------------

static std::vector<char> slurp(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

struct Case {
    int W, H, D;
    std::vector<char> L, R;
};

// Feeds one pair through a matcher and returns its int16 disparities.
static std::vector<int16_t> run_matcher(StereoMatcher &m, const Case &c, FloatImage *keep)
{
    ReadWritePipe<FloatImage> lp, rp;
    m.inputs(&lp, &rp);
    ProcessRunner runner(m);
    runner.run();
    lp.write(FloatImage::FromGray8((const uint8_t *)c.L.data(), c.W, c.H));
    rp.write(FloatImage::FromGray8((const uint8_t *)c.R.data(), c.W, c.H));
    lp.finish();
    rp.finish();
    FloatImage out;
    std::vector<int16_t> disp(c.W * c.H);
    if (m.output()->read(&out)) {
        checkHip(hipMemcpy(disp.data(), out.disparityDev(), disp.size() * 2, hipMemcpyDeviceToHost),
                 "copy disparity");
        if (keep) *keep = out;
    }
    runner.join();
    CHECK(!runner.errorOcurred());
    if (runner.errorOcurred()) fprintf(stderr, "runner: %s\n", runner.lastError().c_str());
    return disp;
}

static bool same(const std::vector<int16_t> &a, const std::vector<char> &b, const char *what)
{
    const bool ok = a.size() * 2 == b.size() && memcmp(a.data(), b.data(), b.size()) == 0;
    if (!ok) fprintf(stderr, "%s: disparities differ from the oracle\n", what);
    return ok;
}

static void test_gpu(const std::string &dir)
{
    Case c;
    FILE *f = fopen((dir + "/shape.txt").c_str(), "r");
    CHECK(f && fscanf(f, "%d %d %d", &c.W, &c.H, &c.D) == 3);
    if (f) fclose(f);
    c.L = slurp(dir + "/left.u8");
    c.R = slurp(dir + "/right.u8");

    { // BT + SGM through DevStereoMatcher, plus the float d/D view
        DevStereoMatcher m;
        m.setMatchingCost(std::make_shared<BirchfieldCostDev>(c.D));
        m.setOptimizer(std::make_shared<SemiGlobalDev>());
        FloatImage out;
        CHECK(m.useMedianFilter()); // on by default, as devstereomatcher.cpp:6-12
        auto d = run_matcher(m, c, &out);
        CHECK(same(d, slurp(dir + "/bt_sgm_pf.i16"), "prefilter+BT+SGM"));
        const float *fv = out.cpuMem();
        bool ok = true;
        for (size_t i = 0; i < d.size(); i++) ok = ok && fv[i] == (float)d[i] / (float)c.D;
        CHECK(ok);
        CHECK(m.name() == "Birchfield Cost+Semi-global device");
        CHECK(m.matchcostBenchmark().probes() == 1 && m.optimizationBenchmark().probes() == 1);
    }
    { // without the prefilter, and the Dilate postfilter stage
        DevStereoMatcher m;
        m.useMedianFilter(false);
        m.setMatchingCost(std::make_shared<BirchfieldCostDev>(c.D));
        m.setOptimizer(std::make_shared<SemiGlobalDev>());
        CHECK(same(run_matcher(m, c, nullptr), slurp(dir + "/bt_sgm.i16"), "BT+SGM raw"));
        DevStereoMatcher m2;
        m2.setMatchingCost(std::make_shared<BirchfieldCostDev>(c.D));
        m2.setOptimizer(std::make_shared<SemiGlobalDev>());
        Dilate dil;
        ReadWritePipe<FloatImage> lp, rp;
        m2.inputs(&lp, &rp);
        dil.input(m2.output());
        ProcessRunner runner(m2);
        runner.run();
        lp.write(FloatImage::FromGray8((const uint8_t *)c.L.data(), c.W, c.H));
        rp.write(FloatImage::FromGray8((const uint8_t *)c.R.data(), c.W, c.H));
        lp.finish();
        rp.finish();
        CHECK(dil.update());
        FloatImage out;
        CHECK(dil.output()->read(&out));
        std::vector<int16_t> d(c.W * c.H);
        checkHip(hipMemcpy(d.data(), out.disparityDev(), d.size() * 2, hipMemcpyDeviceToHost),
                 "copy disparity");
        runner.join();
        CHECK(!runner.errorOcurred());
        CHECK(same(d, slurp(dir + "/bt_sgm_pf_open.i16"), "prefilter+BT+SGM+Dilate"));
        const float *fv = out.cpuMem();
        bool ok = true;
        for (size_t i = 0; i < d.size(); i++) ok = ok && fv[i] == (float)d[i] / (float)c.D;
        CHECK(ok);
    }
    { // the chain as one graph per frame (prefilter [+ Dilate]), two frames each
        for (int post = 0; post < 2; post++) {
            GraphStereoMatcher m(c.D, post != 0);
            ReadWritePipe<FloatImage> lp(2), rp(2);
            m.inputs(&lp, &rp);
            ProcessRunner runner(m);
            runner.run();
            for (int k = 0; k < 2; k++) {
                lp.write(FloatImage::FromGray8((const uint8_t *)c.L.data(), c.W, c.H));
                rp.write(FloatImage::FromGray8((const uint8_t *)c.R.data(), c.W, c.H));
            }
            lp.finish();
            rp.finish();
            const auto want = slurp(dir + (post ? "/bt_sgm_pf_open.i16" : "/bt_sgm_pf.i16"));
            for (int k = 0; k < 2; k++) {
                FloatImage out;
                CHECK(m.output()->read(&out));
                std::vector<int16_t> d(c.W * c.H);
                checkHip(hipMemcpy(d.data(), out.disparityDev(), d.size() * 2,
                                   hipMemcpyDeviceToHost),
                         "copy disparity");
                CHECK(same(d, want, post ? "graph chain + Dilate" : "graph chain"));
            }
            runner.join();
            CHECK(!runner.errorOcurred());
        }
    }
    { // fused BT inside SGM
        SemiglobalLabMatcher m(c.D);
        CHECK(same(run_matcher(m, c, nullptr), slurp(dir + "/bt_sgm.i16"), "SemiglobalLab"));
    }
    { // BT + WTA (lowest index) and tree tie rule
        DevStereoMatcher m;
        m.setMatchingCost(std::make_shared<BirchfieldCostDev>(c.D));
        m.setOptimizer(std::make_shared<WTADev>());
        CHECK(same(run_matcher(m, c, nullptr), slurp(dir + "/bt_wta_pf.i16"), "BT+WTA"));
    }
    { // SSD + DP and SSD + WTA, FastWTAMatcher
        DevStereoMatcher m;
        m.setMatchingCost(std::make_shared<SSDDev>(c.D));
        m.setOptimizer(std::make_shared<DynamicProgDev>(20));
        CHECK(same(run_matcher(m, c, nullptr), slurp(dir + "/ssd_dp_pf.i16"), "SSD+DP"));
        FastWTAMatcher fw(c.D);
        CHECK(same(run_matcher(fw, c, nullptr), slurp(dir + "/ssd_wta.i16"), "FastWTA"));
    }
    { // factory
        CommonStereoMatcherFactory fac;
        fac.maxDisparity(c.D); // SSD + WTA are the reference's defaults
        std::unique_ptr<StereoMatcher> m(fac.createStereoMatcher());
        CHECK(same(run_matcher(*m, c, nullptr), slurp(dir + "/ssd_wta_pf.i16"),
                   "factory SSD+WTA"));
        fac.computeDev(CommonStereoMatcherFactory::CPU);
        bool threw = false;
        try {
            delete fac.createStereoMatcher();
        } catch (const Exception &) {
            threw = true;
        }
        CHECK(threw);
    }
    { // every cost x optimiser pairing of the factory (commonstereomatcherfactory.cpp:39-81)
        const char *costs[3] = {"ssd", "ncc", "bt"};
        const char *opts[3] = {"wta", "dp", "sgm"};
        const CommonStereoMatcherFactory::MatchMode mm[3] = {
            CommonStereoMatcherFactory::SSD, CommonStereoMatcherFactory::CrossCorrelationNorm,
            CommonStereoMatcherFactory::BirchfieldTomasi};
        const CommonStereoMatcherFactory::OptMode om[3][2] = {
            {CommonStereoMatcherFactory::WTA, CommonStereoMatcherFactory::WTAOnCPU},
            {CommonStereoMatcherFactory::DynamicProg, CommonStereoMatcherFactory::DynamicProgOnCPU},
            {CommonStereoMatcherFactory::SemiGlobal, CommonStereoMatcherFactory::SemiGlobalOnCPU}};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int v = 0; v < 2; v++) {
                    CommonStereoMatcherFactory fac;
                    fac.maxDisparity(c.D);
                    fac.matchingCost(mm[i]);
                    fac.optimization(om[j][v]);
                    std::unique_ptr<StereoMatcher> m(fac.createStereoMatcher());
                    const std::string what = std::string("factory ") + costs[i] + "+" + opts[j];
                    CHECK(same(run_matcher(*m, c, nullptr),
                               slurp(dir + "/pair_" + costs[i] + "_" + opts[j] + ".i16"),
                               what.c_str()));
                }
    }
    { // TestSemiGlobalDev.WithSSD and TestSemiGlobalCPU.WithSSD
      // (testsemiglobal.cpp:20-31, testsemiglobalcpu.cpp:8-23): SSD into
      // SGM, with the prefilter, at 64 and 256 disparities
        for (int D : {64, 256}) {
            DevStereoMatcher m;
            m.setMatchingCost(std::make_shared<SSDDev>(D));
            if (D == 64) m.setOptimizer(std::make_shared<SemiGlobalDev>());
            else m.setOptimizer(std::make_shared<SemiGlobalCPU>());
            Case cd = c;
            cd.D = D;
            CHECK(same(run_matcher(m, cd, nullptr),
                       slurp(dir + "/ssd_sgm_pf_" + std::to_string(D) + ".i16"),
                       D == 64 ? "SSDDev(64)+SemiGlobalDev" : "SSDDev(256)+SemiGlobalCPU"));
        }
        DevStereoMatcher w;
        w.setMatchingCost(std::make_shared<SSDDev>(c.D));
        w.setOptimizer(std::make_shared<WTACPU>());
        CHECK(same(run_matcher(w, c, nullptr), slurp(dir + "/ssd_wta_pf.i16"), "SSD+WTACPU"));
        DevStereoMatcher d;
        d.setMatchingCost(std::make_shared<SSDDev>(c.D));
        d.setOptimizer(std::make_shared<DynamicProgCPU>(20));
        CHECK(same(run_matcher(d, c, nullptr), slurp(dir + "/ssd_dp_pf.i16"), "SSD+DynamicProgCPU"));
    }
    { // a cost stage that overrides only updateImpl keeps the float DSI default
        struct ZeroCost : AbstractMatchingCost {
            ZeroCost() : AbstractMatchingCost(4) {}
            void updateImpl(FloatImage, FloatImage, DSIMem dsi)
            {
                checkHip(hipMemset(dsi.mem().ptr, 0, dsi.dim().size() * 4), "clear");
            }
        } zc;
        ReadWritePipe<FloatImage> lp, rp;
        zc.inputs(&lp, &rp);
        lp.write(FloatImage::FromGray8((const uint8_t *)c.L.data(), c.W, c.H));
        rp.write(FloatImage::FromGray8((const uint8_t *)c.R.data(), c.W, c.H));
        CHECK(zc.update());
        DSIMem dsi;
        CHECK(zc.output()->read(&dsi));
        CHECK(dsi.mem().dtype == TDVH_F32 && dsi.mem().layout == TDVH_LAYOUT_STRIDED);
    }
    { // DSIMem download in the reference layout [x][y][d]
        ReadWritePipe<FloatImage> lp, rp;
        BirchfieldCostDev bt(c.D);
        bt.inputs(&lp, &rp);
        lp.write(FloatImage::FromGray8((const uint8_t *)c.L.data(), c.W, c.H));
        rp.write(FloatImage::FromGray8((const uint8_t *)c.R.data(), c.W, c.H));
        CHECK(bt.update());
        DSIMem dsi;
        CHECK(bt.output()->read(&dsi));
        unsigned char *host = (unsigned char *)dsi.toCpuMem();
        const auto want = slurp(dir + "/bt_xyd.u16");
        CHECK(want.size() == dsi.dim().size() * 2 && memcmp(host, want.data(), want.size()) == 0);
        delete[] host;
    }
    { // a device error becomes a tdv::DeviceException
        bool threw = false;
        try {
            checkTdvh(tdvh_sgm(deviceContext(), nullptr, nullptr, nullptr, nullptr, 0, nullptr),
                      "tdvh_sgm");
        } catch (const DeviceException &e) {
            threw = e.status() == TDVH_ERR_INVALID_ARG;
        }
        CHECK(threw);
    }
}

// CtrlWork (ctrlprocess.cpp:27-73): step mode passes one pair per step();
// continuous mode passes pairs while the outputs have room, drops the rest.
static void test_flow_control()
{
    auto frame = [] { return FloatImage::CreateCPU(Dim(2, 2)); };
    {
        ReadWritePipe<FloatImage> lin(8), rin(8);
        CtrlWork cw;
        cw.inputs(&lin, &rin);
        for (int i = 0; i < 3; i++) {
            lin.write(frame());
            rin.write(frame());
        }
        CHECK(cw.mode() == FlowCtrl::Step);
        CHECK(cw.update()); // paused: dropped
        CHECK(cw.passed() == 0 && cw.dropped() == 1);
        cw.step();
        CHECK(cw.update()); // one step: passes
        CHECK(cw.passed() == 1);
        CHECK(cw.update()); // the step is consumed: dropped
        CHECK(cw.passed() == 1 && cw.dropped() == 2);
        FloatImage a, b;
        CHECK(cw.leftImgOutput()->read(&a) && cw.rightImgOutput()->read(&b));
    }
    {
        ReadWritePipe<FloatImage> lin(8), rin(8);
        CtrlWork cw;
        cw.inputs(&lin, &rin);
        cw.continuous();
        for (int i = 0; i < 4; i++) {
            lin.write(frame());
            rin.write(frame());
        }
        lin.finish();
        rin.finish();
        CHECK(cw.update()); // outputs empty: passes
        CHECK(cw.update()); // outputs full (capacity 1, not read): dropped
        CHECK(cw.passed() == 1 && cw.dropped() == 1);
        FloatImage a, b;
        CHECK(cw.leftImgOutput()->read(&a) && cw.rightImgOutput()->read(&b));
        CHECK(cw.update()); // room again: passes
        CHECK(cw.passed() == 2);
        CHECK(cw.leftImgOutput()->read(&a) && cw.rightImgOutput()->read(&b));
        CHECK(cw.update()); // the 4th pair passes
        CHECK(!cw.update()); // inputs finished: outputs finish
        CHECK(cw.leftImgOutput()->read(&a)); // the 4th pair is still readable
        CHECK(!cw.leftImgOutput()->read(&a));
    }
}

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    if (mode == "cpu") {
        test_pipe_connections();
        test_bounded_buffer();
        test_pipe_and_filter();
        test_runner_reports_errors();
        test_flow_control();
    } else if (mode == "gpu" && argc > 2) {
        try {
            test_gpu(argv[2]);
        } catch (const std::exception &e) {
            fprintf(stderr, "exception: %s\n", e.what());
            g_fail++;
        }
    } else {
        fprintf(stderr, "usage: test_stages cpu | gpu <dir>\n");
        return 2;
    }
    printf("%s: %s (%d failures)\n", mode.c_str(), g_fail ? "FAILED" : "ok", g_fail);
    return g_fail ? 1 : 0;
}
