// Runtime of the C++ stage layer: error checks, device processes, runner.
#include <hip/hip_runtime.h>

#include <exception>

#include "tdvh.h"
#include "tdvision/process.hpp"

TDV_NAMESPACE_BEGIN

void checkTdvh(int status, const char *what)
{
    if (status != TDVH_OK)
        throw DeviceException(status, std::string(what) + ": " + tdvh_last_error());
}

void checkHip(int st, const char *what)
{
    if (st != hipSuccess)
        throw DeviceException(TDVH_ERR_DEVICE,
                              std::string(what) + ": " + hipGetErrorString((hipError_t)st));
}

// A unit that throws is dropped; its WriteGuard has finished its output
// pipe on the way out, so the units downstream of it drain and finish
// theirs, and a reader at the end of the chain sees the end instead of
// waiting forever. The first exception is rethrown once the loop ends.
void DeviceProcess::process()
{
    checkHip(hipSetDevice(m_deviceId), "hipSetDevice");
    std::vector<bool> alive(m_units.size(), true);
    std::exception_ptr err;
    bool cont = true;
    while (cont && !m_end) {
        cont = false;
        for (size_t i = 0; i < m_units.size(); i++) {
            if (alive[i]) {
                try {
                    alive[i] = m_units[i]->update();
                } catch (...) {
                    if (!err) err = std::current_exception();
                    alive[i] = false;
                }
            }
            cont = cont || alive[i];
        }
    }
    if (err) std::rethrow_exception(err);
}

ProcessRunner::ProcessRunner(ProcessGroup &group, ExceptionReport *report)
    : m_report(report), m_errorOcurred(false)
{
    for (Process **p = group.processes(); *p; ++p) m_procs.push_back(*p);
}

ProcessRunner::~ProcessRunner() { join(); }

void ProcessRunner::caller(Process *p)
{
    try {
        p->process();
    } catch (const std::exception &ex) {
        {
            std::lock_guard<std::mutex> lock(m_mutex);
            m_errorOcurred = true;
            m_lastError = ex.what();
            if (m_report) m_report->errorOcurred(ex);
        }
        finishAll();
    }
}

void ProcessRunner::run()
{
    for (Process *p : m_procs) m_threads.emplace_back(&ProcessRunner::caller, this, p);
}

void ProcessRunner::join()
{
    for (auto &t : m_threads)
        if (t.joinable()) t.join();
    m_threads.clear();
}

void ProcessRunner::finishAll()
{
    for (Process *p : m_procs) p->finish();
}

TDV_NAMESPACE_END
