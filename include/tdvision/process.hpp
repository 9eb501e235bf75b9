// Work units and processes (workunit.hpp:9-34, process.hpp, cudaprocess.hpp,
// processrunner.hpp of the reference).
#ifndef TDV_PROCESS_HPP
#define TDV_PROCESS_HPP

#include <atomic>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

TDV_NAMESPACE_BEGIN

class WorkUnit {
public:
    WorkUnit() : m_workName("Unknow") {}
    virtual ~WorkUnit() {}
    // Processes one element; false once its inputs are exhausted.
    virtual bool update() = 0;
    const std::string &workName() const { return m_workName; }
    void workName(const std::string &name) { m_workName = name; }

private:
    std::string m_workName;
};

class Process {
public:
    virtual ~Process() {}
    virtual void process() = 0;
    virtual void finish() {}
};

// Runs its units round-robin on one device until all of them return false
// (cudaprocess.cpp:9-30).
class DeviceProcess : public Process {
public:
    explicit DeviceProcess(int deviceId) : m_deviceId(deviceId), m_end(false) {}
    void process();
    void finish() { m_end = true; }
    void addWork(WorkUnit *wu) { m_units.push_back(wu); }

private:
    std::vector<WorkUnit *> m_units;
    int m_deviceId;
    std::atomic<bool> m_end;
};
typedef DeviceProcess CUDAProcess;

// Runs one work unit on its own thread until it returns false (the
// reference's TWorkUnitProcess, workunitprocess.hpp).
template <typename Unit>
class WorkUnitProcess : public Unit, public Process {
public:
    void process()
    {
        while (!m_end && this->update()) {
        }
    }
    void finish() { m_end = true; }

private:
    std::atomic<bool> m_end{false};
};

// Flow control of a frame source (ctrlprocess.hpp:13-57, ctrlprocess.cpp:6-24):
// in Step mode one frame passes per step(); in Continuous mode every frame
// passes while the consumer's input pipes have room, and frames arriving while
// they are full are dropped.
class FlowCtrl {
public:
    enum Mode { Continuous, Step };
    void pause() { m_step = false; }
    void step()
    {
        m_mode = Step;
        m_step = true;
    }
    void stepMode()
    {
        m_mode = Step;
        m_step = false;
    }
    void continuous()
    {
        m_step = true;
        m_mode = Continuous;
    }
    Mode mode() const { return m_mode; }

protected:
    // Whether the next frame may pass (consumes a pending step).
    bool testFlow()
    {
        if (!m_step) return false;
        if (m_mode == Step) m_step = false;
        return true;
    }

private:
    std::atomic<bool> m_step{false};
    std::atomic<Mode> m_mode{Step};
};

// Receives exceptions escaping a process thread (exceptionreport.hpp).
class ExceptionReport {
public:
    virtual ~ExceptionReport() {}
    virtual void errorOcurred(const std::exception &ex) = 0;
};

// Something that owns processes: a null-terminated array (processgroup.hpp).
class ProcessGroup {
public:
    virtual ~ProcessGroup() {}
    virtual Process **processes() = 0;
};

class ArrayProcessGroup : public ProcessGroup {
public:
    ArrayProcessGroup() { m_procs.push_back(nullptr); }
    void addProcess(Process *p) { m_procs.insert(m_procs.end() - 1, p); }
    Process **processes() { return m_procs.data(); }

private:
    std::vector<Process *> m_procs;
};

// One thread per process; an exception in any of them is reported and
// finishes all processes (processrunner.cpp:19-79).
class ProcessRunner {
public:
    ProcessRunner(ProcessGroup &group, ExceptionReport *report = nullptr);
    ~ProcessRunner();
    void run();
    void join();
    void finishAll();
    bool errorOcurred() const { return m_errorOcurred; }
    const std::string &lastError() const { return m_lastError; }

private:
    void caller(Process *p);
    std::vector<Process *> m_procs;
    std::vector<std::thread> m_threads;
    ExceptionReport *m_report;
    std::atomic<bool> m_errorOcurred;
    std::string m_lastError;
    std::mutex m_mutex;
};

TDV_NAMESPACE_END

#endif
