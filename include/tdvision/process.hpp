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
