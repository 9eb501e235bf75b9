// Bounded-buffer pipes between work units (pipe.hpp:15-288 of the reference:
// ReadPipe / WritePipe / ReadWritePipe / WriteGuard, same contracts).
#ifndef TDV_PIPE_HPP
#define TDV_PIPE_HPP

#include <condition_variable>
#include <mutex>
#include <queue>

#include "common.hpp"

TDV_NAMESPACE_BEGIN

class BaseReadPipe {
public:
    virtual ~BaseReadPipe() {}
    // Waits for a read and discards the element.
    virtual void waitRead() = 0;
};

template <typename ReadType>
class ReadPipe : public BaseReadPipe {
public:
    virtual ~ReadPipe() {}
    // Blocks until an element arrives; false once the pipe is finished and
    // empty (the reader may exit).
    virtual bool read(ReadType *outvalue) = 0;
};

class BaseWritePipe {
public:
    virtual ~BaseWritePipe() {}
    // No more data will be written.
    virtual void finish() = 0;
};

template <typename WriteType>
class WritePipe : public BaseWritePipe {
public:
    typedef WriteType WriteValueType;
    virtual ~WritePipe() {}
    virtual void write(WriteType packet) = 0;
};

// Bounded buffer: writers block while maxSize elements are queued.
template <typename ReadType, typename WriteType = ReadType>
class ReadWritePipe : public WritePipe<WriteType>, public ReadPipe<ReadType> {
public:
    explicit ReadWritePipe(size_t maxSize = 1) : m_end(false), m_maxSize(maxSize) {}
    ~ReadWritePipe() { finish(); }

    void write(WriteType value)
    {
        std::unique_lock<std::mutex> lock(m_mutex);
        m_end = false;
        m_notFull.wait(lock, [&] { return m_queue.size() < m_maxSize || m_end; });
        if (!m_end) {
            m_queue.push(static_cast<ReadType>(value));
            m_notEmpty.notify_one();
        }
    }

    bool read(ReadType *out)
    {
        std::unique_lock<std::mutex> lock(m_mutex);
        m_notEmpty.wait(lock, [&] { return !m_queue.empty() || m_end; });
        const bool empty = m_queue.empty();
        if (!empty) {
            *out = m_queue.front();
            m_queue.pop();
        }
        m_notFull.notify_one();
        return !empty;
    }

    void finish()
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        m_end = true;
        m_notEmpty.notify_all();
        m_notFull.notify_all();
    }

    bool isFull() const
    {
        std::lock_guard<std::mutex> lock(m_mutex);
        return m_queue.size() == m_maxSize;
    }

    void reset() { m_end = false; }

    void waitRead()
    {
        ReadType tp;
        (void)read(&tp);
    }

private:
    std::queue<ReadType> m_queue;
    mutable std::mutex m_mutex;
    std::condition_variable m_notEmpty, m_notFull;
    bool m_end;
    size_t m_maxSize;
};

// Finishes the wrapped pipe on scope exit unless something was written.
template <typename WritePipeType>
class WriteGuard : public WritePipe<typename WritePipeType::WriteValueType> {
public:
    typedef typename WritePipeType::WriteValueType WriteValueType;
    explicit WriteGuard(WritePipeType &pipe) : m_pipe(pipe), m_wasWrite(false) {}
    ~WriteGuard()
    {
        if (!m_wasWrite) m_pipe.finish();
    }
    void write(WriteValueType packet)
    {
        m_pipe.write(packet);
        m_wasWrite = true;
    }
    void finish() { m_pipe.finish(); }
    bool wasWrite() const { return m_wasWrite; }

private:
    WritePipeType &m_pipe;
    bool m_wasWrite;
};

TDV_NAMESPACE_END

#endif
