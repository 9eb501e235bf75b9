// Basic types of the C++ stage layer: Dim, exceptions, device checks.
// Mirrors the reference's tdvbasic/common.hpp + dim.hpp + exception.hpp
// surface (dim.hpp:9-93, exception.hpp:12-36, cuerr.hpp:11-23).
#ifndef TDV_COMMON_HPP
#define TDV_COMMON_HPP

#include <cstddef>
#include <stdexcept>
#include <string>

#define TDV_NAMESPACE_BEGIN namespace tdv {
#define TDV_NAMESPACE_END }

TDV_NAMESPACE_BEGIN

class Exception : public std::runtime_error {
public:
    explicit Exception(const std::string &msg) : std::runtime_error(msg) {}
};

// Failure of a device call (the CUException role, cuerr.hpp:11-23).
class DeviceException : public Exception {
public:
    DeviceException(int status, const std::string &msg) : Exception(msg), m_status(status) {}
    int status() const { return m_status; }

private:
    int m_status;
};

// Throws DeviceException when a tdvh_* call returns a negative status.
void checkTdvh(int status, const char *what);
// Throws DeviceException when a HIP runtime call fails.
void checkHip(int hipStatus, const char *what);

class Dim {
public:
    Dim(size_t w = 0, size_t h = 0, size_t d = 1) : m_w(w), m_h(h), m_d(d) {}
    size_t width() const { return m_w; }
    size_t height() const { return m_h; }
    size_t depth() const { return m_d; }
    size_t size() const { return m_w * m_h * m_d; }
    bool operator==(const Dim &o) const { return m_w == o.m_w && m_h == o.m_h && m_d == o.m_d; }
    bool operator!=(const Dim &o) const { return !(*this == o); }

private:
    size_t m_w, m_h, m_d;
};

TDV_NAMESPACE_END

#endif
