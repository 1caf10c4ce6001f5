#pragma once
// =============================================================================
// device.hpp — the host layer's link to the C ABI (hsps_dev.h): a process-wide
// context per device, and status-code -> exception mapping that raises the
// reference's exception types (std::runtime_error for size / diagonal /
// pivot / runtime failures, std::invalid_argument for bad arguments).
// =============================================================================
#include <memory>
#include <stdexcept>
#include <string>

#include "hsps_dev.h"

namespace hsps {

/// Throws the exception type the reference raises in the same situation.
inline void dev_check(int rc) {
    if (rc >= 0) return;
    std::string msg = hsps_last_error();
    if (rc == HSPS_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

/// One hsps_ctx per process (device 0 unless HSPS_DEVICE is set).  A thread
/// may bind its own context instead (ThreadDevice): current() is what the
/// host classes create device objects on.
class DeviceContext {
public:
    static DeviceContext& instance();
    static hsps_ctx current();
    hsps_ctx ctx() const { return ctx_; }
    ~DeviceContext();
    DeviceContext(const DeviceContext&) = delete;
    DeviceContext& operator=(const DeviceContext&) = delete;

private:
    DeviceContext();
    hsps_ctx ctx_ = nullptr;
};

/// Binds a context (its stream and scratch) to the calling thread for the
/// lifetime of the object; objects created meanwhile live on that context.
/// No reference counterpart: used by the solve server's workers, which run
/// solves concurrently, one context and stream each.
class ThreadDevice {
public:
    explicit ThreadDevice(hsps_ctx c);
    ~ThreadDevice();
    ThreadDevice(const ThreadDevice&) = delete;
    ThreadDevice& operator=(const ThreadDevice&) = delete;

private:
    hsps_ctx prev_;
};

/// Owned device allocation (hsps_malloc / hsps_free) on the context current
/// at allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer&& o) noexcept : ctx_(o.ctx_), p_(o.p_), bytes_(o.bytes_) {
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void* get() const { return p_; }
    size_t bytes() const { return bytes_; }
    hsps_ctx ctx() const { return ctx_; }

private:
    hsps_ctx ctx_ = nullptr;
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace hsps
