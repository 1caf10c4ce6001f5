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

/// One hsps_ctx per process (device 0 unless HSPS_DEVICE is set).
class DeviceContext {
public:
    static DeviceContext& instance();
    hsps_ctx ctx() const { return ctx_; }
    ~DeviceContext();
    DeviceContext(const DeviceContext&) = delete;
    DeviceContext& operator=(const DeviceContext&) = delete;

private:
    DeviceContext();
    hsps_ctx ctx_ = nullptr;
};

/// Owned device allocation (hsps_malloc / hsps_free).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes);
    ~DeviceBuffer();
    DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void* get() const { return p_; }
    size_t bytes() const { return bytes_; }

private:
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace hsps
