#pragma once
// =============================================================================
// vector.hpp — dense FP64 vector with the reference's API (vector.hpp:16-70)
// and a lazily synchronised device mirror.
//
// Residency: host and device copies each carry a valid bit.  Host accessors
// download first when only the device copy is current; mutable host access
// invalidates the device copy.  device_data() uploads when needed;
// device_data_mut() marks the device copy as the only current one.  Solvers
// therefore keep x on the device across calls (the adaptive ladder's warm
// start never round-trips through the host).
// =============================================================================
#include <initializer_list>
#include <ostream>
#include <vector>

#include "device.hpp"
#include "types.hpp"

namespace hsps {

class Vector {
public:
    Vector() = default;
    explicit Vector(Index n, Real val = REAL_ZERO);
    Vector(std::initializer_list<Real> il);
    Vector(const std::vector<Real>& data);
    Vector(const Vector& o);
    Vector(Vector&&) noexcept = default;
    Vector& operator=(const Vector& o);
    Vector& operator=(Vector&&) noexcept = default;
    ~Vector() = default;

    Index size() const noexcept { return static_cast<Index>(data_.size()); }
    bool empty() const noexcept { return data_.empty(); }
    Real& operator[](Index i) { host_mut(); return data_[i]; }
    const Real& operator[](Index i) const { host_sync(); return data_[i]; }
    Real* data() { host_mut(); return data_.data(); }
    const Real* data() const { host_sync(); return data_.data(); }

    void resize(Index n, Real val = REAL_ZERO);  ///< assigns n copies of val (vector.cpp:26-28)
    void fill(Real val);
    void zero();

    // BLAS-1 (vector.cpp:39-119); run on the device when the device copy is
    // current, on the host otherwise.
    Real dot(const Vector& other) const;
    Real norm2() const;
    Real norm_inf() const;
    void scale(Real alpha);
    void axpy(Real alpha, const Vector& x);
    void axpby(Real a, const Vector& x, Real b);
    void copy_from(const Vector& src);
    Vector operator+(const Vector& rhs) const;
    Vector operator-(const Vector& rhs) const;
    Vector operator*(Real scalar) const;

    long long mem_bytes() const { return static_cast<long long>(data_.size()) * sizeof(Real); }
    friend std::ostream& operator<<(std::ostream& os, const Vector& v);

    // ---- device residency --------------------------------------------------------
    const double* device_data() const;  ///< uploads if the host copy is newer
    double* device_data_mut();          ///< device copy becomes the current one
    /// The device copy becomes size() doubles copied from d_src (device
    /// memory on this device, already written), enqueued on this vector's
    /// context stream; the host copy becomes stale.
    void assign_device(const double* d_src);
    bool device_current() const { return dev_valid_; }
    bool host_current() const { return host_valid_; }

private:
    void host_sync() const;  ///< download if needed
    void host_mut();         ///< download if needed, then invalidate device
    void ensure_dev() const;

    mutable std::vector<Real> data_;
    mutable DeviceBuffer dev_;
    mutable bool host_valid_ = true;
    mutable bool dev_valid_ = false;
};

Real dot(const Vector& a, const Vector& b);
Real norm2(const Vector& v);
Vector axpy(Real alpha, const Vector& x, const Vector& y);

}  // namespace hsps
