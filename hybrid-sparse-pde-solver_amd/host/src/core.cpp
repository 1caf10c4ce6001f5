// =============================================================================
// core.cpp — DeviceContext, Vector (host/device residency), SparseMatrix
// (reference sparse_matrix.cpp / vector.cpp API; SpMV and BLAS-1 on the device).
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <numeric>

#include "core/sparse_matrix.hpp"
#include "core/vector.hpp"

namespace hsps {

// ---- DeviceContext / DeviceBuffer ---------------------------------------------------
DeviceContext::DeviceContext() {
    const char* d = std::getenv("HSPS_DEVICE");
    dev_check(hsps_ctx_create(d ? std::atoi(d) : 0, &ctx_));
}
DeviceContext::~DeviceContext() {
    if (ctx_) hsps_ctx_destroy(ctx_);
}
DeviceContext& DeviceContext::instance() {
    static DeviceContext c;
    return c;
}

static thread_local hsps_ctx tl_ctx = nullptr;  // ThreadDevice binding of this thread

hsps_ctx DeviceContext::current() { return tl_ctx ? tl_ctx : instance().ctx(); }

ThreadDevice::ThreadDevice(hsps_ctx c) : prev_(tl_ctx) { tl_ctx = c; }
ThreadDevice::~ThreadDevice() { tl_ctx = prev_; }

DeviceBuffer::DeviceBuffer(size_t bytes) : bytes_(bytes) {
    if (bytes) {
        ctx_ = DeviceContext::current();
        dev_check(hsps_malloc(ctx_, bytes, &p_));
    }
}
DeviceBuffer::~DeviceBuffer() {
    if (p_) hsps_free(ctx_, p_);
}
DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        if (p_) hsps_free(ctx_, p_);
        ctx_ = o.ctx_;
        p_ = o.p_;
        bytes_ = o.bytes_;
        o.p_ = nullptr;
        o.bytes_ = 0;
    }
    return *this;
}

static hsps_ctx ctx() { return DeviceContext::current(); }

// ---- Vector ---------------------------------------------------------------------------
Vector::Vector(Index n, Real val) : data_(static_cast<Size>(n), val) {}
Vector::Vector(std::initializer_list<Real> il) : data_(il) {}
Vector::Vector(const std::vector<Real>& d) : data_(d) {}
Vector::Vector(const Vector& o) {
    o.host_sync();
    data_ = o.data_;
}
Vector& Vector::operator=(const Vector& o) {
    if (this != &o) {
        o.host_sync();
        data_ = o.data_;
        host_valid_ = true;
        dev_valid_ = false;
    }
    return *this;
}

void Vector::host_sync() const {
    if (!host_valid_) {
        if (!data_.empty()) dev_check(hsps_memcpy_d2h(dev_.ctx(), data_.data(), dev_.get(), data_.size() * sizeof(Real)));
        host_valid_ = true;
    }
}
void Vector::host_mut() {
    host_sync();
    dev_valid_ = false;
}
void Vector::ensure_dev() const {
    const size_t need = data_.size() * sizeof(Real);
    if (dev_.bytes() < need || !dev_.get()) dev_ = DeviceBuffer(std::max<size_t>(need, 8));
}
const double* Vector::device_data() const {
    ensure_dev();
    if (!dev_valid_) {
        if (!data_.empty()) dev_check(hsps_memcpy_h2d(dev_.ctx(), dev_.get(), data_.data(), data_.size() * sizeof(Real)));
        dev_valid_ = true;
    }
    return static_cast<const double*>(dev_.get());
}
double* Vector::device_data_mut() {
    device_data();
    host_valid_ = false;
    return static_cast<double*>(dev_.get());
}
void Vector::assign_device(const double* d_src) {
    ensure_dev();
    if (!data_.empty()) dev_check(hsps_memcpy_d2d(dev_.ctx(), dev_.get(), d_src, data_.size() * sizeof(Real)));
    dev_valid_ = true;
    host_valid_ = false;
}

void Vector::resize(Index n, Real val) {
    data_.assign(static_cast<Size>(n), val);
    host_valid_ = true;
    dev_valid_ = false;
}
void Vector::fill(Real val) {
    host_mut();
    std::fill(data_.begin(), data_.end(), val);
}
void Vector::zero() { fill(REAL_ZERO); }

Real Vector::dot(const Vector& o) const {
    if (o.size() != size()) throw std::runtime_error("Vector::dot — size mismatch");
    if (dev_valid_ && o.dev_valid_ && size() > 0) {
        double r = 0;
        dev_check(hsps_dot_f64(ctx(), size(), device_data(), o.device_data(), &r));
        return r;
    }
    host_sync();
    o.host_sync();
    Real s = 0;
    for (Index i = 0; i < size(); ++i) s += data_[i] * o.data_[i];
    return s;
}
Real Vector::norm2() const { return std::sqrt(dot(*this)); }
Real Vector::norm_inf() const {
    host_sync();
    Real m = 0;
    for (Real v : data_) m = std::max(m, std::abs(v));
    return m;
}
void Vector::scale(Real a) {
    if (dev_valid_ && !host_valid_) {
        dev_check(hsps_scale_f64(ctx(), size(), a, device_data_mut()));
        return;
    }
    host_mut();
    for (Real& v : data_) v *= a;
}
void Vector::axpy(Real a, const Vector& x) {
    if (x.size() != size()) throw std::runtime_error("Vector::axpy — size mismatch");
    if (dev_valid_ && !host_valid_) {
        dev_check(hsps_axpy_f64(ctx(), size(), a, x.device_data(), device_data_mut()));
        return;
    }
    host_mut();
    x.host_sync();
    for (Index i = 0; i < size(); ++i) data_[i] += a * x.data_[i];
}
void Vector::axpby(Real a, const Vector& x, Real b) {
    if (x.size() != size()) throw std::runtime_error("Vector::axpby — size mismatch");
    if (dev_valid_ && !host_valid_) {
        dev_check(hsps_axpby_f64(ctx(), size(), a, x.device_data(), b, device_data_mut()));
        return;
    }
    host_mut();
    x.host_sync();
    for (Index i = 0; i < size(); ++i) data_[i] = a * x.data_[i] + b * data_[i];
}
void Vector::copy_from(const Vector& src) { *this = src; }
Vector Vector::operator+(const Vector& r) const {
    Vector out(*this);
    out.axpy(1.0, r);
    return out;
}
Vector Vector::operator-(const Vector& r) const {
    Vector out(*this);
    out.axpy(-1.0, r);
    return out;
}
Vector Vector::operator*(Real s) const {
    Vector out(*this);
    out.scale(s);
    return out;
}
std::ostream& operator<<(std::ostream& os, const Vector& v) {
    v.host_sync();
    os << "[";
    for (Index i = 0; i < v.size(); ++i) os << v.data_[i] << (i + 1 < v.size() ? ", " : "");
    return os << "]";
}
Real dot(const Vector& a, const Vector& b) { return a.dot(b); }
Real norm2(const Vector& v) { return v.norm2(); }
Vector axpy(Real alpha, const Vector& x, const Vector& y) {
    Vector r(y);
    r.axpy(alpha, x);
    return r;
}

// ---- SparseMatrix ---------------------------------------------------------------------
SparseMatrix::SparseMatrix(Index rows, Index cols) : rows_(rows), cols_(cols), row_ptr_(rows + 1, 0) {}
SparseMatrix::SparseMatrix(Index rows, Index cols, const std::vector<Index>& rp, const std::vector<Index>& ci,
                           const std::vector<Real>& v)
    : rows_(rows), cols_(cols), row_ptr_(rp), col_idx_(ci), values_(v) {}
SparseMatrix::SparseMatrix(const SparseMatrix& o)
    : rows_(o.rows_), cols_(o.cols_), row_ptr_(o.row_ptr_), col_idx_(o.col_idx_), values_(o.values_), dev_(o.dev_) {}
SparseMatrix& SparseMatrix::operator=(const SparseMatrix& o) {
    if (this != &o) {
        rows_ = o.rows_;
        cols_ = o.cols_;
        row_ptr_ = o.row_ptr_;
        col_idx_ = o.col_idx_;
        values_ = o.values_;
        dev_ = o.dev_;
    }
    return *this;
}

hsps_csr SparseMatrix::device() const {
    if (!dev_) {
        hsps_csr h = nullptr;
        std::vector<Index> rp = row_ptr_.empty() ? std::vector<Index>(1, 0) : row_ptr_;
        dev_check(hsps_csr_create(ctx(), rows_, cols_, nnz(), rp.data(), col_idx_.data(), values_.data(), &h));
        dev_ = std::shared_ptr<hsps_csr_s>(h, [](hsps_csr p) { hsps_csr_destroy(p); });
    }
    return dev_.get();
}

double SparseMatrix::density() const {
    if (rows_ == 0 || cols_ == 0) return 0.0;
    return static_cast<double>(nnz()) / (static_cast<double>(rows_) * cols_);
}
Index SparseMatrix::find_entry(Index i, Index j) const {
    auto b = col_idx_.begin() + row_ptr_[i], e = col_idx_.begin() + row_ptr_[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (it == e || *it != j) ? -1 : static_cast<Index>(it - col_idx_.begin());
}
Real SparseMatrix::get(Index i, Index j) const {
    const Index k = find_entry(i, j);
    return k >= 0 ? values_[k] : REAL_ZERO;
}
void SparseMatrix::set(Index i, Index j, Real v) {
    const Index k = find_entry(i, j);
    if (k < 0) throw std::runtime_error("SparseMatrix::set — entry does not exist");
    values_[k] = v;
    dev_.reset();
}
bool SparseMatrix::has_entry(Index i, Index j) const { return find_entry(i, j) >= 0; }

void SparseMatrix::spmv(const Vector& x, Vector& y) const {
    if (x.size() != cols_) throw std::runtime_error("SpMV: x size != cols");
    if (y.size() != rows_) y.resize(rows_, REAL_ZERO);
    if (rows_ == 0) return;
    const double* xd = x.device_data();
    dev_check(hsps_spmv_f64(device(), xd, y.device_data_mut()));
}
Vector SparseMatrix::operator*(const Vector& x) const {
    Vector y(rows_, REAL_ZERO);
    spmv(x, y);
    return y;
}
void SparseMatrix::extract_diagonal(Vector& d) const {
    d.resize(rows_, REAL_ZERO);
    Real* p = d.data();
    for (Index i = 0; i < rows_; ++i)
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k)
            if (col_idx_[k] == i) { p[i] = values_[k]; break; }
}
Vector SparseMatrix::diagonal() const {
    Vector d;
    extract_diagonal(d);
    return d;
}
Real SparseMatrix::spectral_radius_estimate(int iters) const {  // sparse_matrix.cpp:115-133
    Vector v(rows_);
    for (Index i = 0; i < rows_; ++i) v[i] = (i % 2 == 0) ? 1.0 : -0.5;
    Real lambda = 1.0;
    Vector Av(rows_);
    for (int it = 0; it < iters; ++it) {
        spmv(v, Av);
        lambda = Av.norm2();
        if (lambda < REAL_EPS) break;
        Av.scale(1.0 / lambda);
        v = Av;
    }
    return lambda;
}

SparseMatrix transpose(const SparseMatrix& A) {
    std::vector<Index> rp(A.cols_ + 1, 0), ci(A.nnz());
    std::vector<Real> v(A.nnz());
    for (Index k = 0; k < A.nnz(); ++k) rp[A.col_idx_[k] + 1]++;
    for (Index j = 0; j < A.cols_; ++j) rp[j + 1] += rp[j];
    std::vector<Index> pos(rp.begin(), rp.end() - 1);
    for (Index i = 0; i < A.rows_; ++i)
        for (Index k = A.row_ptr_[i]; k < A.row_ptr_[i + 1]; ++k) {
            const Index p = pos[A.col_idx_[k]]++;
            ci[p] = i;
            v[p] = A.values_[k];
        }
    return SparseMatrix(A.cols_, A.rows_, rp, ci, v);
}

SparseMatrix mat_mat_product(const SparseMatrix& A, const SparseMatrix& B) {
    if (B.rows_ != A.cols_) throw std::runtime_error("mat_mat_product: inner dimensions mismatch");
    std::vector<Index> rp(A.rows_ + 1, 0), ci;
    std::vector<Real> v;
    std::vector<Real> acc(B.cols_, 0.0);
    std::vector<char> seen(B.cols_, 0);
    std::vector<Index> touched;
    for (Index i = 0; i < A.rows_; ++i) {
        for (Index pa = A.row_ptr_[i]; pa < A.row_ptr_[i + 1]; ++pa) {
            const Index j = A.col_idx_[pa];
            const Real av = A.values_[pa];
            for (Index pb = B.row_ptr_[j]; pb < B.row_ptr_[j + 1]; ++pb) {
                const Index l = B.col_idx_[pb];
                if (!seen[l]) { seen[l] = 1; touched.push_back(l); }
                acc[l] = std::fma(av, B.values_[pb], acc[l]);
            }
        }
        std::sort(touched.begin(), touched.end());
        for (Index l : touched) {
            if (std::abs(acc[l]) > REAL_EPS * REAL_EPS) { ci.push_back(l); v.push_back(acc[l]); }
            acc[l] = 0.0;
            seen[l] = 0;
        }
        touched.clear();
        rp[i + 1] = static_cast<Index>(ci.size());
    }
    return SparseMatrix(A.rows_, B.cols_, rp, ci, v);
}

SparseMatrix SparseMatrix::from_coo(Index rows, Index cols, const std::vector<Index>& r, const std::vector<Index>& c,
                                    const std::vector<Real>& val) {
    std::vector<Size> ord(r.size());
    std::iota(ord.begin(), ord.end(), Size(0));
    std::stable_sort(ord.begin(), ord.end(), [&](Size a, Size b) { return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b]; });
    std::vector<Index> rp(rows + 1, 0), ci;
    std::vector<Real> v;
    ci.reserve(r.size());
    v.reserve(r.size());
    Index lr = -1, lc = -1;
    for (Size o : ord) {
        if (r[o] == lr && c[o] == lc) {
            v.back() += val[o];  // duplicates accumulate (sparse_matrix.cpp:264-268)
        } else {
            ci.push_back(c[o]);
            v.push_back(val[o]);
            rp[r[o] + 1]++;
            lr = r[o];
            lc = c[o];
        }
    }
    for (Index i = 0; i < rows; ++i) rp[i + 1] += rp[i];
    return SparseMatrix(rows, cols, rp, ci, v);
}

// 5-point builder shared by the generators: entries of row (i,j) in ascending
// column order (i-1,j), (i,j-1), (i,j), (i,j+1), (i+1,j).
template <class F>
static SparseMatrix five_point(Index n, F coef) {
    const Index N = n * n;
    std::vector<Index> rp(N + 1, 0), ci;
    std::vector<Real> v;
    ci.reserve(5 * (size_t)N);
    v.reserve(5 * (size_t)N);
    for (Index i = 0; i < n; ++i)
        for (Index j = 0; j < n; ++j) {
            Real w, e, s, no, d;
            coef(i, j, d, w, e, s, no);
            const Index row = i * n + j;
            if (i > 0) { ci.push_back(row - n); v.push_back(s); }
            if (j > 0) { ci.push_back(row - 1); v.push_back(w); }
            ci.push_back(row);
            v.push_back(d);
            if (j < n - 1) { ci.push_back(row + 1); v.push_back(e); }
            if (i < n - 1) { ci.push_back(row + n); v.push_back(no); }
            rp[row + 1] = static_cast<Index>(ci.size());
        }
    return SparseMatrix(N, N, rp, ci, v);
}

SparseMatrix SparseMatrix::poisson_2d(Index n) {  // sparse_matrix.cpp:288-311
    return five_point(n, [](Index, Index, Real& d, Real& w, Real& e, Real& s, Real& no) {
        d = 4.0; w = e = s = no = -1.0;
    });
}

SparseMatrix SparseMatrix::convection_diffusion_2d(Index n, Real eps, Real bx, Real by) {  // :317-350
    const Real h = 1.0 / (n + 1), eh = eps / (h * h), bxh = bx / (2.0 * h), byh = by / (2.0 * h);
    return five_point(n, [&](Index i, Index j, Real& d, Real& w, Real& e, Real& s, Real& no) {
        d = 4.0 * eh;
        w = -eh + std::max(bxh, 0.0);
        e = -eh - std::min(bxh, 0.0);
        s = -eh + std::max(byh, 0.0);
        no = -eh - std::min(byh, 0.0);
        // diagonal accumulates in the reference's push order: x- then x+, y- then y+
        if (j > 0) d += std::max(-bxh, 0.0);
        if (j < n - 1) d += std::max(bxh, 0.0);
        if (i > 0) d += std::max(-byh, 0.0);
        if (i < n - 1) d += std::max(byh, 0.0);
    });
}

SparseMatrix SparseMatrix::anisotropic_2d(Index n, Real eps) {
    return five_point(n, [&](Index, Index, Real& d, Real& w, Real& e, Real& s, Real& no) {
        d = 2.0 * eps + 2.0; w = e = -eps; s = no = -1.0;
    });
}

bool SparseMatrix::is_symmetric(Real tol) const {
    if (rows_ != cols_) return false;
    for (Index i = 0; i < rows_; ++i)
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k) {
            const Real v = values_[k];
            if (std::abs(v - get(col_idx_[k], i)) > tol * (1.0 + std::abs(v))) return false;
        }
    return true;
}
Real SparseMatrix::frobenius_norm() const {
    Real s = 0;
    for (Real v : values_) s += v * v;
    return std::sqrt(s);
}
void SparseMatrix::print_info(std::ostream& os) const {
    os << "SparseMatrix [" << rows_ << " × " << cols_ << "]  nnz=" << nnz() << "  density=" << std::scientific
       << std::setprecision(3) << density() << "  mem=" << mem_bytes() / 1024 << " KB\n";
}
long long SparseMatrix::mem_bytes() const {
    return (long long)row_ptr_.size() * sizeof(Index) + (long long)col_idx_.size() * sizeof(Index) +
           (long long)values_.size() * sizeof(Real);
}
Index SparseMatrix::bandwidth() const {
    Index bw = 0;
    for (Index i = 0; i < rows_; ++i)
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k) bw = std::max(bw, std::abs(i - col_idx_[k]));
    return bw;
}
double SparseMatrix::nnz_variance() const {
    if (rows_ == 0) return 0.0;
    const double mean = static_cast<double>(nnz()) / rows_;
    double var = 0;
    for (Index i = 0; i < rows_; ++i) {
        const double d = (row_ptr_[i + 1] - row_ptr_[i]) - mean;
        var += d * d;
    }
    return var / rows_;
}
double SparseMatrix::diagonal_fraction() const {
    if (nnz() == 0) return 0.0;
    int c = 0;
    for (Index i = 0; i < rows_; ++i)
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k)
            if (col_idx_[k] == i) { ++c; break; }
    return static_cast<double>(c) / nnz();
}
double SparseMatrix::lower_upper_ratio() const {
    long long lo = 0, hi = 0;
    for (Index i = 0; i < rows_; ++i)
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k) {
            if (col_idx_[k] < i) ++lo;
            if (col_idx_[k] > i) ++hi;
        }
    return hi > 0 ? static_cast<double>(lo) / hi : 1.0;
}
double SparseMatrix::estimated_amg_coarsening_ratio(Real th) const {
    double strong = 0;
    for (Index i = 0; i < rows_; ++i) {
        Real mx = 0;
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k)
            if (col_idx_[k] != i) mx = std::max(mx, std::abs(values_[k]));
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k)
            if (col_idx_[k] != i && std::abs(values_[k]) >= th * mx) strong += 1.0;
    }
    return 1.0 / (1.0 + (rows_ > 0 ? strong / rows_ : 0.0));
}
std::pair<Real, Real> SparseMatrix::spectral_interval(int steps) const {
    const int K = std::min(rows_, steps);
    Vector q(rows_, 1.0 / std::sqrt(static_cast<Real>(rows_))), qp(rows_, 0.0), Aq(rows_);
    std::vector<Real> al, be;
    Real bj = 0;
    for (int j = 0; j < K; ++j) {
        spmv(q, Aq);
        const Real aj = q.dot(Aq);
        al.push_back(aj);
        {
            Real* a = Aq.data();
            const Real* qq = q.data();
            const Real* pp = qp.data();
            for (Index i = 0; i < rows_; ++i) a[i] -= aj * qq[i] + bj * pp[i];
        }
        bj = Aq.norm2();
        if (bj < REAL_EPS * 1e6) break;
        be.push_back(bj);
        qp = q;
        q = Aq;
        q.scale(1.0 / bj);
    }
    const int sz = static_cast<int>(al.size());
    Real lo = al[0], hi = al[0];
    for (int i = 0; i < sz; ++i) {
        Real r = 0;
        if (i > 0) r += std::abs(be[i - 1]);
        if (i < sz - 1) r += std::abs(be[i]);
        lo = std::min(lo, al[i] - r);
        hi = std::max(hi, al[i] + r);
    }
    return {lo, hi};
}
int SparseMatrix::count_zero_diagonal(Real tol) const {
    int c = 0;
    for (Index i = 0; i < rows_; ++i) {
        bool found = false;
        for (Index k = row_ptr_[i]; k < row_ptr_[i + 1]; ++k)
            if (col_idx_[k] == i) {
                if (std::abs(values_[k]) <= tol) ++c;
                found = true;
                break;
            }
        if (!found) ++c;
    }
    return c;
}

}  // namespace hsps
