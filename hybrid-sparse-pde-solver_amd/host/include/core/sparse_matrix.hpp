#pragma once
// =============================================================================
// sparse_matrix.hpp — CSR matrix with the reference's API
// (sparse_matrix.hpp:34-170): int32 row_ptr / col_idx, FP64 values, columns
// sorted.  The device copy (hsps_csr: FP64 + FP32 values, diagonal, row-block
// SpMV schedule) is created on first device use and dropped when the values
// are modified through the mutable accessors.
// =============================================================================
#include <memory>
#include <ostream>
#include <string>
#include <utility>
#include <vector>

#include "types.hpp"
#include "vector.hpp"

namespace hsps {

class SparseMatrix;
SparseMatrix transpose(const SparseMatrix& A);
SparseMatrix mat_mat_product(const SparseMatrix& A, const SparseMatrix& B);

class SparseMatrix {
public:
    SparseMatrix() = default;
    SparseMatrix(Index rows, Index cols);
    SparseMatrix(Index rows, Index cols, const std::vector<Index>& row_ptr, const std::vector<Index>& col_idx,
                 const std::vector<Real>& values);
    SparseMatrix(const SparseMatrix& o);
    SparseMatrix(SparseMatrix&&) noexcept = default;
    SparseMatrix& operator=(const SparseMatrix& o);
    SparseMatrix& operator=(SparseMatrix&&) noexcept = default;
    ~SparseMatrix() = default;

    Index rows() const noexcept { return rows_; }
    Index cols() const noexcept { return cols_; }
    Index nnz() const noexcept { return static_cast<Index>(values_.size()); }
    bool empty() const noexcept { return values_.empty(); }
    double density() const;

    Real get(Index i, Index j) const;
    void set(Index i, Index j, Real v);
    bool has_entry(Index i, Index j) const;

    const std::vector<Index>& row_ptr() const noexcept { return row_ptr_; }
    const std::vector<Index>& col_idx() const noexcept { return col_idx_; }
    const std::vector<Real>& values() const noexcept { return values_; }
    std::vector<Real>& values() noexcept { dev_.reset(); return values_; }

    /// y = A x on the device (sparse_matrix.cpp:68-87 semantics: throws on
    /// x.size() != cols, resizes y).
    void spmv(const Vector& x, Vector& y) const;
    Vector operator*(const Vector& x) const;

    void extract_diagonal(Vector& diag) const;
    Vector diagonal() const;
    Real spectral_radius_estimate(int power_iters = 20) const;

    friend SparseMatrix mat_mat_product(const SparseMatrix& A, const SparseMatrix& B);
    friend SparseMatrix transpose(const SparseMatrix& A);

    static SparseMatrix from_coo(Index rows, Index cols, const std::vector<Index>& row_vec,
                                 const std::vector<Index>& col_vec, const std::vector<Real>& val_vec);
    static SparseMatrix poisson_2d(Index n);
    static SparseMatrix convection_diffusion_2d(Index n, Real epsilon = 1.0, Real beta_x = 0.0, Real beta_y = 0.0);
    /// -eps u_xx - u_yy, 5-point, unscaled (no reference counterpart).
    static SparseMatrix anisotropic_2d(Index n, Real epsilon);

    bool is_symmetric(Real tol = 1e-10) const;
    Real frobenius_norm() const;
    void print_info(std::ostream& os) const;
    long long mem_bytes() const;
    Index bandwidth() const;
    double nnz_variance() const;
    double diagonal_fraction() const;
    double lower_upper_ratio() const;
    double estimated_ilu_fill() const { return 1.0; }
    double estimated_amg_coarsening_ratio(Real strength_thresh = 0.25) const;
    std::pair<Real, Real> spectral_interval(int lanczos_steps = 20) const;
    int count_zero_diagonal(Real tol = 1e-15) const;

    /// Device handle (created on demand; shared by copies until modified).
    hsps_csr device() const;
    /// Drops this object's reference to the device copy (the next device use
    /// uploads again); no reference counterpart.
    void release_device() const { dev_.reset(); }

private:
    Index rows_ = 0;
    Index cols_ = 0;
    std::vector<Index> row_ptr_;
    std::vector<Index> col_idx_;
    std::vector<Real> values_;
    Index find_entry(Index i, Index j) const;
    mutable std::shared_ptr<struct hsps_csr_s> dev_;
};

}  // namespace hsps
