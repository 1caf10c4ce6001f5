// SparseMatrix / Vector API through the device path; cases follow the
// reference's tests/test_sparse_matrix.cpp:28-117.
#include "core/sparse_matrix.hpp"
#include "hsps_test.hpp"

using namespace hsps;

TEST(spmv_identity) {
    auto I = SparseMatrix::from_coo(3, 3, {0, 1, 2}, {0, 1, 2}, {1.0, 1.0, 1.0});
    Vector x = {1.0, 2.0, 3.0}, y;
    I.spmv(x, y);
    for (int i = 0; i < 3; ++i) expect_near(y[i], x[i], 0.0, "I x = x");
}
TEST(spmv_tridiagonal) {
    auto T = SparseMatrix::from_coo(4, 4, {0, 0, 1, 1, 1, 2, 2, 2, 3, 3}, {0, 1, 0, 1, 2, 1, 2, 3, 2, 3},
                                    {2, -1, -1, 2, -1, -1, 2, -1, -1, 2});
    Vector y = T * Vector(4, 1.0);
    expect_near(y[0], 1, 0, "y0"); expect_near(y[1], 0, 0, "y1"); expect_near(y[2], 0, 0, "y2"); expect_near(y[3], 1, 0, "y3");
}
TEST(poisson_symmetric_diag4) {
    auto A = SparseMatrix::poisson_2d(10);
    expect_true(A.rows() == 100 && A.nnz() == 5 * 100 - 4 * 10, "nnz = 5N - 4n");
    expect_true(A.is_symmetric(), "symmetric");
    for (int i = 0; i < 100; ++i) expect_near(A.get(i, i), 4.0, 0, "diag 4");
}
TEST(transpose_and_product) {
    auto A = SparseMatrix::from_coo(2, 3, {0, 0, 1}, {0, 2, 1}, {1.0, 2.0, 3.0});
    auto T = transpose(A);
    expect_true(T.rows() == 3 && T.cols() == 2, "shape");
    expect_near(T.get(2, 0), 2.0, 0, "T(2,0)");
    auto I = SparseMatrix::from_coo(3, 3, {0, 1, 2}, {0, 1, 2}, {1, 1, 1});
    auto II = mat_mat_product(I, I);
    expect_true(II.nnz() == 3, "I*I nnz");
}
TEST(coo_duplicates_accumulate) {
    auto A = SparseMatrix::from_coo(2, 2, {0, 0, 1}, {0, 0, 1}, {1.0, 2.0, 5.0});
    expect_near(A.get(0, 0), 3.0, 0, "1+2");
}
TEST(convdiff_nonsymmetric) {
    auto A = SparseMatrix::convection_diffusion_2d(8, 0.1, 1.0, 0.5);
    expect_true(!A.is_symmetric(), "non-symmetric");
}
TEST(spmv_size_mismatch_throws) {
    auto A = SparseMatrix::poisson_2d(4);
    Vector x(3), y;
    expect_throw<std::runtime_error>([&] { A.spmv(x, y); }, "x size != cols");
}
TEST(vector_blas_and_residency) {
    Vector a = {1, 2, 3}, b = {4, 5, 6};
    expect_near(a.dot(b), 32, 0, "dot");
    a.axpy(2.0, b);
    expect_near(a[2], 15, 0, "axpy");
    auto A = SparseMatrix::poisson_2d(16);
    Vector x(256, 1.0), y;
    A.spmv(x, y);  // y now lives on the device
    expect_true(y.device_current() && !y.host_current(), "result stays on the device");
    expect_near(y.dot(y), 56 * 1.0 + 4 * 4.0, 1e-12, "56 edge rows give 1, 4 corners give 2");
}
int main() { return run_all(); }
