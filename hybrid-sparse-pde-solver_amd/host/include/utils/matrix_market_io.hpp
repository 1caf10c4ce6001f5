#pragma once
// matrix_market_io.hpp — MatrixMarket reader/writer (reference
// matrix_market_io.hpp:80-130): coordinate matrices with symmetric /
// skew-symmetric expansion, array vectors.
#include <string>
#include <utility>
#include <vector>

#include "../core/sparse_matrix.hpp"
#include "../core/vector.hpp"

namespace hsps {

struct MatrixMarketHeader {
    enum class Format { COORDINATE, ARRAY };
    enum class Field { REAL, INTEGER, COMPLEX, PATTERN };
    enum class Symmetry { GENERAL, SYMMETRIC, SKEW_SYMMETRIC, HERMITIAN };
    Format format = Format::COORDINATE;
    Field field = Field::REAL;
    Symmetry symmetry = Symmetry::GENERAL;
    int rows = 0, cols = 0, nnz = 0, nnz_full = 0;
    bool is_square() const { return rows == cols; }
    bool is_symmetric() const { return symmetry == Symmetry::SYMMETRIC; }
};

class MatrixMarketIO {
public:
    static SparseMatrix load_matrix(const std::string& path);
    static Vector load_vector(const std::string& path, int expected_n = 0);
    static MatrixMarketHeader read_header(const std::string& path);
    static std::pair<SparseMatrix, Vector> load_problem(const std::string& matrix_path,
                                                        const std::string& rhs_path = "");
    static Vector make_default_rhs(int n, Real value = 1.0);
    static void save_vector(const std::string& path, const Vector& v, const std::string& comment = "");
    static void save_matrix(const std::string& path, const SparseMatrix& A, const std::string& comment = "");
    static std::vector<std::string> list_mtx_files(const std::string& dir);

private:
    static MatrixMarketHeader parse_header_line(const std::string& line);
};

}  // namespace hsps
