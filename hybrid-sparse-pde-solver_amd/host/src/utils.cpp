// =============================================================================
// utils.cpp — MatrixMarket I/O (matrix_market_io.cpp:45-376 behaviour):
// coordinate matrices with symmetric / skew-symmetric expansion through
// from_coo (duplicates accumulate), array vectors padded with 1.0 to the
// expected size, unit RHS when none is given.
// =============================================================================
#include <algorithm>
#include <cctype>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "utils/matrix_market_io.hpp"

namespace hsps {

static std::ifstream open_in(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("MatrixMarketIO: cannot open " + path);
    return f;
}

MatrixMarketHeader MatrixMarketIO::parse_header_line(const std::string& line) {
    std::istringstream iss(line);
    std::vector<std::string> t;
    std::string w;
    while (iss >> w) {
        std::transform(w.begin(), w.end(), w.begin(), [](unsigned char c) { return std::tolower(c); });
        t.push_back(w);
    }
    if (t.size() < 4 || t[0] != "%%matrixmarket") throw std::runtime_error("MatrixMarketIO: bad banner: " + line);
    MatrixMarketHeader h;
    h.format = t[2] == "array" ? MatrixMarketHeader::Format::ARRAY : MatrixMarketHeader::Format::COORDINATE;
    if (t[3] == "integer") h.field = MatrixMarketHeader::Field::INTEGER;
    else if (t[3] == "complex") h.field = MatrixMarketHeader::Field::COMPLEX;
    else if (t[3] == "pattern") h.field = MatrixMarketHeader::Field::PATTERN;
    if (t.size() > 4) {
        if (t[4] == "symmetric") h.symmetry = MatrixMarketHeader::Symmetry::SYMMETRIC;
        else if (t[4] == "skew-symmetric") h.symmetry = MatrixMarketHeader::Symmetry::SKEW_SYMMETRIC;
        else if (t[4] == "hermitian") h.symmetry = MatrixMarketHeader::Symmetry::HERMITIAN;
    }
    return h;
}

MatrixMarketHeader MatrixMarketIO::read_header(const std::string& path) {
    auto f = open_in(path);
    std::string line;
    std::getline(f, line);
    MatrixMarketHeader h = parse_header_line(line);
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream iss(line);
        if (h.format == MatrixMarketHeader::Format::COORDINATE) {
            iss >> h.rows >> h.cols >> h.nnz;
            h.nnz_full = h.symmetry == MatrixMarketHeader::Symmetry::GENERAL ? h.nnz : h.nnz * 2 - h.rows;
        } else {
            iss >> h.rows >> h.cols;
            h.nnz = h.nnz_full = h.rows * h.cols;
        }
        break;
    }
    return h;
}

SparseMatrix MatrixMarketIO::load_matrix(const std::string& path) {
    auto f = open_in(path);
    std::string line;
    std::getline(f, line);
    if (line.size() < 2 || line[0] != '%' || line[1] != '%')
        throw std::runtime_error(path + ": first line is not a MatrixMarket header (expected %%MatrixMarket)");
    MatrixMarketHeader h = parse_header_line(line);
    if (h.format != MatrixMarketHeader::Format::COORDINATE)
        throw std::runtime_error(path + ": dense (array) MatrixMarket file; a sparse matrix needs the coordinate format");
    if (h.field == MatrixMarketHeader::Field::COMPLEX)
        throw std::runtime_error(path + ": complex-valued MatrixMarket matrix; only real, integer and pattern fields are read");
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream iss(line);
        iss >> h.rows >> h.cols >> h.nnz;
        break;
    }
    std::vector<Index> r, c;
    std::vector<Real> v;
    r.reserve(2 * (size_t)h.nnz);
    c.reserve(2 * (size_t)h.nnz);
    v.reserve(2 * (size_t)h.nnz);
    int read = 0;
    while (read < h.nnz && std::getline(f, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream iss(line);
        int i, j;
        double x = 1.0;
        iss >> i >> j;
        if (h.field != MatrixMarketHeader::Field::PATTERN) iss >> x;
        --i;
        --j;
        r.push_back(i); c.push_back(j); v.push_back(x);
        if (i != j && h.symmetry == MatrixMarketHeader::Symmetry::SYMMETRIC) { r.push_back(j); c.push_back(i); v.push_back(x); }
        if (i != j && h.symmetry == MatrixMarketHeader::Symmetry::SKEW_SYMMETRIC) { r.push_back(j); c.push_back(i); v.push_back(-x); }
        ++read;
    }
    return SparseMatrix::from_coo(h.rows, h.cols, r, c, v);
}

Vector MatrixMarketIO::load_vector(const std::string& path, int expected_n) {
    auto f = open_in(path);
    std::string line;
    std::vector<Real> vals;
    int n_entries = -1;
    bool banner = false;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        if (line[0] == '%') { banner = true; continue; }
        if (banner && n_entries < 0) {
            std::istringstream iss(line);
            int rows = 0, cols = 1;
            iss >> rows >> cols;
            n_entries = rows * std::max(cols, 1);
            continue;
        }
        std::istringstream iss(line);
        double x;
        while (iss >> x) vals.push_back(x);
    }
    if (n_entries > 0 && (int)vals.size() > n_entries) vals.resize(n_entries);
    if (expected_n > 0 && (int)vals.size() != expected_n) vals.resize(expected_n, REAL_ONE);
    return Vector(vals);
}

Vector MatrixMarketIO::make_default_rhs(int n, Real value) { return Vector(n, value); }

std::pair<SparseMatrix, Vector> MatrixMarketIO::load_problem(const std::string& mp, const std::string& rp) {
    SparseMatrix A = load_matrix(mp);
    Vector b;
    if (!rp.empty()) {
        try {
            b = load_vector(rp, A.rows());
            if (b.size() != A.rows()) b = make_default_rhs(A.rows());
        } catch (const std::exception&) {
            b = make_default_rhs(A.rows());
        }
    } else {
        b = make_default_rhs(A.rows());
    }
    return {std::move(A), std::move(b)};
}

void MatrixMarketIO::save_vector(const std::string& path, const Vector& v, const std::string& comment) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("MatrixMarketIO: cannot write " + path);
    f << "%%MatrixMarket matrix array real general\n";
    if (!comment.empty()) f << "% " << comment << "\n";
    f << v.size() << " 1\n" << std::setprecision(17);
    const Real* d = v.data();
    for (Index i = 0; i < v.size(); ++i) f << d[i] << "\n";
}

void MatrixMarketIO::save_matrix(const std::string& path, const SparseMatrix& A, const std::string& comment) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("MatrixMarketIO: cannot write " + path);
    f << "%%MatrixMarket matrix coordinate real general\n";
    if (!comment.empty()) f << "% " << comment << "\n";
    f << A.rows() << " " << A.cols() << " " << A.nnz() << "\n" << std::setprecision(17);
    for (Index i = 0; i < A.rows(); ++i)
        for (Index k = A.row_ptr()[i]; k < A.row_ptr()[i + 1]; ++k)
            f << i + 1 << " " << A.col_idx()[k] + 1 << " " << A.values()[k] << "\n";
}

std::vector<std::string> MatrixMarketIO::list_mtx_files(const std::string& dir) {
    std::vector<std::string> out;
    for (const auto& e : std::filesystem::directory_iterator(dir)) {
        const auto p = e.path().string();
        if (e.path().extension() == ".mtx" && p.find("_rhs") == std::string::npos) out.push_back(p);
    }
    std::sort(out.begin(), out.end());
    return out;
}

}  // namespace hsps
