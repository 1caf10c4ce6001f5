// =============================================================================
// host_setup.cpp — host-side setup stages (see host_setup.hpp).
// =============================================================================
#include "host_setup.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <utility>

namespace hsps_host {

static constexpr double kEps = std::numeric_limits<double>::epsilon();

std::vector<int> aggregate_strong(int n, const std::vector<int>& srp, const std::vector<int>& sci, int* n_coarse) {
    std::vector<int> agg(n, -1);
    int nagg = 0;
    for (int i = 0; i < n; ++i) {
        if (agg[i] >= 0) continue;
        agg[i] = nagg;
        for (int k = srp[i]; k < srp[i + 1]; ++k) {
            const int j = sci[k];
            if (agg[j] < 0) agg[j] = nagg;
        }
        ++nagg;
    }
    if (n_coarse) *n_coarse = nagg;
    return agg;
}

HostCsr ilu0(const HostCsr& A) {
    HostCsr F = A;
    const int n = A.rows;
    std::vector<int> diag(n, -1), pos(n, -1);
    for (int i = 0; i < n; ++i)
        for (int k = F.rp[i]; k < F.rp[i + 1]; ++k)
            if (F.ci[k] == i) { diag[i] = k; break; }
    for (int i = 1; i < n; ++i) {
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p) pos[F.ci[p]] = p;
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p) {
            const int k = F.ci[p];
            if (k >= i) break;
            const double dk = diag[k] >= 0 ? F.v[diag[k]] : 0.0;
            if (std::abs(dk) < kEps) throw std::runtime_error("ILU(0): near-zero pivot at row " + std::to_string(k));
            F.v[p] /= dk;
            const double lik = F.v[p];
            for (int q = F.rp[k]; q < F.rp[k + 1]; ++q) {
                const int j = F.ci[q];
                if (j <= k) continue;
                const int pij = pos[j];
                if (pij >= 0) F.v[pij] -= lik * F.v[q];
            }
        }
        for (int p = F.rp[i]; p < F.rp[i + 1]; ++p) pos[F.ci[p]] = -1;
    }
    return F;
}

// ILUT(tau, p), ilu.cpp:94-220.  The reference scans every column k < i of
// a dense work row; the same arithmetic in the same order visits only the
// active columns, taken from a min-heap (an elimination only activates
// columns j > k, so the pops come out in increasing order, each column once),
// and resets only the columns the row touched.  Same factors, bit for bit;
// O(nnz(L\U) * row length * log) instead of O(n^2).
HostCsr ilut(const HostCsr& A, double tau_rel, int fill_p) {
    const int n = A.rows;
    std::vector<std::vector<std::pair<int, double>>> L(n), U(n);
    std::vector<double> w(n, 0.0), udiag(n, 0.0);
    std::vector<char> act(n, 0), seen(n, 0);
    std::vector<int> touched, heap, upper;
    auto touch = [&](int j) {
        if (!seen[j]) { seen[j] = 1; touched.push_back(j); }
    };
    auto push = [&](int j) {  // min-heap of pending pivot columns
        heap.push_back(j);
        std::push_heap(heap.begin(), heap.end(), std::greater<int>());
    };
    for (int i = 0; i < n; ++i) {
        touched.clear();
        double rn = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const int j = A.ci[p];
            w[j] = A.v[p];
            act[j] = 1;
            touch(j);
            rn += A.v[p] * A.v[p];
            if (j < i) push(j);
        }
        const double tau = tau_rel * std::sqrt(rn);
        std::vector<std::pair<int, double>> Lr, Ur;
        // pivot columns k < i in increasing order (ilu.cpp:121-146)
        while (!heap.empty()) {
            std::pop_heap(heap.begin(), heap.end(), std::greater<int>());
            const int k = heap.back();
            heap.pop_back();
            if (!act[k]) continue;
            if (std::abs(w[k]) < tau) { w[k] = 0.0; act[k] = 0; continue; }
            const double ukk = udiag[k];
            if (std::abs(ukk) < kEps) { act[k] = 0; continue; }
            const double lik = w[k] / ukk;
            w[k] = lik;
            for (const auto& e : U[k]) {
                const int j = e.first;
                if (j <= k) continue;
                if (!act[j] && std::abs(lik * e.second) >= tau) {
                    act[j] = 1;
                    if (j < i) push(j);
                }
                w[j] -= lik * e.second;
                touch(j);
            }
            if (act[k] && std::abs(w[k]) >= tau) Lr.emplace_back(k, w[k]);  // ilu.cpp:151-155, ascending k
        }
        double d = w[i];
        if (std::abs(d) < kEps) d = (d >= 0 ? 1.0 : -1.0) * kEps * 1e6;
        Ur.emplace_back(i, d);
        upper.clear();
        for (int j : touched)
            if (j > i && act[j] && std::abs(w[j]) >= tau) upper.push_back(j);
        std::sort(upper.begin(), upper.end());
        for (int j : upper) Ur.emplace_back(j, w[j]);
        for (int j : touched) {
            w[j] = 0.0;
            act[j] = 0;
            seen[j] = 0;
        }
        w[i] = 0.0;
        act[i] = 0;
        const int keep = A.rp[i + 1] - A.rp[i] + fill_p;
        auto limit = [keep](std::vector<std::pair<int, double>>& row) {
            if (keep <= 0 || (int)row.size() <= keep) return;
            std::partial_sort(row.begin(), row.begin() + keep, row.end(),
                              [](const auto& a, const auto& b) { return std::abs(a.second) > std::abs(b.second); });
            row.resize(keep);
            std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        };
        limit(Lr);
        limit(Ur);
        udiag[i] = 0.0;
        for (const auto& e : Ur)
            if (e.first == i) { udiag[i] = e.second; break; }
        L[i] = std::move(Lr);
        U[i] = std::move(Ur);
    }
    HostCsr F;
    F.rows = F.cols = n;
    F.rp.assign(n + 1, 0);
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += L[i].size() + U[i].size();
    F.ci.reserve(total);
    F.v.reserve(total);
    for (int i = 0; i < n; ++i) {
        for (const auto& e : L[i]) { F.ci.push_back(e.first); F.v.push_back(e.second); }
        for (const auto& e : U[i]) { F.ci.push_back(e.first); F.v.push_back(e.second); }
        F.rp[i + 1] = (int)F.ci.size();
    }
    return F;
}

std::vector<int> tri_levels(const HostCsr& F, bool lower, int* n_levels) {
    const int n = F.rows;
    std::vector<int> lev(n, 0);
    int nl = 0;
    if (lower) {
        for (int i = 0; i < n; ++i) {
            int l = 0;
            for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {
                const int j = F.ci[k];
                if (j < i) l = std::max(l, lev[j] + 1);
            }
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int l = 0;
            for (int k = F.rp[i]; k < F.rp[i + 1]; ++k) {
                const int j = F.ci[k];
                if (j > i) l = std::max(l, lev[j] + 1);
            }
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    }
    if (n_levels) *n_levels = n ? nl : 0;
    return lev;
}

}  // namespace hsps_host
