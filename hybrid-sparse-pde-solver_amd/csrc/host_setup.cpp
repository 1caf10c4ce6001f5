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
// row's active columns: the lower ones in a small sorted list, walked in
// increasing order while an elimination of k inserts the columns it
// activates (always > k, so behind the walk's position), and the upper ones
// collected and sorted at the end.  A bitmap answers "active?" in O(1), and
// only the columns the row touched are reset.  Same factors, bit for bit;
// O(nnz(L\U) * row length) per row instead of O(n).
HostCsr ilut(const HostCsr& A, double tau_rel, int fill_p) {
    const int n = A.rows;
    // finished rows, flat: L strictly lower, U from the diagonal on
    std::vector<long long> lp(n + 1, 0), up(n + 1, 0);
    std::vector<int> lc, uc;
    std::vector<double> lv, uv;
    lc.reserve(A.ci.size());
    lv.reserve(A.ci.size());
    uc.reserve(A.ci.size());
    uv.reserve(A.ci.size());
    std::vector<double> w(n, 0.0), udiag(n, 0.0);
    std::vector<unsigned long long> act((size_t)n / 64 + 1, 0);  // active columns of the work row
    std::vector<char> seen(n, 0);
    std::vector<int> touched, lower, upper;
    std::vector<std::pair<int, double>> Lr, Ur;
    auto on = [&](int j) { return (act[(size_t)j >> 6] >> (j & 63)) & 1ull; };
    auto set = [&](int j) { act[(size_t)j >> 6] |= 1ull << (j & 63); };
    auto clr = [&](int j) { act[(size_t)j >> 6] &= ~(1ull << (j & 63)); };
    auto touch = [&](int j) {
        if (!seen[j]) { seen[j] = 1; touched.push_back(j); }
    };
    for (int i = 0; i < n; ++i) {
        touched.clear();
        lower.clear();
        upper.clear();
        Lr.clear();
        Ur.clear();
        double rn = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            const int j = A.ci[p];
            w[j] = A.v[p];
            if (!on(j)) {
                set(j);
                if (j < i) lower.push_back(j);
                else if (j > i) upper.push_back(j);
            }
            touch(j);
            rn += A.v[p] * A.v[p];
        }
        std::sort(lower.begin(), lower.end());
        const double tau = tau_rel * std::sqrt(rn);
        // pivot columns k < i in increasing order (ilu.cpp:121-146)
        for (size_t t = 0; t < lower.size(); ++t) {
            const int k = lower[t];
            if (std::abs(w[k]) < tau) { w[k] = 0.0; clr(k); continue; }
            const double ukk = udiag[k];
            if (std::abs(ukk) < kEps) { clr(k); continue; }
            const double lik = w[k] / ukk;
            w[k] = lik;
            for (long long q = up[k]; q < up[k + 1]; ++q) {
                const int j = uc[q];
                if (j <= k) continue;
                if (!on(j) && std::abs(lik * uv[q]) >= tau) {
                    set(j);
                    if (j < i) lower.insert(std::upper_bound(lower.begin() + (long)t + 1, lower.end(), j), j);
                    else if (j > i) upper.push_back(j);
                }
                w[j] -= lik * uv[q];
                touch(j);
            }
            if (std::abs(w[k]) >= tau) Lr.emplace_back(k, w[k]);  // ilu.cpp:151-155, ascending k
        }
        double d = w[i];
        if (std::abs(d) < kEps) d = (d >= 0 ? 1.0 : -1.0) * kEps * 1e6;
        Ur.emplace_back(i, d);
        std::sort(upper.begin(), upper.end());
        for (int j : upper)
            if (std::abs(w[j]) >= tau) Ur.emplace_back(j, w[j]);
        for (int j : touched) {
            w[j] = 0.0;
            clr(j);
            seen[j] = 0;
        }
        w[i] = 0.0;
        clr(i);
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
        for (const auto& e : Lr) { lc.push_back(e.first); lv.push_back(e.second); }
        for (const auto& e : Ur) { uc.push_back(e.first); uv.push_back(e.second); }
        lp[i + 1] = (long long)lc.size();
        up[i + 1] = (long long)uc.size();
    }
    HostCsr F;
    F.rows = F.cols = n;
    F.rp.assign(n + 1, 0);
    F.ci.reserve(lc.size() + uc.size());
    F.v.reserve(lc.size() + uc.size());
    for (int i = 0; i < n; ++i) {
        F.ci.insert(F.ci.end(), lc.begin() + lp[i], lc.begin() + lp[i + 1]);
        F.v.insert(F.v.end(), lv.begin() + lp[i], lv.begin() + lp[i + 1]);
        F.ci.insert(F.ci.end(), uc.begin() + up[i], uc.begin() + up[i + 1]);
        F.v.insert(F.v.end(), uv.begin() + up[i], uv.begin() + up[i + 1]);
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
