// =============================================================================
// host_setup.cpp — host-side setup stages (see host_setup.hpp).
// =============================================================================
#include "host_setup.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
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
namespace {
// Factor rows of one thread: chunks that never move, so other threads can
// read a finished row through its pointers while this one appends.
struct RowStore {
    std::vector<std::unique_ptr<int[]>> ci;
    std::vector<std::unique_ptr<double[]>> v;
    size_t cap = 0, used = 0;
    std::pair<int*, double*> take(size_t m) {
        if (ci.empty() || used + m > cap) {
            cap = std::max<size_t>(m, size_t(1) << 20);
            ci.emplace_back(new int[cap]);
            v.emplace_back(new double[cap]);
            used = 0;
        }
        std::pair<int*, double*> r(ci.back().get() + used, v.back().get() + used);
        used += m;
        return r;
    }
};

constexpr int kIlutBlock = 512;        // rows per block at least (and the factor's lower bandwidth)
constexpr int kIlutPublish = 16;       // a block's progress is published every this many rows
struct alignas(64) Progress {          // one cache line per block's count (no false sharing)
    std::atomic<int> rows{0};
};
constexpr int kIlutMinRows = 65536;    // below this one thread
constexpr int kIlutMaxThreads = 8;

int ilut_threads(int n, int nb) {
    int t = 1;
    if (const char* e = std::getenv("HSPS_ILUT_THREADS")) {
        t = std::max(1, std::atoi(e));
    } else if (n >= kIlutMinRows) {
        t = std::min(kIlutMaxThreads, std::max(1, omp_get_max_threads()));
    }
    return std::max(1, std::min(t, nb));
}
}  // namespace

// Rows run in blocks of S >= the lower bandwidth bw of A, dealt round-robin to
// threads.  Row i reads the finished U rows k of the columns it eliminates,
// all in [i - bw, i) (a fill column lies between a pivot row and i), so in
// its own block (done by its own thread, in order) or the previous block:
// for those it waits until that block's thread has published row k (a
// per-block count of finished rows, release / acquire).  The lowest
// unfinished block never waits, so every wait ends.  Each row's arithmetic
// is the sequential scan's, so the factors are the same bits for any number
// of threads (HSPS_ILUT_THREADS sets it; by default up to 8 from 65,536 rows).
HostCsr ilut(const HostCsr& A, double tau_rel, int fill_p) {
    const int n = A.rows;
    int bw = 0;
    for (int i = 0; i < n; ++i)
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p)
            if (A.ci[p] < i) bw = std::max(bw, i - A.ci[p]);
    const int S = std::max(bw, kIlutBlock);
    const int nb = (n + S - 1) / S;
    const int T = ilut_threads(n, nb);
    // finished rows: L strictly lower, U from the diagonal on
    std::vector<const int*> lcp(n), ucp(n);
    std::vector<const double*> lvp(n), uvp(n);
    std::vector<int> llen(n, 0), ulen(n, 0);
    std::vector<double> udiag(n, 0.0);
    std::unique_ptr<Progress[]> done(new Progress[std::max(nb, 1)]);
    std::vector<RowStore> stores(T);
    // a thread that throws (allocation) stops the others' waits; the first
    // exception is rethrown once every thread has returned
    std::atomic<bool> failed{false};
    std::vector<std::exception_ptr> errs(T);
    auto worker = [&](int t) {
        RowStore& st = stores[t];
        std::vector<double> w(n, 0.0);
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
        for (int b = t; b < nb; b += T) {
            const int r0 = b * S, r1 = std::min(n, r0 + S);
            int prev = 0;  // rows of block b - 1 seen published
            for (int i = r0; i < r1; ++i) {
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
                for (size_t q0 = 0; q0 < lower.size(); ++q0) {
                    const int k = lower[q0];
                    if (std::abs(w[k]) < tau) { w[k] = 0.0; clr(k); continue; }
                    if (k < r0 && k - (b - 1) * S >= prev) {  // row k of the previous block: wait until published
                        const int need = k - (b - 1) * S + 1;
                        for (int spin = 0; (prev = done[b - 1].rows.load(std::memory_order_acquire)) < need; ++spin) {
                            if (failed.load(std::memory_order_relaxed)) return;
                            if (spin < 1024) __builtin_ia32_pause();
                            else std::this_thread::yield();
                        }
                    }
                    const double ukk = udiag[k];
                    if (std::abs(ukk) < kEps) { clr(k); continue; }
                    const double lik = w[k] / ukk;
                    w[k] = lik;
                    const int* uck = ucp[k];
                    const double* uvk = uvp[k];
                    for (int q = 0; q < ulen[k]; ++q) {
                        const int j = uck[q];
                        if (j <= k) continue;
                        if (!on(j) && std::abs(lik * uvk[q]) >= tau) {
                            set(j);
                            if (j < i) lower.insert(std::upper_bound(lower.begin() + (long)q0 + 1, lower.end(), j), j);
                            else if (j > i) upper.push_back(j);
                        }
                        w[j] -= lik * uvk[q];
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
                double di = 0.0;
                for (const auto& e : Ur)
                    if (e.first == i) { di = e.second; break; }
                udiag[i] = di;
                auto put = [&](const std::vector<std::pair<int, double>>& row, const int** cp, const double** vp, int* len) {
                    auto m = st.take(row.size());
                    for (size_t e = 0; e < row.size(); ++e) {
                        m.first[e] = row[e].first;
                        m.second[e] = row[e].second;
                    }
                    *cp = m.first;
                    *vp = m.second;
                    *len = (int)row.size();
                };
                put(Lr, &lcp[i], &lvp[i], &llen[i]);
                put(Ur, &ucp[i], &uvp[i], &ulen[i]);
                if ((i - r0 + 1) % kIlutPublish == 0 || i + 1 == r1)
                    done[b].rows.store(i - r0 + 1, std::memory_order_release);
            }
        }
    };
    auto guarded_worker = [&](int t) {
        try {
            worker(t);
        } catch (...) {
            errs[t] = std::current_exception();
            failed.store(true);
        }
    };
    if (T == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        th.reserve(T - 1);
        for (int t = 1; t < T; ++t) th.emplace_back(guarded_worker, t);
        guarded_worker(0);
        for (auto& x : th) x.join();
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
    }
    HostCsr F;
    F.rows = F.cols = n;
    F.rp.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) F.rp[i + 1] = F.rp[i] + llen[i] + ulen[i];
    F.ci.resize(F.rp[n]);
    F.v.resize(F.rp[n]);
#pragma omp parallel for schedule(static) if (n >= kIlutMinRows)
    for (int i = 0; i < n; ++i) {
        const int o = F.rp[i];
        std::copy(lcp[i], lcp[i] + llen[i], F.ci.begin() + o);
        std::copy(lvp[i], lvp[i] + llen[i], F.v.begin() + o);
        std::copy(ucp[i], ucp[i] + ulen[i], F.ci.begin() + o + llen[i]);
        std::copy(uvp[i], uvp[i] + ulen[i], F.v.begin() + o + llen[i]);
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
