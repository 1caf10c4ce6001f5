// =============================================================================
// host_setup.hpp — host-side setup stages of the solve path (compiled by the
// host compiler with OpenMP): SA-AMG greedy aggregation, ILU(0)/ILUT
// factorisation, triangular level schedules and the coarsest-level dense
// operator.  Decisions (aggregates, ILU patterns, pivots) follow
// the reference exactly; see each function's citation.
// =============================================================================
#pragma once

#include <vector>

namespace hsps_host {

struct HostCsr {
    int rows = 0, cols = 0;
    std::vector<int> rp{0}, ci;
    std::vector<double> v;
    int nnz() const { return (int)v.size(); }
};

// Greedy aggregation (amg.cpp:48-60) over a precomputed strength pattern
// (srp/sci: the strong neighbours j != i of each row, |a_ij| >= theta *
// max_{l!=i}|a_il|, in row order): each unassigned row in order seeds an
// aggregate that its unassigned strong neighbours join.
std::vector<int> aggregate_strong(int n, const std::vector<int>& srp, const std::vector<int>& sci, int* n_coarse);

// ILU(0) on A's pattern (ilu.cpp:41-89) / ILUT(tau, p) (ilu.cpp:94-220).
// Throws std::runtime_error on a near-zero pivot with the reference's message.
HostCsr ilu0(const HostCsr& A);
HostCsr ilut(const HostCsr& A, double tau, int fill_p);

// Dependency levels of the unit-lower (lower=true) or upper part of an L\U
// CSR.  Returns level of each row; *n_levels receives the count.
std::vector<int> tri_levels(const HostCsr& LU, bool lower, int* n_levels);

}  // namespace hsps_host
