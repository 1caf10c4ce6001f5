// =============================================================================
// precond.hpp — device preconditioner objects (Jacobi, ILU(0)/ILUT, SA-AMG).
// =============================================================================
#pragma once

#include <memory>
#include <vector>

#include "hsps_internal.hpp"
#include "tri_band.hpp"

namespace hsps_dev {

// Level-scheduled triangular factor: rows grouped into dependency levels
// (host-computed), stored as a row permutation + level offsets.
struct TriSchedule {
    int n_levels = 0;
    std::vector<int> h_level_ptr;  // n_levels+1 offsets into order
    std::vector<int> h_order;
    DevBuf order;                  // rows sorted by level
    DevBuf level_ptr;
    int max_level_width = 0;
};

// The off-diagonal entries a sweep uses, in level order, K per row (ELL,
// padded): the sweep then reads a row's structure with independent loads, and
// a dependency on the previous level (kept in LDS by the sweep) as an LDS
// index.  K = 0 when some row has more than 8 such entries.
struct TriEll {
    int K = 0;
    int lds = 0;      // every level fits the LDS window
    int all_lds = 0;  // every dependency is on the previous level (read from LDS)
    DevBuf dep;       // per position and slot: row j (>= 0), -(1 + index in the previous level), or INT_MAX
    DevBuf val64, val32;
    DevBuf invd64, invd32;  // backward sweep: 1 / U(i,i) per position
};

struct IluFactor {
    std::unique_ptr<Csr> lu;  // L\U in one CSR (unit L implicit)
    DevBuf diag_pos;          // position of U(i,i) in lu (or -1)
    DevBuf inv_diag64, inv_diag32;
    TriSchedule fwd, bwd;
    TriEll fwd_ell, bwd_ell;
    DevBuf tmp64, tmp32;      // y of L y = r
    DevBuf pos64, pos32;      // a sweep's input in level order
    DevBuf opos64, opos32;    // a narrow sweep's output in level order
    TriBand fwd_band, bwd_band;  // wavefront sweeps, when the factor is banded (else ok = 0)
    int l_empty = 0;             // no entry below the diagonal: the forward sweep is y = r
    DevBuf band_err;             // int: set if a band sweep's wait gave up
};

struct AmgLevelDev {
    std::unique_ptr<Csr> A, P, R;
    DevBuf dinv64, dinv32;
    std::vector<int> agg;  // host aggregate map (levels < L-1)
    // per-level work vectors (FP64 and FP32 views)
    DevBuf x64, x64b, rhs64, res64;
    DevBuf x32, x32b, rhs32, res32;
};

// Wall time of the AMG setup stages, seconds.
struct AmgSetupTimes {
    double strength = 0, aggregate = 0, prolongation = 0, transpose = 0, galerkin = 0, coarse = 0;
};

struct Precond {
    int kind = HSPS_PRECOND_NONE;
    Ctx* ctx = nullptr;
    int n = 0;
    hsps_csr_s* A0 = nullptr;  // the operator it was built for (retained until destroy)
    unsigned long long A0_version = 0;  // A0->values_version at setup
    // Jacobi
    DevBuf dinv64, dinv32;
    double jacobi_omega = 1.0;
    // ILU
    std::unique_ptr<IluFactor> ilu;
    // AMG
    std::vector<AmgLevelDev> amg;
    double amg_omega = 0.67;
    int pre = 2, post = 2;
    int coarse_n = 0;
    int coarse_ld = 0;   // row stride of coarse_inv (the size padded to 64)
    DevBuf coarse_inv;   // FP64 solve operator of the coarsest level (row-major; both precisions)
    AmgSetupTimes amg_times;
    // caller-written (HSPS_PRECOND_USER): the callback and FP64 staging
    // vectors for the FP32 loops of mixed precision
    hsps_user_apply_fn user_fn = nullptr;
    void* user_data = nullptr;
    DevBuf user_r64, user_z64;
};

// Builds M.amg (levels, P, R, D^-1) for A on the device (amg_setup.hip).
void amg_setup_device(Precond& M, const Csr& A, const hsps_params& p, AmgSetupTimes* times);

}  // namespace hsps_dev

struct hsps_precond_s : hsps_dev::Precond {};
