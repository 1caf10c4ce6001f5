// =============================================================================
// dense_lu.hpp — the coarsest AMG level's dense solve operator on the device
// (dense_lu.hip).
// =============================================================================
#pragma once

#include "hsps_internal.hpp"

namespace hsps_dev {

// Rows of the padded dense array for an n x n operator (a multiple of 64).
int dense_pad(int n);

// G (row-major, *ld = dense_pad(A.rows) doubles per row; its top-left
// A.rows x A.rows block is the operator) such that G b is the reference's
// direct_solve(A, b) (amg.cpp:260-291), up to rounding: LU with partial
// pivoting on the device, FP64 MFMA for the blocked updates.  Synchronises s.
void dense_inverse_device(Ctx* ctx, const Csr& A, DevBuf& G, int* ld, hipStream_t s);

}  // namespace hsps_dev
