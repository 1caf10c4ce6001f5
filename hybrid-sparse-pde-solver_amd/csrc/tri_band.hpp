#pragma once
// =============================================================================
// tri_band.hpp — wavefront triangular sweeps for banded factors (ILU of
// grid-ordered matrices), the alternative to level scheduling.
//
// A sweep over rows i' (forward: i' = i; backward: i' = n-1-i, so both are
// lower-triangular in i') whose dependencies all lie in [i'-W, i'-1] is laid
// out as a grid of rows of W columns: i' = g W + j.  A dependency then sits
// in the same grid row (column j-a, "own", a >= 1) or in the previous one
// (column j+b, b >= 0).  Lane l of a wave owns grid row g = 64 band + l and
// solves column j at its step tau = j + S l (S = 1 + max b), so every
// dependency was computed by itself a steps ago or by lane l-1 (S-b) steps
// ago: both kept in short register histories, the neighbour's passed by a
// lane shift each step.  Waves take bands of 64 grid rows in turn; lane 0 of
// a band reads the previous band's last grid row from the sweep's output
// from an LDS ring that band's wave fills as it goes (one workgroup; waits
// only on earlier bands, so the earliest unfinished band always proceeds).
// Critical path about W + S x (grid rows) steps of a few instructions,
// against 2 W levels separated by workgroup barriers for level scheduling.
// Every row is the same expression in the same order as the level sweeps
// (ascending columns, then the 1/U(i,i) scale), so the results are the
// same bits.
//
// Vectors and factor entries are stored "skewed": element (band, tau, lane)
// at ((band T + tau) 64 + lane), T = W + 63 S rounded up to the handoff
// chunk, so a wave's loads and stores at one step are one contiguous run.
// =============================================================================
#include <vector>

#include "host_setup.hpp"
#include "hsps_internal.hpp"

namespace hsps_dev {

struct TriBand {
    int ok = 0;       // the sweep qualifies (banded, short histories, no more than 8 entries a row)
    int K = 0;        // entry slots per row: 2 (then H = 1), 4 or 8 (H = 4)
    int H = 0;        // history depth per source
    int W = 0, S = 1; // grid width and lane skew
    int NB = 0, T = 0;          // bands, steps per band (a multiple of the handoff chunk)
    int NW64 = 0, NW32 = 0;     // waves of the workgroup (FP64 / FP32 sweeps: LDS rings of W values each)
    long long n_sk = 0;         // skewed positions (NB * T * 64)
    DevBuf val64, val32;        // [n_sk * K] entry values in summation order
    DevBuf code;                // [n_sk * K] history slot of each entry; kBandNone for an empty slot
    DevBuf invd64, invd32;      // backward sweep: 1 / U(i,i) per position
    DevBuf sk_of_row;           // [n] skewed position of natural row i
    DevBuf in64, in32;          // [n_sk] sweep input, skewed; padding positions stay zero
    DevBuf out64, out32;        // [n_sk] sweep output, skewed
    int mode = 0;               // 1: one workgroup (LDS handoff); 2: a workgroup per 4 waves (global handoff)
    int five = 0;               // two fixed-source slots (five-point ILU(0)): values in slot order, no codes
    int NWT = 0;                // waves of the multi-workgroup sweep
    long long n_slot_words = 0;
    DevBuf slots;               // multi-workgroup handoff slots: [NWT][2 rounds][W][2] tagged 64-bit words
    DevBuf con;                 // [NWT] consumer progress: pair-stream round x W + columns read
};

// The band layout of a sweep of F (lower: the strictly lower entries; upper:
// the strictly upper ones, reversed), or ok = 0 when it does not qualify.
TriBand make_tri_band(Ctx* c, const hsps_host::HostCsr& F, bool lower, const std::vector<double>& invd);

// One band sweep of in (natural row order) into out (natural row order):
// gather into the skewed input, the wavefront kernel, scatter back.  err
// (device int) is set if a wait ever gives up (never, by construction).
template <class T>
void tri_band_sweep(const TriBand& B, bool upper, int n, const T* in, T* out, int* err, hipStream_t s,
                    const int* skip);

}  // namespace hsps_dev
