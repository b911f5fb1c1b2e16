// prof.hpp -- optional HIP-event timing of the solver's batched launches.
//
// When enabled (pisces_prof_enable), every batched launch family of the
// block-Davidson engine is bracketed by a pair of HIP events recorded on the
// stream it runs on, together with the algorithmic HBM bytes of that launch
// (each operand read once, each result written once).  Totals are resolved on
// read.  Off by default: no events are recorded then.
#pragma once

#include "common.hpp"

namespace pisces {

enum ProfKind {
    PK_HX64 = 0,  // H*X, FP64 form (both plane passes)
    PK_HXLO,      // H*X, FP32 / split-BF16 form (both plane passes)
    PK_GRAM,      // block Gram partial sums + reduction
    PK_COMB,      // block combinations Y = A W
    PK_EIG,       // projected eigensolve
    PK_RESID,     // Ritz vectors, residuals and their norms (one fused pass)
    PK_PREC,      // preconditioner (diagonal or kinetic fast diagonalisation)
    PK_CHOL,      // Cholesky / orthogonalisation factors (small)
    PK_POT,       // electron-water potential on the grid (+ diag H)
    PK_SCF,       // induced-dipole solve of the water cluster
    PK_POL3,      // self-consistent polarization on the grid (PolType 3)
    PK_COUNT
};

bool prof_enabled();

// Records a start event at construction and a stop event at destruction.
class ProfScope {
public:
    ProfScope(int kind, hipStream_t s, double bytes);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

private:
    int slot_ = -1;
    hipStream_t s_ = nullptr;
};

}  // namespace pisces
