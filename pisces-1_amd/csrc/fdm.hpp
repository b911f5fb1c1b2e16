// fdm.hpp -- kinetic (fast-diagonalisation) preconditioner.
#pragma once

#include "common.hpp"

namespace pisces {

// dst = scale * (Tx (+) Ty (+) Tz - shift)^{-1} src for one grid vector, with
// |eigenvalue - shift| floored at `floor`.  Q[a]: eigenvectors of T_a (column
// major, n_a x n_a), lam[a]: its eigenvalues.
struct FdmTask {
    int n[3];
    int pad;
    const double *Q[3];
    const double *lam[3];
    const double *src;
    double *dst;
    double shift, scale, floor;
};

void fdm_launch(const FdmTask *d_tasks, int ntasks, int max_n, hipStream_t s);

}  // namespace pisces
