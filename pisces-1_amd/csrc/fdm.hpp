// fdm.hpp -- kinetic (fast-diagonalisation) preconditioner.
#pragma once

#include "common.hpp"

namespace pisces {

// dst = scale * (Tx (+) Ty (+) Tz - shift)^{-1} src for one grid vector, with
// |lam_x + lam_y + lam_z - shift| floored at `floor`, applied through the
// eigenbases of the 1-D kinetic matrices (ax[a].q64 / qt64 / split BF16 and
// ax[a].lam).  absden: divide by |lam_x + lam_y + lam_z - shift| instead (the
// absolute value of T - shift: positive definite wherever shift lies, the
// same operator when shift is below the kinetic spectrum).  work: ngp BF16
// values of scratch (split-BF16 form only).  The FP64 form works in place in
// dst.
struct FdmTask {
    AxisMats ax[3];
    const void *src;  // FP64, or FP32 if src32 (split-BF16 form only)
    void *dst;        // FP64, or FP32 if dst32 (split-BF16 form only)
    uint16_t *work;   // BF16 intermediate between the passes (split-BF16 form only)
    double shift, scale, floor;
    int src32, dst32, absden;
};

// precision: PISCES_FP64 (FP64 matrix cores) or PISCES_BF16X2 (split BF16)
void fdm_launch(const FdmTask *d_tasks, int ntasks, int max_n, int precision, hipStream_t s);

// The split-BF16 form with its three passes in one XCD-ordered launch (each
// pass reads the intermediate the previous one left in the L2).  The plan
// (fdm_fused_plan_ints ints, built on the host from the host copy of the
// tasks) must reach the device unchanged with every launch: it carries zeroed
// counters.
int fdm_fused_plan_ints(int ntasks);
void fdm_fused_plan(const FdmTask *h_tasks, int ntasks, int *plan);
void fdm_fused_launch(const FdmTask *d_tasks, int *d_plan, const int *h_plan, int max_n,
                      hipStream_t s);
// default on (PISCES_FDM_FUSED=0: three launches)
bool fdm_fused_enabled();

}  // namespace pisces
