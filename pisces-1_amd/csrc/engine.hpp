// engine.hpp -- the batched block-Davidson engine.
//
// Many independent eigenproblems (grids of different sizes, different
// potentials) advance together: each Davidson step is one fixed sequence of
// batched launches over every live instance, with a single device-to-host
// transfer of the per-root Ritz values and residual norms.  The same engine
// runs one instance (pisces_solve_block) or many (the server).
//
// Algorithm (per instance; k = block size, nroots <= k wanted):
//   basis  = [X | C]: X the k current Ritz vectors (FP64, orthonormal) with
//            HX = H X kept in FP64; C the corrections of this macro iteration,
//            stored in the solve precision with HC = H C from the matching
//            H*X form (FP64, FP32 MFMA or split-BF16 MFMA).
//   step   : G = [X|C]^T H [X|C] extended by its new columns (FP64 sums);
//            lowest k eigenpairs of G; U = [X|C] y, HU = [HX|HC] y (FP64);
//            R = HU - theta U and ||R|| (FP64); Davidson corrections
//            R_i / (theta_i - diag H) for the unconverged columns,
//            orthogonalised against [X|C] (two CGS passes) and among
//            themselves (Cholesky QR, twice), appended to C.
//   restart: when C is full: X = U (re-orthonormalised), HX = HU; the
//            low-precision forms then recompute HX = H X in FP64 (defect
//            correction), so residuals and eigenvalues are FP64-exact.  The
//            last search directions P = C Y_C, orthonormalised against the
//            Ritz coefficients in the coefficient space of the old basis, open
//            the new C (GD+k); H P comes from [HX|HC] by the same coefficients.
//   converged: ||R_i|| < 10^-ptol for i < nroots, judged on an FP64-fresh HX
//            (every form: a restart that recomputes H X in FP64 confirms it).
#pragma once

#include <memory>
#include <vector>

#include "bd.hpp"
#include "common.hpp"

namespace pisces {

// Grid data of one instance on the device (non-owning).
struct BdGrid {
    int n[3] = {0, 0, 0};
    long long N = 0;
    AxisMats ax[3];             // kinetic operators and eigenbases of the axes
    const double *V64 = nullptr;
    const float *V32 = nullptr;
    const double *D = nullptr;  // diag(H)
};

struct BdResult {
    int status = 0;   // >= 0: converged roots; < 0 error
    int nroots = 0;
    std::vector<double> evals, resid;
    int n_iter = 0, n_mv64 = 0, n_mvlo = 0, n_restart = 0;
};

class BdEngine {
public:
    // precond: 1 Davidson diagonal correction, 3 kinetic (fast diagonalisation)
    BdEngine(int device, hipStream_t s, int precision, int block, int maxsub, int ptol,
             int max_iter, int precond);
    ~BdEngine();
    // Adds an instance; start: nroots host vectors (more columns are filled
    // with seeded random vectors up to the block size).  Returns its slot.
    int add(const BdGrid &g, int nroots, const double *start, int nstart, unsigned seed,
            std::shared_ptr<void> keepalive = nullptr, bool start_on_device = false);
    // Advances every live instance by one step; returns the number still live.
    int step();
    // The two halves of step(): step_a() queues the Ritz part of the step
    // (projected eigensolve, Ritz vectors, residual norms and their transfer to
    // the host) and returns without waiting; step_b() waits for the transfer,
    // decides per instance (converged, restart, corrections) and queues the
    // corrections.  Engines on different streams overlap their device work
    // and their host decisions when every step_a() is called before the
    // step_b()s.  step_b() must follow step_a() before the engine is used again.
    void step_a();
    int step_b();
    hipStream_t stream() const { return s_; }
    bool finished(int slot) const;
    const BdResult &result(int slot) const;
    // steps of the slot whose second orthogonalisation round ran on the stored
    // data (the FP32-stored forms; waits for the stream)
    int data_rounds(int slot);
    // Ritz vectors of a finished instance (nroots x N doubles, host).
    void vectors(int slot, double *out);
    // Device pointer to the Ritz vectors of a finished instance (valid until release).
    const double *vectors_device(int slot) const;
    // Frees pooled buffers (of released instances) beyond max_bytes in total.
    void trim_pool(size_t max_bytes);
    size_t pool_bytes() const;
    void release(int slot);
    size_t bytes_per_instance(long long N) const;
    int live() const;
    int precision() const { return prec_; }

private:
    struct Inst;
    struct Arena;
    std::vector<std::unique_ptr<Inst>> inst_;
    int dev_;
    hipStream_t s_;
    int prec_, k_, maxsub_, ptol_, max_iter_, precond_;
    std::unique_ptr<Arena> arena_[2];  // task descriptors of phases A and B
    dbuf dpart_;
    dbuf dgather_;
    double *hsmall_ = nullptr;
    size_t hsmall_cap_ = 0;
    bool trace_ = false;  // PISCES_TRACE=1: per-step Ritz values of the first instance
    double refresh_below_ = 1e-3;  // late restarts refresh H X in FP64 below this residual
    // second orthogonalisation round on the data where the first round's
    // squared amplification exceeds amp2_max_ (default 1e3; PISCES_AMP2_MAX)
    bool data_round2_ = true;
    double amp2_max_ = 1e3;
    void grow_part(size_t bytes);
    std::vector<dbuf> pool_;  // released instance buffers, reused by add()
    std::vector<Inst *> act_, rs_, lt_;  // the step between step_a() and step_b()
    bool mid_ = false;                   // step_a() queued, step_b() not yet run
    dbuf dinit_;  // keep mask (ints, all ones) and, at byte kInitEye, the k x k identity
    static constexpr size_t kInitEye = 512;
    dbuf take_buf(size_t bytes);
};

}  // namespace pisces
