/*
 * pisces_hip.h -- C ABI of the MI355X (gfx950) implementation of the PISCES
 * excess-electron eigensolve: sine-DVR H*X, the DPP electron-water potential on
 * the grid, and the Davidson eigensolver with its block operations.
 *
 * Conventions shared by every entry point:
 *   - Plain C types only; the caller owns host buffers, a context owns its
 *     device buffers.  No global state: contexts are independent and may be
 *     used from different threads concurrently (one thread per context).
 *   - Grid vectors are x-fastest, index i + nx*(j + ny*k), exactly the layout of
 *     DVR::wavefn / v_diag in the reference (src/mtx.cpp:15-19).  A block of
 *     nvec vectors is nvec such columns back to back (the reference's B and Z,
 *     src/davdriver.cpp:34-39).
 *   - Lengths are bohr, energies hartree.
 *   - Return codes: 0 ok; < 0 error (see pisces_strerror); > 0 "not converged"
 *     where documented.  No entry point calls exit().
 *   - Pointers named d_* are device pointers (hipMalloc'd or torch CUDA
 *     tensors); all others are host pointers.
 */
#ifndef PISCES_HIP_H
#define PISCES_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pisces_ctx pisces_ctx;

/* precision of the H*X products inside a solve */
enum {
    PISCES_FP64 = 0,   /* plain form: FP64 everywhere                         */
    PISCES_FP32 = 1,   /* FP32 MFMA contractions, FP64 block ops and residuals */
    PISCES_BF16X2 = 2  /* split hi/lo BF16 MFMA contractions, FP32 accumulate  */
};

/* correction (preconditioner) choices of pisces_solve_block */
enum { PISCES_PRECOND_DIAG = 1, PISCES_PRECOND_KINETIC = 3 };

enum {
    PISCES_OK = 0,
    PISCES_ERR_ARG = -1,
    PISCES_ERR_HIP = -2,
    PISCES_ERR_NOMEM = -3,
    PISCES_ERR_UNSUPPORTED = -4,
    PISCES_ERR_STATE = -5,
    PISCES_ERR_NUMERIC = -6
};

/* Version / build string of the library. */
const char *pisces_version(void);
const char *pisces_strerror(int code);
/* Message of the last failing call on this thread (empty if none). */
const char *pisces_last_error(void);
/* Selects the H*X kernels of one context: 1 matrix cores (default: the passes
 * in separate launches unless PISCES_HX_FUSED=1 / PISCES_FDM_FUSED=1), 2 matrix
 * cores in separate launches always, 3 matrix cores with the split-BF16 H*X
 * and kinetic preconditioner each in one XCD-ordered launch (slower than 1 on
 * the served batch; PISCES_ERR_UNSUPPORTED on a device without eight XCDs),
 * 0 LDS-tiled VALU (FP64 and FP32; the split-BF16 form as for 2).  For
 * cross-checks; no reference counterpart. */
int pisces_debug_set_hx_variant(pisces_ctx *ctx, int variant);
/* Number of times a wait inside the one-launch split-BF16 H*X reached its spin
 * limit on the current device since the library was loaded (0 unless the
 * ordering it relies on broke; the launch then goes on).  No reference
 * counterpart. */
unsigned pisces_debug_hx_spin_timeouts(void);
/* Inverse of a symmetric n x n matrix (row-major) as the PolType 3 set-up
 * inverts the polarization matrix (ComputeInvA, src/MolPolAux.cpp:39-136, with
 * dsytrf/dsytri there): method 0 Cholesky where positive definite, else
 * Gauss-Jordan with partial pivoting (what pisces_build_potential_scp uses);
 * 1 Cholesky only; 2 Gauss-Jordan only.  Host arithmetic, no device needed.
 * Errors: PISCES_ERR_ARG, PISCES_ERR_NUMERIC (singular; for method 1 also not
 * positive definite). */
int pisces_debug_sym_inverse(int n, const double *a, int method, double *inv);

/* Launch timing: when enabled, the solver brackets each of its batched launch
 * families with HIP events on the stream they run on and accumulates time,
 * launch count and algorithmic HBM bytes (operands read once, results written
 * once) per family.  kind: 0 H*X FP64, 1 H*X FP32/BF16X2, 2 Gram, 3 block
 * combination, 4 projected eigensolve, 5 Ritz vectors + residuals, 6
 * preconditioner, 7 small factors (Cholesky, orthogonalisation), 8 grid
 * potential, 9 induced-dipole solve.  Measurement hook
 * for the benchmark; no reference counterpart. */
int pisces_prof_enable(int on);
int pisces_prof_reset(void);
int pisces_prof_read(int kind, double *ms_total, long long *launches, double *bytes_total);
const char *pisces_prof_name(int kind);
/* Gram kernel configuration picked by the start-up autotuner of the block
 * solver (variant 0 VALU register tiles, 1 matrix cores with 16-row blocks,
 * 2 matrix cores with 64-row runs; rows per workgroup).  Valid after the first
 * block solve or server creation. */
int pisces_tuning_gram(int *variant, int *rows);

/* Create a context for one sine-DVR grid of n[0] x n[1] x n[2] points on boxes
 * of length L[3] (bohr), centred at the origin; computes the 1-D grids and
 * kinetic matrices (sine_dvr, src/sine_dvr.cpp:45-131) and uploads them.
 * Replaces DVR::SetupDVR for DVRType 0 (src/DVR.cpp:64-161, :524-651).
 * device: HIP device ordinal. */
int pisces_create(pisces_ctx **ctx, const int n[3], const double L[3], int device);
/* The same with the DVR type of DVR::SetupDVR (src/DVR.h:56-91, `type`):
 * 0 the sine DVR (as pisces_create); 1 the harmonic-oscillator DVR (ho_dvr,
 * src/ho_dvr.cpp), with L[a] the oscillator frequency Omega0 of axis a
 * (src/GetInput.cpp:133-140), mass 1; 20 the Colbert-Miller DVR (cm_ham,
 * src/cm_dvr.cpp) on [-L/2, L/2]; 3 the FFT kinetic energy of DVRType 3
 * (Tdiag, src/KE_diag.cpp; applied through FFTW by VectorFFT::apply,
 * src/VectorFFT.hpp:32, src/VectorFFTW.cpp), on the same grid points.  The
 * kinetic operator of type 3 is diagonal in the discrete Fourier basis of each
 * axis; here it is applied as the equivalent real circulant matrix per axis
 * through the same matrix-core axis contractions.  For non-cubic grids this
 * follows the per-axis definition (the reference's 3-D index of KE_diag is
 * only right for cubic grids, src/VectorFFTW.cpp:83).  diag(H) of type 3 takes
 * the circulant's diagonal, the mean of the axis' Fourier kinetic energies.
 * Type 3 needs even grid sizes: for odd n, Tdiag's wave numbers (a < n/2: a,
 * else a - n) are not symmetric about 0 and the multiplier is not real on real
 * vectors.  Other types and odd sizes: PISCES_ERR_UNSUPPORTED. */
int pisces_create_ex(pisces_ctx **ctx, const int n[3], const double L[3], int device,
                     int dvr_type);
void pisces_destroy(pisces_ctx *ctx);

/* Grid points (x[n0], y[n1], z[n2]) and kinetic matrices (n_a*n_a each,
 * symmetric) of the context; either pointer may be NULL. */
int pisces_grid(const pisces_ctx *ctx, double *xyz, double *T);

/* Set the diagonal potential from a host array of ngp doubles (x fastest).
 * Replaces DVR::ComputePotential's product (v_diag) when the caller builds the
 * potential itself. */
int pisces_set_potential(pisces_ctx *ctx, const double *v_diag);
int pisces_get_potential(const pisces_ctx *ctx, double *v_diag);

/* Build the electron-water potential on the grid on the GPU from the lists
 * that ClusterAnion hands to Potential::Setup (src/ClusterAnion.cpp:186-246;
 * Potential::Setup src/Potential.h:38-42) and evaluate it at every grid point
 * (DVR::ComputePotential sampling 1, src/DVR.cpp:661-859, with
 * Potential::Evaluate / EvaluateDPPGTOP, src/Potential.cpp:128-182, 844-956).
 *   flags[6]  Potential, DampType, SigmaOFlag, Polarization, AdiabaticPolPot,
 *             InternalParam (PotFlag of src/GetInput.cpp:43-48)
 *   potpara   PotPara[0..10] (src/GetInput.cpp:53-61)
 *   sites     3*nsites, 4 per water (O,H1,H2,M)
 *   q/iq      charges and their site indices; d/id dipoles and site indices
 *   ion_*     optional erf-damped point charges (Bloomfield form,
 *             src/NaCl.cpp:285-318); nion may be 0.
 * Supported: Potential 1-4, DampType 1-2, Polarization 0-1 (and adiabatic);
 * Polarization 3 goes through pisces_build_potential_scp. */
int pisces_build_potential(pisces_ctx *ctx, const int flags[6], const double *potpara,
                           int nsites, const double *sites, int nq, const double *q,
                           const int *iq, int nd, const double *d, const int *id,
                           int nion, const double *ion_xyz, const double *ion_q,
                           double ion_r);

/* The same with self-consistent polarizable sites (Polarization flag 3,
 * flags[3] == 3): Potential::Setup with its polarizable-site arguments
 * (src/Potential.h:38-42: np, alpha, ip, Epc; SetupDPPGTOP case 3,
 * src/Potential.cpp:387-432, InvA from ComputeInvA, src/MolPolAux.cpp:39-136)
 * and VpolWaterWater (src/Potential.h:85, set at src/ClusterAnion.cpp:173).
 * The water model's point dipoles are not taken: PolType 3 ignores them
 * (src/Potential.cpp:327-335).  Vpol at a grid point is
 * SelfConsistentPolarizability (src/Potential.cpp:1495-1592):
 * -E^T InvA E / 2 - vpol_ww, E the electron's damped field at the sites plus
 * epc.  Where a grid point coincides with a site, that site's electron field
 * is taken as its limit 0 (the reference: 0 for DampType 2, 0 * inf for 1).
 *   np sites ip[] (indices into sites), alpha[np], epc[3*np]; 3*np <= 576.
 * Errors: PISCES_ERR_ARG (flags[3] != 3, bad lists), PISCES_ERR_NUMERIC
 * (singular polarization matrix), PISCES_ERR_UNSUPPORTED (too many sites). */
int pisces_build_potential_scp(pisces_ctx *ctx, const int flags[6], const double *potpara,
                               int nsites, const double *sites, int nq, const double *q,
                               const int *iq, int np, const double *alpha, const int *ip,
                               const double *epc, double vpol_ww, int nion,
                               const double *ion_xyz, const double *ion_q, double ion_r);

/* Water model on the GPU: rigid DPP monomers from O,H1,H2 Cartesian input in
 * Angstrom (Water::Reposition + UpdatePositions, src/Water.cpp:695-943), the
 * FP64 self-consistent induced dipoles (WaterCluster::CalcInducedDipoles,
 * src/DPP.cpp:760-1065), and the lists above (GetLists :402-498,
 * ReportInducedDipoles :505-529), then the grid potential.
 * Outputs (optional, may be NULL): sites[12*nw], mu[9*nw].
 * flags[3] == 3 builds the self-consistent polarization of the O,H1,H2 sites
 * (ReportPolSitesAndField, src/DPP.cpp:540-642) with VpolWaterWater = the
 * water-water polarization energy (src/ClusterAnion.cpp:173).
 * Returns > 0 if the dipole SCF did not converge (the potential is still built,
 * as in the reference, which only warns). */
int pisces_build_cluster_potential(pisces_ctx *ctx, int nw, const double *xyz_ang,
                                   const int flags[6], const double *potpara, int nion,
                                   const double *ion_xyz_bohr, const double *ion_q,
                                   double ion_r, double *sites, double *mu);

/* Energy of the neutral water cluster, E0 of the reference's energy block
 * (WaterCluster::CalcEnergy -> CalcIntermolecularPotential, src/DPP.cpp:
 * 251-262, 1648-1700; printed as Eneutral, src/ClusterAnion.cpp:306-311):
 * the water-water electrostatic, polarization and DPP repulsion/dispersion
 * terms summed over pairs.  sites: 12*nw (bohr, O,H1,H2,M per water) and mu:
 * 9*nw (induced dipoles of O,H1,H2), as pisces_build_cluster_potential returns
 * them.  energies[4]: E0, electrostatic, polarization, dispersion (hartree).
 * Host computation; no GPU is used. */
int pisces_water_energy(int nw, const double *sites, const double *mu, double *energies);

/* Electron side of the analytic gradient (DVR::ComputeGradient's grid loop,
 * src/DVR.cpp:1099-1176, over Potential::EvaluateDPPGTOPGradient,
 * src/Gradients.cpp:79-238) for the context's last potential, which must be a
 * PolType 3 one with Potential 3 and DampType 2 (the flag set the reference's
 * gradient is written for) and no ions: for the wavefunction psi (host, ngp
 * doubles, x fastest; DVR::wavefn of the state),
 *   gel[3 nsites] = sum_g psi_g^2 dV(r_g)/dR_site (charges, Slater cores and
 *                   the electron's field derivative at the polarizable sites
 *                   dotted with mu = InvA (E + Epc)),
 *   tmu[3 np]     = sum_g psi_g^2 mu(r_g),
 *   mcm[(3 np)^2] = sum_g psi_g^2 mu(r_g) mu(r_g)^T (row-major),
 * with np and nsites those of the potential.  PISCES_ERR_STATE without such a
 * potential. */
int pisces_gradient_terms(pisces_ctx *ctx, const double *psi, double *gel, double *tmu,
                          double *mcm);

/* Replaces DVR::ComputeGradient (src/DVR.cpp:1099-1176) with its arguments
 * (src/ClusterAnion.cpp:407): gradient[3 nsites] holds the water-water
 * gradient from WaterCluster::CalcGradient on entry and the total on return:
 * += gel, += FinalGradient's terms (src/Gradients.cpp:240-268) from dEfield
 * (WaterCluster::Calc_dInducedDipoles, 3 nsites x 3 np) and dT_x/y/z
 * (BuildSuperdT, (3 np)^2 each), -= polgrad (SubtractWWGradient).  nsites: 4
 * per water, the potential's sites. */
int pisces_compute_gradient(pisces_ctx *ctx, const double *psi, int nsites, double *gradient,
                            const double *dT_x, const double *dT_y, const double *dT_z,
                            const double *polgrad, const double *dEfield);

/* Dual-grid method (ClusterAnion with Polarization 6, src/ClusterAnion.cpp:
 * 207-299), cubic grids.  pisces_dual_grid_coarse: the coarse grid of an n^3
 * grid on length L -- (n + 1) / 2 points at twice the spacing for odd n,
 * (n / 2 + 1) / 3 * 2 points at three times it for even n (:209-232).
 * pisces_dual_grid_extend: a coarse vector (nc^3 doubles, host) carried to the
 * fine grid of ctx (ngp doubles, host); kind 0 a wavefunction (ExtendWfDouble /
 * ExtendWfThree, src/DVR.cpp:1446-1722: the start vector of SVFlag 3), 1 a
 * potential (InterpolVdouble / InterpolVtriple, :1240-1421).
 * pisces_dual_grid_potential: the potential of Polarization 6
 * (DVR::ComputePotential, src/DVR.cpp:776-830): on ctx's grid, whose exact
 * potential must have been built, every point on the coarse lattice or farther
 * than rtol (PotPara[10], default 20 bohr) from every site of that build takes
 * the coarse potential carried over; the others keep their exact value. */
int pisces_dual_grid_coarse(const int n[3], const double L[3], int nc[3], double Lc[3]);
int pisces_dual_grid_extend(pisces_ctx *ctx, const int nc[3], const double *coarse, int kind,
                            double *fine);
int pisces_dual_grid_potential(pisces_ctx *ctx, const int nc[3], const double *coarse_v,
                               double rtol);

/* Y[:,v] = H X[:,v] for v < nvec, host buffers (ngp*nvec doubles each).
 * Replaces DVR::MatrixTimesVector (src/mtx.cpp:22-108) called once per vector
 * by davdriver (src/davdriver.cpp:67-78); the context's potential is used. */
int pisces_apply_h(pisces_ctx *ctx, const double *X, double *Y, int nvec);

/* Same on device buffers: d_X, d_Y hold nvec vectors with leading dimension
 * ngp, in the given precision (FP64: double; FP32 and BF16X2: float). */
int pisces_apply_h_device(pisces_ctx *ctx, const void *d_X, void *d_Y, int nvec,
                          int precision);

/* Host-buffer H*X in a chosen precision: X is converted on the device to the
 * precision's storage type, multiplied, and Y converted back to double.  No
 * reference counterpart (the reference has only the FP64 product). */
int pisces_apply_h_ex(pisces_ctx *ctx, const double *X, double *Y, int nvec, int precision);

/* Times `iters` H*X applications on `nvec` device-resident vectors (uniform in
 * [-1,1], seeded) with HIP events on the context's stream after `warmup`
 * untimed ones; *ms_per_apply receives the mean time of one application (both
 * passes).  Measurement hook for the benchmark; no reference counterpart. */
int pisces_time_hx(pisces_ctx *ctx, int nvec, int precision, int warmup, int iters,
                   double *ms_per_apply);

/* Kinetic preconditioner t = scale * (Tx (+) Ty (+) Tz - shift)^{-1} r for nvec
 * host vectors (nvec*ngp doubles each way), applied through the eigenbases of
 * the 1-D kinetic matrices with |lam - shift| floored at `floor`; precision
 * PISCES_FP64 or PISCES_BF16X2 (split-BF16 products, BF16 intermediates between
 * the three plane passes: about 3e-3 relative, enough for a Davidson
 * correction, which only steers the search).  The block solver's correction step; exposed
 * for testing.  No reference counterpart (the reference corrects with the
 * diagonal, src/Davidson.cpp:557-566). */
int pisces_kinetic_precond(pisces_ctx *ctx, int nvec, const double *r, double shift,
                           double scale, double floor, int precision, double *t);

/* Allocation-free form of pisces_kinetic_precond.  It has no counterpart in
 * the reference; it is for callers that cannot allocate device memory (code
 * under stream capture, serving loops that forbid hipMalloc).  d_r and d_t are
 * distinct device arrays of nvec*ngp doubles.  The per-vector task
 * descriptors and, in the split-BF16 form, the BF16 intermediate between the
 * three plane passes live in a fixed array of pisces_noalloc_scratch_bytes()
 * bytes (32 MiB) that each GPU holds until the process ends, so the split-BF16
 * form takes at most about 16.7 million grid points x vectors (64^3 x 63
 * vectors); larger requests return PISCES_ERR_ARG before any launch.  All
 * calls on one GPU share that array: calls that overlap in time on one GPU
 * (other threads, other contexts) race.  The array's address on a device is
 * looked up when the first context is created there.  Results equal
 * pisces_kinetic_precond's bit for bit (same kernels); synchronises the
 * context's stream before returning. */
size_t pisces_noalloc_scratch_bytes(void);
int pisces_kinetic_precond_noalloc(pisces_ctx *ctx, int nvec, const double *d_r, double shift,
                                   double scale, double floor, int precision, double *d_t);

/* Times `iters` applications of the kinetic preconditioner (all three passes)
 * to `nvec` device-resident vectors, as pisces_time_hx does for H*X: FP64
 * vectors for PISCES_FP64, FP32 vectors with the split-BF16 form for
 * PISCES_BF16X2.  Measurement hook; no reference counterpart. */
int pisces_time_precond(pisces_ctx *ctx, int nvec, int precision, int warmup, int iters,
                        double *ms_per_apply);

/* Plain device-to-device copy of `bytes` (multiple of 16) by a streaming
 * kernel, `reps` timed launches after one warm-up; *ms = mean launch time.
 * Calibrates hardware byte counters and the achievable HBM bandwidth.
 * Measurement hook; no reference counterpart. */
int pisces_copy_probe(int device, long long bytes, int reps, double *ms);

/* Read-only stream over `bytes` (multiple of 16) of device memory, with the
 * nontemporal load hint if `nontemporal` != 0; `reps` timed launches after
 * one warm-up, *ms = mean launch time.  The read bandwidth a plain stream
 * reaches under each cache policy.  Measurement hook; no reference
 * counterpart. */
int pisces_stream_probe(int device, long long bytes, int reps, int nontemporal, double *ms);

/* diag(H) on the host, ngp doubles.  Replaces DVR::ComputeDiagonal
 * (src/DVR.cpp:1808-1881). */
int pisces_diag(pisces_ctx *ctx, double *diag);

/* Half-box particle-in-a-box start vector (DVR::ParticleInAnDBoxWf,
 * src/DVR.cpp:1725-1805), ngp doubles on the host. */
int pisces_piab(pisces_ctx *ctx, double *wf);

typedef struct {
    int nconv;        /* converged roots                                      */
    int n_matvec;     /* H*x products (vectors) in the final precision         */
    int n_matvec_lo;  /* H*x products in the low precision (mixed forms)       */
    int n_iter;       /* Davidson iterations (subspace expansions)             */
    int n_macro;      /* macro iterations (restarts + 1)                      */
    int n_orth_data;  /* block solver, FP32-stored forms: steps whose second
                         orthogonalisation round ran on the stored columns
                         (the rest took it from the Gram algebra); 0 otherwise */
    double seconds;   /* wall time of the solve                                */
    double resid[64]; /* final FP64 residual norms ||H u - lambda u||, per root */
} pisces_stats;

/* Davidson eigensolve on the GPU with the context's potential.
 * Replaces DVR::Diagonalize case 2 -> davdriver -> Davidson (src/DVR.cpp:
 * 490-504, src/davdriver.cpp:28-96, src/Davidson.cpp:61-376):
 *   nroots    eigenpairs wanted (nStates)
 *   maxsub    maximal subspace size (maxSubspace)
 *   maxmacro  maximal macro iterations (maxIter)
 *   ptol      residual threshold 10^-ptol (pTol)
 *   corrflag  0 residual, 1 Davidson, 2 zeroth-order Jacobi-Davidson correction
 *   vecs      in: start vectors (nroots*ngp); out: Ritz vectors
 *   evals     out: eigenvalues (nroots)
 *   precision PISCES_FP64 follows the reference algorithm step by step;
 *             PISCES_FP32/BF16X2 run the low-precision contractions with FP64
 *             defect correction (FP64 residuals every refresh_every iterations
 *             and at convergence).
 * Returns the number of converged roots (>= 0) or an error code. */
int pisces_solve(pisces_ctx *ctx, int nroots, int maxsub, int maxmacro, int ptol,
                 int corrflag, int precision, double *vecs, double *evals,
                 pisces_stats *stats);

/* Block Davidson on the GPU (the solver behind the benchmark and the server):
 * `block` Ritz pairs are tracked and corrected per iteration (block >= nroots;
 * the extra ones guard the wanted roots against near-degenerate neighbours),
 * the subspace holds at most maxsub vectors and is thick-restarted from the
 * block's Ritz vectors.  precision selects the storage of the correction
 * basis and the H*X form used on it (FP64, FP32 MFMA or split-BF16 MFMA); the
 * Ritz block and its H-image are FP64 throughout and, in the low-precision
 * forms, H*X of the Ritz block is recomputed in FP64 at every restart (defect
 * correction), so reported eigenvalues and residuals are FP64-exact.
 * Converged when ||H u_i - theta_i u_i|| < 10^-ptol for i < nroots.  The
 * FP32-stored forms converged to ptol 8 on every case tried; at ptol 9 they
 * converge on most but stalled on a separable (degenerate) model potential,
 * which the FP64 form solves: ask for FP64 where 1e-9 is required.
 * precond: PISCES_PRECOND_DIAG, the reference's Davidson correction
 * r / (theta - diag H) (src/Davidson.cpp:557-566), or PISCES_PRECOND_KINETIC,
 * t = -(T - theta)^{-1} r applied exactly through the eigenbases of the 1-D
 * sine-DVR kinetic matrices (fast diagonalisation; an extension).
 * vecs: in nroots start vectors (ngp doubles each; the rest of the block is
 * seeded random), out nroots Ritz vectors.  Returns converged roots or an
 * error.  Generalises the reference's one-vector-per-root Davidson
 * (src/Davidson.cpp:61-376) to blocks; it has no direct counterpart. */
int pisces_solve_block(pisces_ctx *ctx, int nroots, int block, int maxsub, int max_iter,
                       int ptol, int precond, int precision, double *vecs, double *evals,
                       pisces_stats *stats);

/* Reverse-communication Davidson with the reference's exact calling
 * convention and semantics (src/Davidson.h:2-15): return 1 asks the caller to
 * set Z[inout[1]*ndim ..] = H B[inout[0]*ndim ..] for inout[2] vectors, return
 * 0 means done (converged vectors in B[0..nroots), eigenvalues in evals).
 * B, Z, diag and davwork are host arrays owned by the caller (davwork of
 * pisces_davidson_work_size doubles); the block operations run on the GPU of
 * ctx.  Unlike the reference it keeps its state in ctx (re-entrant). */
int pisces_davidson_work_size(int ndim, int maxsub, int nroots, int corrflag);
int pisces_davidson(pisces_ctx *ctx, int ndim, int maxsub, int nroots, int maxmacro,
                    int tol, int corrflag, int verbose, double *evals, int *nConv,
                    double *B, double *Z, const double *diag, double *davwork,
                    int *inout);

/* ---------------------------------------------------------------------------
 * Batched / served solver (no reference counterpart: the reference solves one
 * cluster at a time).  Many independent instances of (possibly) different grid
 * sizes are advanced together, one set of batched kernels per Davidson step,
 * with per-instance convergence masks.
 * ------------------------------------------------------------------------- */
typedef struct pisces_server pisces_server;

typedef struct {
    int n[3];            /* grid points                                   */
    double L[3];         /* box lengths (bohr)                            */
    int nw;              /* waters                                        */
    const double *xyz;   /* 9*nw, Angstrom (O,H1,H2)                      */
    int nion;            /* optional ions (bohr positions, charges)       */
    const double *ion_xyz;
    const double *ion_q;
    int nroots;          /* states wanted                                 */
    int warm;            /* 1: start from the chain's previous subspace   */
    long long chain;     /* chain id for warm starts (-1: none)           */
    int want_vectors;    /* 1: keep the Ritz vectors for pisces_server_vectors */
} pisces_request;

typedef struct {
    long long id;        /* request id returned by pisces_server_submit   */
    int status;          /* nconv >= 0, or an error code                  */
    int nroots;
    double evals[8];
    double resid[8];
    int n_iter;
    int n_matvec;
    double latency_s;    /* submit -> completion                          */
} pisces_result;

/* hbm_cap_bytes: admission cap on device memory used by live instances;
 * max_slots: instances advanced together; precision as for pisces_solve. */
int pisces_server_create(pisces_server **srv, int device, size_t hbm_cap_bytes,
                         int max_slots, int precision, int maxsub, int ptol);
void pisces_server_destroy(pisces_server *srv);
/* Queue a request; returns its id (>= 0) or an error code. */
long long pisces_server_submit(pisces_server *srv, const pisces_request *req);
/* Advance every live instance by one Davidson step, admitting queued requests
 * as slots and memory free up.  Returns the number of live instances. */
int pisces_server_step(pisces_server *srv);
/* Copy out up to max finished results; returns how many were copied. */
int pisces_server_poll(pisces_server *srv, pisces_result *out, int max);
/* Device bytes an instance of this request would hold (the admission cost). */
size_t pisces_server_footprint(const pisces_server *srv, const pisces_request *req);
/* Fetch the Ritz vectors of a finished request that asked for them
 * (want_vectors; nroots*ngp doubles); those of the 64 most recently finished
 * such requests are kept on the host. */
int pisces_server_vectors(pisces_server *srv, long long id, double *vecs);
/* Requests queued and not yet admitted; device bytes held by live instances. */
int pisces_server_pending(const pisces_server *srv);
size_t pisces_server_used_bytes(const pisces_server *srv);

#ifdef __cplusplus
}
#endif
#endif
