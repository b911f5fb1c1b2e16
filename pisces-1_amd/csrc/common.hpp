// common.hpp -- shared declarations of the PISCES MI355X library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pisces_hip.h"

namespace pisces {

// Error handling: every HIP call goes through PC_HIP, which turns a failure into
// a thrown hip_error; the C-ABI layer converts exceptions into return codes.
struct hip_error {
    hipError_t err;
    const char *what;
    int line;
};
struct api_error {
    int code;
    std::string msg;
};

#define PC_HIP(call)                                                     \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) throw ::pisces::hip_error{e_, #call, __LINE__}; \
    } while (0)

#define PC_REQUIRE(cond, code, msg)                                      \
    do {                                                                 \
        if (!(cond)) throw ::pisces::api_error{(code), (msg)};          \
    } while (0)

void set_last_error(const std::string &s);

// Device buffer owned by RAII.
struct dbuf {
    void *p = nullptr;
    size_t bytes = 0;
    dbuf() = default;
    explicit dbuf(size_t b) { alloc(b); }
    dbuf(const dbuf &) = delete;
    dbuf &operator=(const dbuf &) = delete;
    dbuf(dbuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    dbuf &operator=(dbuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~dbuf() { release(); }
    void alloc(size_t b) {
        release();
        if (b) PC_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    void ensure(size_t b) { if (b > bytes) alloc(b); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// --------------------------------------------------------------- axis operators
// Device operand matrices of one grid axis with n points, zero padded to
// np = n rounded up to a multiple of 32, row-major (element (r, c) at r*np + c):
// the kinetic matrix T (symmetric) in FP64, FP32 and split BF16 (hi = bf16(t),
// lo = bf16(t - hi) of the FP32 value), and the eigenvectors Q (T = Q diag(lam)
// Q^T, eigenvectors in the columns) and Q^T in FP64 and split BF16.  BF16 is
// held as its 16-bit pattern.
struct AxisMats {
    int n = 0, np = 0;
    const double *t64 = nullptr;
    const float *t32 = nullptr;
    const uint16_t *th = nullptr, *tl = nullptr;
    const double *q64 = nullptr, *qt64 = nullptr;
    const uint16_t *qh = nullptr, *ql = nullptr, *qth = nullptr, *qtl = nullptr;
    const double *lam = nullptr;
};

// Owner of one axis's device matrices.
struct AxisStore {
    dbuf buf;
    AxisMats m;
    // T: n x n FP64 kinetic matrix (row-major, symmetric); uploads on stream s
    // and waits for the copy.
    void build(int n, const double *T, hipStream_t s);
};

inline int axis_pad(int n) { return (n + 31) & ~31; }

// --------------------------------------------------------------- H*X
// One instance's share of an H*X launch: Y[:,v] = H X[:,v], v < nvec.
struct HxJob {
    int nx, ny, nz, nvec;
    AxisMats ax[3];     // kinetic operators of x, y, z
    const void *V;      // diagonal potential (ngp), job precision
    const void *X;      // vectors, X + v*ld
    void *Y;
    long long ld;
};

// precision codes: PISCES_FP64 / PISCES_FP32 / PISCES_BF16X2
// variant 0 selects the LDS-tiled VALU kernels (FP64 and FP32; a cross-check
// of the matrix-core kernels, which every other value selects)
void hx_launch(const HxJob *d_jobs, int njobs, int max_planes_x, int max_planes_z,
               int max_n, int precision, hipStream_t s, int variant = 1);
// Split-BF16 H*X with both passes in one launch, ordered per XCD so that
// pass B reads X and Y from the L2 that pass A left them in.  The plan (ints,
// hx_fused_plan_ints of them) is built on the host from the host copy of the
// jobs and must reach the device unchanged (it carries zeroed counters):
// every launch takes a freshly uploaded plan.
int hx_fused_plan_ints(const HxJob *h_jobs, int njobs);
void hx_fused_plan(const HxJob *h_jobs, int njobs, int *plan);
void hx_fused_launch(const HxJob *h_jobs, const HxJob *d_jobs, int njobs, const int *h_plan,
                     int *d_plan, int max_n, hipStream_t s);
// waits of the fused launch that hit their spin limit (0 unless something is broken)
unsigned hx_fused_spin_timeouts();
// whether split-BF16 H*X takes the one-launch form (default; PISCES_HX_FUSED=0: two launches)
bool hx_fused_enabled();
bool one_launch_available();
// matrix-core plane kernels (both passes)
void hx_mfma_launch(const HxJob *d_jobs, int njobs, int max_planes_a, int max_planes_b,
                    int max_n, int precision, hipStream_t s);

// --------------------------------------------------------------- potential
struct PotParams {
    int damp_type;     // 1 Gaussian, 2 effective distance
    int rep_type;      // 0 none, 1 Gaussians exp(-a R^2), 2 Slaters exp(-a R)
    int pol_type;      // 0 none, 1 non-interacting polarizable M sites
    int adiabatic;
    double qdamp, ddamp, pdamp, alpha, eps;
    double ion_r;
};
// Builds the repulsive-core table and PotParams from PotFlag/PotPara exactly as
// Potential::SetupDPPGTOP does; returns false for unsupported flag sets.
bool potential_setup(const int *flags, const double *pp, int nsites, PotParams &P,
                     std::vector<double> &rep, std::vector<int> &irep,
                     std::vector<int> &ipps);

struct PotDev {
    int nsites, nq, nd, nrep, npps, nion;
    const double *sites;   // 3*nsites
    const double *q; const int *iq;
    const double *d; const int *id;
    const double *rep; const int *irep;
    const int *ipps;
    const double *ion_xyz; const double *ion_q;
};
void potential_grid_launch(const PotParams &P, const PotDev &D, int nx, int ny, int nz,
                           const double *d_x, const double *d_y, const double *d_z,
                           double *d_v, float *d_v32, hipStream_t s);

// Several clusters in one launch (one task per cluster, same PotParams).
struct PotTask {
    PotDev D;
    int nx, ny, nz;
    const double *gx, *gy, *gz;
    double *v;
    float *v32;
};
void potential_grid_batched_launch(const PotParams &P, const PotTask *d_tasks, int ntasks,
                                   long long max_ng, hipStream_t s);
int potential_max_sites();

// ------------------------------------- self-consistent polarization (PolType 3)
// The polarizable sites of Potential::SetupDPPGTOP case 3 on the device.
// n3p = 3 np rounded up to a multiple of 16; epc and inva are zero past 3 np.
struct Pol3Dev {
    int np, n3p, damp_type, pad;
    double pdamp, vpolww;
    const double *rpol;  // 3*np site positions
    const double *epc;   // n3p: permanent-charge field at the sites
    const double *inva;  // n3p x n3p, row-major
};
constexpr int kPol3MaxN = 576;  // largest n3p (192 polarizable sites, 64 waters)
inline int pol3_padded(int np) { return (3 * np + 15) / 16 * 16; }
// Host: InvA = (alpha^-1 - T)^-1 (Thole a = 0.3 between every pair), both
// triangles taken from the lower one, zero-padded to n3p x n3p.  false if the
// matrix is singular.
bool pol3_inva_host(int np, const double *rpol, const double *alpha, int n3p, double *inva);
bool sym_inverse(int n, std::vector<double> &A, int n3p, double *inva, int method);
// Device: d_v[g] += -E^T InvA E / 2 - vpolww at every grid point, E the
// electron's damped field at the sites plus epc.
void pol3_grid_launch(const Pol3Dev &D, int nx, int ny, int nz, const double *d_x,
                      const double *d_y, const double *d_z, double *d_v, hipStream_t s);

// --------------------------------------------------------------- water model
// Host: rigid-monomer projection of O,H1,H2 (Angstrom) to 4 sites (bohr).
bool water_sites_host(int nw, const double *xyz_ang, double *sites);
// Device: induced-dipole SCF; d_sites 12*nw, outputs d_mu/d_epc 9*nw.
void water_scf_launch(int nw, const double *d_sites, double *d_mu, double *d_epc,
                      double *d_tensor, int *d_info, hipStream_t s);
struct ScfTask {
    int nw;
    const double *sites;
    double *mu, *epc, *tens;
    int *info;
};
void water_scf_batched_launch(const ScfTask *d_tasks, int ntasks, hipStream_t s);
int water_scf_max_waters();
// Host: energy of the neutral cluster (E0, electrostatic, polarization,
// dispersion) from the sites (12*nw, bohr) and the induced dipoles (9*nw).
void water_energy_host(int nw, const double *sites, const double *mu, double out[4]);
// Charges of the DPP monomer (H1, H2, M) and polarizabilities (O, H1, H2).
void water_lists_host(int nw, std::vector<double> &q, std::vector<int> &iq,
                      std::vector<int> &id);

// --------------------------------------------------------------- host helpers
void sine_dvr_full(int n, double mass, double x0, double xnp1, double *x, double *T);
// DVRType 1: harmonic-oscillator DVR of frequency w; DVRType 20: Colbert-Miller
void ho_dvr_full(int n, double w, double *x, double *T);
void cm_dvr_full(int n, double mass, double rl, double ru, double *x, double *T);
// DVRType 3: the FFT kinetic energy as the real circulant n x n matrix
void fft_kinetic_full(int n, double mass, double x0, double xnp1, double *x, double *T);

}  // namespace pisces
