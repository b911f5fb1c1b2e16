// context.hpp -- the per-grid context behind the C ABI.
#pragma once

#include <functional>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "blockops.hpp"
#include "common.hpp"

// Device tables of the last PolType 3 potential built on a context (Potential
// 3, DampType 2, Polarization 3): what the analytic gradient evaluates.
struct GradState {
    bool valid = false;
    int nsites = 0, np = 0, n3p = 0;
    double qdamp = 0, pdamp = 0;
    pisces::dbuf site4, srep, rpol, epc, inva;
    std::vector<int> ip;
};

// One sine-DVR grid on one device, with its potential and solver workspace.
// Re-entrant: everything lives here, nothing in globals.
struct pisces_ctx {
    int device = 0;
    int dvr_type = 0;  // 0 sine DVR, 3 FFT kinetic energy
    hipStream_t stream = nullptr;
    int n[3] = {0, 0, 0};
    double L[3] = {0, 0, 0};
    long long ngp = 0;
    int maxn = 0;
    std::vector<double> x[3], T[3];  // host copies
    pisces::dbuf dx[3], dT64[3];
    pisces::AxisStore ax[3];  // padded operand matrices of the plane kernels
    pisces::dbuf dV64, dV32;
    bool has_v = false;
    int hx_variant = 1;  // H*X kernels: 1 matrix cores (default), 2 multi-launch only, 3 one-launch, 0 VALU

    // job ring for H*X launches
    static constexpr int kJobRing = 256;
    pisces::dbuf djobs, dplan;
    pisces::HxJob *hjobs = nullptr;  // pinned
    int job_slot = 0;

    // solver workspace (grown on demand)
    pisces::dbuf wB, wZ, wU, wZV, wR, wC, wJ, wdiag, wpart, wsmall, winfo, wX, wY;
    pisces::dbuf wB32, wZ32, wC32;
    double *hsmall = nullptr;  // pinned, kSmall doubles
    static constexpr int kSmall = 1 << 16;

    // reverse-communication Davidson state (pisces_davidson); type-erased so
    // that the state's definition stays in its own translation unit
    std::shared_ptr<void> rc;

    GradState gs;
    std::vector<double> last_sites;  // sites of the last potential build (3 per site)

    // device address of the library's fixed scratch array on this device
    // (pisces_kinetic_precond_noalloc), looked up at creation
    void *scratch = nullptr;

    pisces_ctx() = default;
    ~pisces_ctx();
    pisces_ctx(const pisces_ctx &) = delete;
    pisces_ctx &operator=(const pisces_ctx &) = delete;

    // device job slot for one H*X launch of this grid
    const pisces::HxJob *push_job(const pisces::HxJob &j);
    void hx(const void *X, void *Y, long long ld, int nvec, int precision);
    void sync();
};

namespace pisces {
// Address of the allocation-free entry point's scratch array on `device` (the
// current device); looked up once per device.
void *noalloc_scratch(int device);
// The electron side of the analytic gradient for the context's PolType 3
// potential and the wavefunction psi (host, ngp): gel[3 nsites], tmu[3 np],
// mcm[(3 np)^2] row-major.
void gradient_terms(pisces_ctx *c, const double *psi_host, double *gel, double *tmu, double *mcm);
// Dual grid: the coarse grid of a cubic fine grid (n points on length L), and
// the coarse -> fine transfer of a device vector (kind 0 wavefunction, 1
// potential; with sites, only the points farther than rtol from every site or
// on the coarse lattice are written).
bool dual_coarse_grid(int n, double L, int &nc, double &Lc);
void dual_transfer(pisces_ctx *c, int nc, const double *d_coarse, int kind, const double *sites,
                   int n_sites, double rtol, double *d_fine);
// Runs f, converting library exceptions into C-ABI return codes.
template <typename F>
int guarded_call(F &&f)
{
    try {
        return f();
    } catch (const hip_error &e) {
        set_last_error(std::string("HIP error ") + hipGetErrorString(e.err) + " in " + e.what);
        return PISCES_ERR_HIP;
    } catch (const api_error &e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return PISCES_ERR_NOMEM;
    }
}
// Matrix-free hooks of the FP64 reference-mirroring solver: with matvec set,
// Y = H X for nvec device columns (X at subspace column col0) is obtained from
// it instead of the context's H*X; with host_diag set, diag(H) is uploaded
// from it instead of being built from the potential.
struct Fp64Hooks {
    std::function<void(const double *X, double *Y, int col0, int nvec)> matvec;
    const double *host_diag = nullptr;
};
int solve_fp64(pisces_ctx *c, int nroots, int maxsub, int maxmacro, int ptol, int corrflag,
               double *vecs, double *evals, pisces_stats *st, const Fp64Hooks *hooks = nullptr);
int solve_mixed(pisces_ctx *c, int nroots, int maxsub, int maxmacro, int ptol, int corrflag,
                int precision, double *vecs, double *evals, pisces_stats *st);
}  // namespace pisces
