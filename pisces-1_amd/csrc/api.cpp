// api.cpp -- the C ABI (include/pisces_hip.h): contexts, H*X, potential builds,
// the water model and the single-instance solves.
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "context.hpp"
#include "fdm.hpp"
#include "prof.hpp"

using namespace pisces;

namespace pisces {
const char *last_error();
}

// ------------------------------------------------------------------ context
pisces_ctx::~pisces_ctx()
{
    rc.reset();  // stops a reverse-communication solve still waiting for its caller
    if (device >= 0) static_cast<void>(hipSetDevice(device));
    if (stream) static_cast<void>(hipStreamSynchronize(stream));
    if (hjobs) static_cast<void>(hipHostFree(hjobs));
    if (hsmall) static_cast<void>(hipHostFree(hsmall));
    if (stream) static_cast<void>(hipStreamDestroy(stream));
}

const HxJob *pisces_ctx::push_job(const HxJob &j)
{
    if (job_slot == kJobRing) {  // host staging slots are reused: drain first
        PC_HIP(hipStreamSynchronize(stream));
        job_slot = 0;
    }
    const int s = job_slot++;
    hjobs[s] = j;
    HxJob *d = djobs.as<HxJob>() + s;
    PC_HIP(hipMemcpyAsync(d, &hjobs[s], sizeof(HxJob), hipMemcpyHostToDevice, stream));
    return d;
}

void pisces_ctx::hx(const void *X, void *Y, long long ld, int nvec, int precision)
{
    if (nvec <= 0) return;
    HxJob j;
    j.nx = n[0]; j.ny = n[1]; j.nz = n[2]; j.nvec = nvec;
    const bool f64 = (precision == PISCES_FP64);
    for (int a = 0; a < 3; ++a) j.ax[a] = ax[a].m;
    j.V = f64 ? dV64.p : dV32.p;
    j.X = X; j.Y = Y; j.ld = ld;
    const HxJob *d = push_job(j);
    if (precision == PISCES_BF16X2 && (hx_variant == 3 || (hx_variant == 1 && hx_fused_enabled()))) {
        // the one-launch form: its plan (with zeroed counters) is uploaded from
        // pageable memory, which the copy has consumed when it returns
        std::vector<int> plan(hx_fused_plan_ints(&j, 1));
        hx_fused_plan(&j, 1, plan.data());
        dplan.ensure(sizeof(int) * plan.size());
        PC_HIP(hipMemcpyAsync(dplan.p, plan.data(), sizeof(int) * plan.size(),
                              hipMemcpyHostToDevice, stream));
        hx_fused_launch(&j, d, 1, plan.data(), dplan.as<int>(), maxn, stream);
        return;
    }
    hx_launch(d, 1, n[2] * nvec, n[1] * nvec, maxn, precision, stream, hx_variant);
}

void pisces_ctx::sync() { PC_HIP(hipStreamSynchronize(stream)); }

// ------------------------------------------------------------------ helpers
namespace {

template <typename F>
int guarded(F &&f)
{
    return guarded_call(std::forward<F>(f));
}

void upload(dbuf &b, const void *h, size_t bytes, hipStream_t s)
{
    b.ensure(bytes);
    PC_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
}

// The kinetic preconditioner on device tasks (host copies in h): the split-
// BF16 form in one XCD-ordered launch where the context's H*X variant is the
// default, else in three launches.
void precond_launch(pisces_ctx *c, const std::vector<FdmTask> &h, const FdmTask *d, int precision)
{
    const int n = (int)h.size();
    if (precision == PISCES_BF16X2 && (c->hx_variant == 3 || (c->hx_variant == 1 && fdm_fused_enabled()))) {
        std::vector<int> plan(fdm_fused_plan_ints(n));
        fdm_fused_plan(h.data(), n, plan.data());
        c->dplan.ensure(sizeof(int) * plan.size());
        PC_HIP(hipMemcpyAsync(c->dplan.p, plan.data(), sizeof(int) * plan.size(),
                              hipMemcpyHostToDevice, c->stream));
        fdm_fused_launch(d, c->dplan.as<int>(), plan.data(), c->maxn, c->stream);
        return;
    }
    fdm_launch(d, n, c->maxn, precision, c->stream);
}

void set_potential_device(pisces_ctx *c)
{
    // the FP32 copy for the low-precision H*X forms
    c->dV32.ensure(sizeof(float) * c->ngp);
    to_f32(c->dV64.as<double>(), c->dV32.as<float>(), c->ngp, c->stream);
    c->has_v = true;
}

}  // namespace

extern "C" {

const char *pisces_version(void) { return "pisces-mi355x 0.1 (gfx950)"; }

const char *pisces_strerror(int code)
{
    switch (code) {
    case PISCES_OK: return "ok";
    case PISCES_ERR_ARG: return "invalid argument";
    case PISCES_ERR_HIP: return "HIP runtime error";
    case PISCES_ERR_NOMEM: return "out of memory";
    case PISCES_ERR_UNSUPPORTED: return "unsupported option";
    case PISCES_ERR_STATE: return "invalid state";
    case PISCES_ERR_NUMERIC: return "numerical failure";
    default: return code > 0 ? "not converged" : "unknown error";
    }
}

const char *pisces_last_error(void) { return pisces::last_error(); }

int pisces_create(pisces_ctx **out, const int n[3], const double L[3], int device)
{
    return pisces_create_ex(out, n, L, device, 0);
}

int pisces_create_ex(pisces_ctx **out, const int n[3], const double L[3], int device, int dvr_type)
{
    if (!out || !n || !L) return PISCES_ERR_ARG;
    if (dvr_type != 0 && dvr_type != 1 && dvr_type != 3 && dvr_type != 20) {
        set_last_error("DVR type must be 0 (sine), 1 (harmonic oscillator), 3 (FFT) or 20 "
                       "(Colbert-Miller)");
        return PISCES_ERR_UNSUPPORTED;
    }
    // Tdiag's wave numbers m = a (a < n/2), a - n are symmetric only for even n
    if (dvr_type == 3 && ((n[0] | n[1] | n[2]) & 1)) {
        set_last_error("DVR type 3 needs even grid sizes");
        return PISCES_ERR_UNSUPPORTED;
    }
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (n[a] < 2 || n[a] > 128 || !(L[a] > 0)) {
            set_last_error("grid dimensions must be 2..128 and lengths positive");
            return PISCES_ERR_ARG;
        }
    pisces_ctx *c = new (std::nothrow) pisces_ctx();
    if (!c) return PISCES_ERR_NOMEM;
    int rc = guarded([&] {
        c->device = device;
        c->dvr_type = dvr_type;
        PC_HIP(hipSetDevice(device));
        c->scratch = noalloc_scratch(device);
        PC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->ngp = 1;
        for (int a = 0; a < 3; ++a) {
            c->n[a] = n[a];
            c->L[a] = L[a];
            c->ngp *= n[a];
            c->maxn = std::max(c->maxn, n[a]);
            c->x[a].resize(n[a]);
            c->T[a].resize((size_t)n[a] * n[a]);
            // DVR::ComputeGridPointsAndKineticEnergy (src/DVR.cpp:583-606) on
            // [-L/2, L/2], mass 1: sine_dvr, or Tdiag for DVRType 3
            if (dvr_type == 3)
                fft_kinetic_full(n[a], 1.0, -0.5 * L[a], 0.5 * L[a], c->x[a].data(), c->T[a].data());
            else if (dvr_type == 1)  // L[a] is the oscillator frequency (Omega0)
                ho_dvr_full(n[a], L[a], c->x[a].data(), c->T[a].data());
            else if (dvr_type == 20)
                cm_dvr_full(n[a], 1.0, -0.5 * L[a], 0.5 * L[a], c->x[a].data(), c->T[a].data());
            else
                sine_dvr_full(n[a], 1.0, -0.5 * L[a], 0.5 * L[a], c->x[a].data(), c->T[a].data());
            upload(c->dx[a], c->x[a].data(), sizeof(double) * n[a], c->stream);
            upload(c->dT64[a], c->T[a].data(), sizeof(double) * n[a] * n[a], c->stream);
            c->ax[a].build(n[a], c->T[a].data(), c->stream);
        }
        PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->hjobs),
                             sizeof(HxJob) * pisces_ctx::kJobRing, hipHostMallocDefault));
        PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->hsmall),
                             sizeof(double) * pisces_ctx::kSmall, hipHostMallocDefault));
        c->djobs.alloc(sizeof(HxJob) * pisces_ctx::kJobRing);
        c->dV64.alloc(sizeof(double) * c->ngp);
        return 0;
    });
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return PISCES_OK;
}

void pisces_destroy(pisces_ctx *c) { delete c; }

int pisces_grid(const pisces_ctx *c, double *xyz, double *T)
{
    if (!c) return PISCES_ERR_ARG;
    size_t ox = 0, ot = 0;
    for (int a = 0; a < 3; ++a) {
        if (xyz) std::memcpy(xyz + ox, c->x[a].data(), sizeof(double) * c->n[a]);
        if (T) std::memcpy(T + ot, c->T[a].data(), sizeof(double) * c->n[a] * c->n[a]);
        ox += c->n[a];
        ot += (size_t)c->n[a] * c->n[a];
    }
    return PISCES_OK;
}

int pisces_set_potential(pisces_ctx *c, const double *v)
{
    if (!c || !v) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        PC_HIP(hipMemcpyAsync(c->dV64.p, v, sizeof(double) * c->ngp, hipMemcpyHostToDevice, c->stream));
        set_potential_device(c);
        c->sync();
        return 0;
    });
}

int pisces_get_potential(const pisces_ctx *c, double *v)
{
    if (!c || !v) return PISCES_ERR_ARG;
    if (!c->has_v) return PISCES_ERR_STATE;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        PC_HIP(hipMemcpyAsync(v, c->dV64.p, sizeof(double) * c->ngp, hipMemcpyDeviceToHost, c->stream));
        PC_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

namespace {

// Potential::Setup + DVR::ComputePotential: the site terms on the grid, then
// (np > 0, PolType 3) the self-consistent polarization added on top
int build_potential_impl(pisces_ctx *c, const int flags[6], const double *potpara, int nsites,
                         const double *sites, int nq, const double *q, const int *iq, int nd,
                         const double *d, const int *id, int np, const double *alpha,
                         const int *ip, const double *epc, double vpol_ww, int nion,
                         const double *ion_xyz, const double *ion_q, double ion_r)
{
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        PotParams P{};
        std::vector<double> rep;
        std::vector<int> irep, ipps;
        PC_REQUIRE(potential_setup(flags, potpara, nsites, P, rep, irep, ipps),
                   PISCES_ERR_UNSUPPORTED, "unsupported potential flags");
        PC_REQUIRE((P.pol_type == 3) == (np > 0), PISCES_ERR_ARG,
                   "PolType 3 takes the polarizable-site lists (pisces_build_potential_scp)");
        P.ion_r = ion_r;
        c->last_sites.assign(sites, sites + 3 * (size_t)nsites);
        dbuf ds, dq, diq, dd, did, drep, direp, dipps, dix, diq2, drp, depc, dinv;
        upload(ds, sites, sizeof(double) * 3 * nsites, c->stream);
        if (nq) { upload(dq, q, sizeof(double) * nq, c->stream); upload(diq, iq, sizeof(int) * nq, c->stream); }
        if (nd) { upload(dd, d, sizeof(double) * 3 * nd, c->stream); upload(did, id, sizeof(int) * nd, c->stream); }
        if (!rep.empty()) {
            upload(drep, rep.data(), sizeof(double) * rep.size(), c->stream);
            upload(direp, irep.data(), sizeof(int) * irep.size(), c->stream);
        }
        if (!ipps.empty()) upload(dipps, ipps.data(), sizeof(int) * ipps.size(), c->stream);
        if (nion) {
            upload(dix, ion_xyz, sizeof(double) * 3 * nion, c->stream);
            upload(diq2, ion_q, sizeof(double) * nion, c->stream);
        }
        PotDev D{nsites, nq, nd, (int)irep.size(), (int)ipps.size(), nion,
                 ds.as<double>(), dq.as<double>(), diq.as<int>(), dd.as<double>(), did.as<int>(),
                 drep.as<double>(), direp.as<int>(), dipps.as<int>(), dix.as<double>(),
                 diq2.as<double>()};
        potential_grid_launch(P, D, c->n[0], c->n[1], c->n[2], c->dx[0].as<double>(),
                              c->dx[1].as<double>(), c->dx[2].as<double>(), c->dV64.as<double>(),
                              nullptr, c->stream);
        if (np > 0) {
            // SetupDPPGTOP case 3 (src/Potential.cpp:387-432): InvA of the
            // polarizable sites on the host, the grid sum on the matrix cores
            const int n3p = pol3_padded(np);
            PC_REQUIRE(n3p <= kPol3MaxN, PISCES_ERR_UNSUPPORTED, "too many polarizable sites");
            std::vector<double> rp((size_t)3 * np), ep((size_t)n3p, 0.0), inv((size_t)n3p * n3p);
            for (int i = 0; i < np; ++i)
                for (int k = 0; k < 3; ++k) {
                    rp[3 * i + k] = sites[3 * ip[i] + k];
                    ep[3 * i + k] = epc[3 * i + k];
                }
            PC_REQUIRE(pol3_inva_host(np, rp.data(), alpha, n3p, inv.data()), PISCES_ERR_NUMERIC,
                       "singular polarization matrix");
            upload(drp, rp.data(), sizeof(double) * rp.size(), c->stream);
            upload(depc, ep.data(), sizeof(double) * ep.size(), c->stream);
            upload(dinv, inv.data(), sizeof(double) * inv.size(), c->stream);
            Pol3Dev Q{np, n3p, P.damp_type, 0, P.pdamp, vpol_ww, drp.as<double>(),
                      depc.as<double>(), dinv.as<double>()};
            {
                ProfScope ps(PK_POL3, c->stream, 0.0);
                pol3_grid_launch(Q, c->n[0], c->n[1], c->n[2], c->dx[0].as<double>(),
                                 c->dx[1].as<double>(), c->dx[2].as<double>(),
                                 c->dV64.as<double>(), c->stream);
            }
            // what the analytic gradient evaluates (EvaluateDPPGTOPGradient is
            // written for Potential 3 with DampType 2): per site its charge and
            // Slater cores, the polarizable sites, their field and InvA
            std::vector<double> s4((size_t)4 * nsites, 0.0), sr((size_t)4 * nsites, 0.0);
            bool ok = flags[0] == 3 && P.damp_type == 2 && nion == 0;
            for (int s = 0; s < nsites; ++s)
                for (int k = 0; k < 3; ++k) s4[4 * s + k] = sites[3 * s + k];
            for (int i = 0; i < nq; ++i) {
                ok = ok && s4[4 * iq[i] + 3] == 0.0;  // one charge per site
                s4[4 * iq[i] + 3] = q[i];
            }
            std::vector<int> nr((size_t)nsites, 0);
            for (size_t i = 0; i < irep.size(); ++i) {
                const int s = irep[i];
                if (nr[s] >= 2) { ok = false; break; }
                sr[4 * s + 2 * nr[s]] = rep[2 * i];
                sr[4 * s + 2 * nr[s] + 1] = rep[2 * i + 1];
                ++nr[s];
            }
            c->gs.valid = false;
            if (ok) {
                GradState &G = c->gs;
                G.nsites = nsites;
                G.np = np;
                G.n3p = n3p;
                G.qdamp = P.qdamp;
                G.pdamp = P.pdamp;
                upload(G.site4, s4.data(), sizeof(double) * s4.size(), c->stream);
                upload(G.srep, sr.data(), sizeof(double) * sr.size(), c->stream);
                G.rpol = std::move(drp);
                G.epc = std::move(depc);
                G.inva = std::move(dinv);
                G.ip.assign(ip, ip + np);
                G.valid = true;
            }
        } else {
            c->gs.valid = false;
        }
        set_potential_device(c);
        c->sync();
        return 0;
    });
}

bool lists_ok(int n, const int *idx, int nsites)
{
    for (int i = 0; i < n; ++i)
        if (idx[i] < 0 || idx[i] >= nsites) return false;
    return true;
}

}  // namespace

int pisces_build_potential(pisces_ctx *c, const int flags[6], const double *potpara, int nsites,
                           const double *sites, int nq, const double *q, const int *iq, int nd,
                           const double *d, const int *id, int nion, const double *ion_xyz,
                           const double *ion_q, double ion_r)
{
    if (!c || !flags || !potpara || nsites <= 0 || !sites || nq < 0 || nd < 0 || nion < 0)
        return PISCES_ERR_ARG;
    if ((nq && (!q || !iq)) || (nd && (!d || !id)) || (nion && (!ion_xyz || !ion_q)))
        return PISCES_ERR_ARG;
    if (!lists_ok(nq, iq, nsites) || !lists_ok(nd, id, nsites)) return PISCES_ERR_ARG;
    return build_potential_impl(c, flags, potpara, nsites, sites, nq, q, iq, nd, d, id, 0,
                                nullptr, nullptr, nullptr, 0.0, nion, ion_xyz, ion_q, ion_r);
}

int pisces_build_potential_scp(pisces_ctx *c, const int flags[6], const double *potpara,
                               int nsites, const double *sites, int nq, const double *q,
                               const int *iq, int np, const double *alpha, const int *ip,
                               const double *epc, double vpol_ww, int nion,
                               const double *ion_xyz, const double *ion_q, double ion_r)
{
    if (!c || !flags || !potpara || nsites <= 0 || !sites || nq < 0 || np <= 0 || nion < 0)
        return PISCES_ERR_ARG;
    if ((nq && (!q || !iq)) || !alpha || !ip || !epc || (nion && (!ion_xyz || !ion_q)))
        return PISCES_ERR_ARG;
    if (flags[3] != 3) return PISCES_ERR_ARG;
    if (!lists_ok(nq, iq, nsites) || !lists_ok(np, ip, nsites)) return PISCES_ERR_ARG;
    for (int i = 0; i < np; ++i)
        if (!(alpha[i] > 0.0)) {
            set_last_error("polarizabilities must be positive");
            return PISCES_ERR_ARG;
        }
    // the water model's point dipoles are not passed: PolType 3 ignores them
    // (src/Potential.cpp:327-335)
    return build_potential_impl(c, flags, potpara, nsites, sites, nq, q, iq, 0, nullptr, nullptr,
                                np, alpha, ip, epc, vpol_ww, nion, ion_xyz, ion_q, ion_r);
}

int pisces_build_cluster_potential(pisces_ctx *c, int nw, const double *xyz_ang,
                                   const int flags[6], const double *potpara, int nion,
                                   const double *ion_xyz_bohr, const double *ion_q, double ion_r,
                                   double *sites_out, double *mu_out)
{
    if (!c || nw <= 0 || !xyz_ang || !flags || !potpara || nion < 0) return PISCES_ERR_ARG;
    if (nion && (!ion_xyz_bohr || !ion_q)) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        std::vector<double> sites((size_t)12 * nw);
        PC_REQUIRE(water_sites_host(nw, xyz_ang, sites.data()), PISCES_ERR_NUMERIC,
                   "water projection produced non-finite coordinates");
        const int ns = 3 * nw;
        dbuf dsites, dmu, depc, dtens, dinfo;
        upload(dsites, sites.data(), sizeof(double) * sites.size(), c->stream);
        dmu.alloc(sizeof(double) * 3 * ns);
        depc.alloc(sizeof(double) * 3 * ns);
        dtens.alloc(sizeof(double) * 6 * (size_t)ns * ns);
        dinfo.alloc(sizeof(int) * 2);
        water_scf_launch(nw, dsites.as<double>(), dmu.as<double>(), depc.as<double>(),
                         dtens.as<double>(), dinfo.as<int>(), c->stream);
        int info[2] = {0, 0};
        PC_HIP(hipMemcpyAsync(info, dinfo.p, sizeof(info), hipMemcpyDeviceToHost, c->stream));
        std::vector<double> mu((size_t)3 * ns);
        PC_HIP(hipMemcpyAsync(mu.data(), dmu.p, sizeof(double) * mu.size(), hipMemcpyDeviceToHost, c->stream));
        PC_HIP(hipStreamSynchronize(c->stream));
        if (sites_out) std::memcpy(sites_out, sites.data(), sizeof(double) * sites.size());
        if (mu_out) std::memcpy(mu_out, mu.data(), sizeof(double) * mu.size());
        std::vector<double> q;
        std::vector<int> iq, id;
        water_lists_host(nw, q, iq, id);
        int rc;
        if (flags[3] == 3) {
            // ReportPolSitesAndField (src/DPP.cpp:540-642): the sites O,H1,H2
            // with their polarizabilities and permanent-charge field; and
            // VpolWaterWater = the water-water polarization energy
            // (src/ClusterAnion.cpp:173; WaterWaterPolarizationEnergy,
            // src/DPP.cpp:1935-2024, summed over all pairs: -mu.Epc / 2)
            std::vector<double> epc((size_t)3 * ns);
            PC_HIP(hipMemcpyAsync(epc.data(), depc.p, sizeof(double) * epc.size(), hipMemcpyDeviceToHost,
                                  c->stream));
            PC_HIP(hipStreamSynchronize(c->stream));
            const double a3 = 1.0 / (0.52917720859 * 0.52917720859 * 0.52917720859);
            std::vector<double> alpha((size_t)ns);
            std::vector<int> ip((size_t)ns);
            double eww = 0.0;
            for (int i = 0; i < ns; ++i) {
                alpha[i] = (i % 3 == 0 ? 1.22 : 0.28) * a3;  // src/ttm_data.h:51-52
                ip[i] = 4 * (i / 3) + i % 3;
            }
            for (int k = 0; k < 3 * ns; ++k) eww += mu[k] * epc[k];
            rc = pisces_build_potential_scp(c, flags, potpara, 4 * nw, sites.data(), (int)q.size(),
                                            q.data(), iq.data(), ns, alpha.data(), ip.data(),
                                            epc.data(), -0.5 * eww, nion, ion_xyz_bohr, ion_q,
                                            ion_r);
        } else {
            rc = pisces_build_potential(c, flags, potpara, 4 * nw, sites.data(), (int)q.size(),
                                        q.data(), iq.data(), (int)id.size(), mu.data(), id.data(),
                                        nion, ion_xyz_bohr, ion_q, ion_r);
        }
        if (rc < 0) return rc;
        return info[0] ? 0 : 1;
    });
}

int pisces_gradient_terms(pisces_ctx *c, const double *psi, double *gel, double *tmu, double *mcm)
{
    if (!c || !psi || !gel || !tmu || !mcm) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        gradient_terms(c, psi, gel, tmu, mcm);
        return 0;
    });
}

int pisces_compute_gradient(pisces_ctx *c, const double *psi, int nsites, double *gradient,
                            const double *dT_x, const double *dT_y, const double *dT_z,
                            const double *polgrad, const double *dEfield)
{
    if (!c || !psi || !gradient || !dT_x || !dT_y || !dT_z || !polgrad || !dEfield)
        return PISCES_ERR_ARG;
    if (!c->gs.valid) {
        set_last_error("the gradient needs a PolType 3 potential built on this context");
        return PISCES_ERR_STATE;
    }
    if (nsites != c->gs.nsites || c->gs.np * 4 != nsites * 3) {
        set_last_error("nsites must be the potential's site count, 4 per water, 3 polarizable");
        return PISCES_ERR_ARG;
    }
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const int np = c->gs.np, n3 = 3 * np, n9 = 9 * np;
        std::vector<double> loc((size_t)3 * nsites), tmu((size_t)n3), mcm((size_t)n3 * n3);
        gradient_terms(c, psi, loc.data(), tmu.data(), mcm.data());
        // Potential::FinalGradient (src/Gradients.cpp:240-268): the derivative
        // of the permanent-charge field dotted with sum psi^2 mu, and for the
        // polarizable sites (not M) the dipole-tensor derivative against
        // sum psi^2 mu mu^T over the rows of the site's atom
        for (int j = 0; j < nsites; ++j) {
            for (int d = 0; d < 3; ++d) {
                const double *de = dEfield + (size_t)n3 * (3 * j + d);
                double s = 0.0;
                for (int k = 0; k < n3; ++k) s += de[k] * tmu[k];
                loc[3 * j + d] -= s;
            }
            if ((j + 1) % 4 != 0) {
                const size_t off = (size_t)(3 * (j / 4) + j % 4) * n9;
                const double *T[3] = {dT_x, dT_y, dT_z};
                for (int d = 0; d < 3; ++d) {
                    double s = 0.0;
                    for (int k = 0; k < n9; ++k) s += T[d][off + k] * mcm[off + k];
                    loc[3 * j + d] -= s;
                }
            }
        }
        // the caller's water-water gradient += electron part, then
        // SubtractWWGradient (src/Gradients.cpp:270-284; src/DVR.cpp:1165-1175)
        for (int k = 0; k < 3 * nsites; ++k) gradient[k] = gradient[k] + loc[k] - polgrad[k];
        return 0;
    });
}

int pisces_dual_grid_coarse(const int n[3], const double L[3], int nc[3], double Lc[3])
{
    if (!n || !L || !nc || !Lc) return PISCES_ERR_ARG;
    if (n[1] != n[0] || n[2] != n[0] || L[1] != L[0] || L[2] != L[0]) {
        set_last_error("the dual grid needs a cubic grid");
        return PISCES_ERR_UNSUPPORTED;
    }
    int m;
    double Lm;
    if (!dual_coarse_grid(n[0], L[0], m, Lm)) return PISCES_ERR_ARG;
    for (int a = 0; a < 3; ++a) {
        nc[a] = m;
        Lc[a] = Lm;
    }
    return PISCES_OK;
}

int pisces_dual_grid_extend(pisces_ctx *c, const int nc[3], const double *coarse, int kind,
                            double *fine)
{
    if (!c || !nc || !coarse || !fine || (kind != 0 && kind != 1)) return PISCES_ERR_ARG;
    if (nc[1] != nc[0] || nc[2] != nc[0]) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const long long ncg = (long long)nc[0] * nc[0] * nc[0];
        dbuf dc(sizeof(double) * ncg), df(sizeof(double) * c->ngp);
        PC_HIP(hipMemcpyAsync(dc.p, coarse, sizeof(double) * ncg, hipMemcpyHostToDevice, c->stream));
        dual_transfer(c, nc[0], dc.as<double>(), kind, nullptr, 0, 0.0, df.as<double>());
        PC_HIP(hipMemcpyAsync(fine, df.p, sizeof(double) * c->ngp, hipMemcpyDeviceToHost, c->stream));
        PC_HIP(hipStreamSynchronize(c->stream));
        return 0;
    });
}

int pisces_dual_grid_potential(pisces_ctx *c, const int nc[3], const double *coarse_v, double rtol)
{
    if (!c || !nc || !coarse_v || nc[1] != nc[0] || nc[2] != nc[0]) return PISCES_ERR_ARG;
    if (!c->has_v || c->last_sites.empty()) {
        set_last_error("build the fine grid's potential first");
        return PISCES_ERR_STATE;
    }
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const long long ncg = (long long)nc[0] * nc[0] * nc[0];
        dbuf dc(sizeof(double) * ncg);
        PC_HIP(hipMemcpyAsync(dc.p, coarse_v, sizeof(double) * ncg, hipMemcpyHostToDevice, c->stream));
        dual_transfer(c, nc[0], dc.as<double>(), 1, c->last_sites.data(),
                      (int)(c->last_sites.size() / 3), rtol, c->dV64.as<double>());
        set_potential_device(c);
        c->sync();
        return 0;
    });
}

int pisces_water_energy(int nw, const double *sites, const double *mu, double *energies)
{
    if (nw < 0 || (nw > 0 && (!sites || !mu)) || !energies) return PISCES_ERR_ARG;
    water_energy_host(nw, sites, mu, energies);
    return PISCES_OK;
}

int pisces_apply_h(pisces_ctx *c, const double *X, double *Y, int nvec)
{
    if (!c || !X || !Y || nvec <= 0) return PISCES_ERR_ARG;
    if (!c->has_v) return PISCES_ERR_STATE;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const size_t bytes = sizeof(double) * c->ngp * nvec;
        c->wX.ensure(bytes);
        c->wY.ensure(bytes);
        PC_HIP(hipMemcpyAsync(c->wX.p, X, bytes, hipMemcpyHostToDevice, c->stream));
        c->hx(c->wX.p, c->wY.p, c->ngp, nvec, PISCES_FP64);
        PC_HIP(hipMemcpyAsync(Y, c->wY.p, bytes, hipMemcpyDeviceToHost, c->stream));
        c->sync();
        return 0;
    });
}

int pisces_apply_h_device(pisces_ctx *c, const void *dX, void *dY, int nvec, int precision)
{
    if (!c || !dX || !dY || nvec <= 0) return PISCES_ERR_ARG;
    if (precision < PISCES_FP64 || precision > PISCES_BF16X2) return PISCES_ERR_ARG;
    if (!c->has_v) return PISCES_ERR_STATE;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        c->hx(dX, dY, c->ngp, nvec, precision);
        c->sync();
        return 0;
    });
}

int pisces_apply_h_ex(pisces_ctx *c, const double *X, double *Y, int nvec, int precision)
{
    if (!c || !X || !Y || nvec <= 0) return PISCES_ERR_ARG;
    if (precision < PISCES_FP64 || precision > PISCES_BF16X2) return PISCES_ERR_ARG;
    if (precision == PISCES_FP64) return pisces_apply_h(c, X, Y, nvec);
    if (!c->has_v) return PISCES_ERR_STATE;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const long long n = c->ngp * (long long)nvec;
        c->wX.ensure(sizeof(double) * n);
        c->wY.ensure(sizeof(double) * n);
        c->wB32.ensure(sizeof(float) * n);
        c->wZ32.ensure(sizeof(float) * n);
        PC_HIP(hipMemcpyAsync(c->wX.p, X, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        to_f32(c->wX.as<double>(), c->wB32.as<float>(), n, c->stream);
        c->hx(c->wB32.p, c->wZ32.p, c->ngp, nvec, precision);
        to_f64(c->wZ32.as<float>(), c->wY.as<double>(), n, c->stream);
        PC_HIP(hipMemcpyAsync(Y, c->wY.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        c->sync();
        return 0;
    });
}

int pisces_time_hx(pisces_ctx *c, int nvec, int precision, int warmup, int iters, double *ms)
{
    if (!c || nvec <= 0 || iters <= 0 || warmup < 0 || !ms) return PISCES_ERR_ARG;
    if (precision < PISCES_FP64 || precision > PISCES_BF16X2) return PISCES_ERR_ARG;
    if (!c->has_v) return PISCES_ERR_STATE;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const long long n = c->ngp * (long long)nvec;
        const size_t es = precision == PISCES_FP64 ? sizeof(double) : sizeof(float);
        dbuf X(es * n), Y(es * n);
        fill_uniform(X.p, n, precision == PISCES_FP64, 12345u, c->stream);
        for (int w = 0; w < warmup; ++w) c->hx(X.p, Y.p, c->ngp, nvec, precision);
        hipEvent_t e0, e1;
        PC_HIP(hipEventCreate(&e0));
        PC_HIP(hipEventCreate(&e1));
        PC_HIP(hipEventRecord(e0, c->stream));
        for (int w = 0; w < iters; ++w) c->hx(X.p, Y.p, c->ngp, nvec, precision);
        PC_HIP(hipEventRecord(e1, c->stream));
        PC_HIP(hipEventSynchronize(e1));
        float t = 0;
        PC_HIP(hipEventElapsedTime(&t, e0, e1));
        static_cast<void>(hipEventDestroy(e0));
        static_cast<void>(hipEventDestroy(e1));
        *ms = (double)t / iters;
        return 0;
    });
}

int pisces_time_precond(pisces_ctx *c, int nvec, int precision, int warmup, int iters, double *ms)
{
    if (!c || nvec <= 0 || iters <= 0 || warmup < 0 || !ms) return PISCES_ERR_ARG;
    if (precision != PISCES_FP64 && precision != PISCES_BF16X2) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const bool f64 = precision == PISCES_FP64;
        const long long n = c->ngp * (long long)nvec;
        const size_t es = f64 ? sizeof(double) : sizeof(float);
        dbuf R(es * n), Tb(es * n), W(sizeof(uint16_t) * n), tasks(sizeof(FdmTask) * nvec);
        fill_uniform(R.p, n, f64, 4321u, c->stream);
        std::vector<FdmTask> h(nvec);
        for (int v = 0; v < nvec; ++v) {
            FdmTask &f = h[v];
            for (int a = 0; a < 3; ++a) f.ax[a] = c->ax[a].m;
            f.src = R.as<unsigned char>() + es * c->ngp * v;
            f.dst = Tb.as<unsigned char>() + es * c->ngp * v;
            f.src32 = f.dst32 = f64 ? 0 : 1;
            f.work = W.as<uint16_t>() + c->ngp * v;
            f.shift = -0.05;
            f.scale = -1.0;
            f.floor = 1e-3;
        }
        PC_HIP(hipMemcpyAsync(tasks.p, h.data(), sizeof(FdmTask) * nvec, hipMemcpyHostToDevice,
                              c->stream));
        for (int w = 0; w < warmup; ++w) precond_launch(c, h, tasks.as<FdmTask>(), precision);
        hipEvent_t e0, e1;
        PC_HIP(hipEventCreate(&e0));
        PC_HIP(hipEventCreate(&e1));
        PC_HIP(hipEventRecord(e0, c->stream));
        for (int w = 0; w < iters; ++w) precond_launch(c, h, tasks.as<FdmTask>(), precision);
        PC_HIP(hipEventRecord(e1, c->stream));
        PC_HIP(hipEventSynchronize(e1));
        float t = 0;
        PC_HIP(hipEventElapsedTime(&t, e0, e1));
        static_cast<void>(hipEventDestroy(e0));
        static_cast<void>(hipEventDestroy(e1));
        *ms = (double)t / iters;
        return 0;
    });
}

int pisces_copy_probe(int device, long long bytes, int reps, double *ms)
{
    if (bytes <= 0 || bytes % 16 || reps <= 0 || !ms) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(device));
        hipStream_t s;
        PC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        dbuf a(bytes), b(bytes);
        fill_uniform(a.p, bytes / 8, true, 7u, s);
        copy_probe(a.p, b.p, bytes, s);  // warm-up
        hipEvent_t e0, e1;
        PC_HIP(hipEventCreate(&e0));
        PC_HIP(hipEventCreate(&e1));
        PC_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) copy_probe(a.p, b.p, bytes, s);
        PC_HIP(hipEventRecord(e1, s));
        PC_HIP(hipEventSynchronize(e1));
        float t = 0;
        PC_HIP(hipEventElapsedTime(&t, e0, e1));
        static_cast<void>(hipEventDestroy(e0));
        static_cast<void>(hipEventDestroy(e1));
        PC_HIP(hipStreamDestroy(s));
        *ms = (double)t / reps;
        return 0;
    });
}

int pisces_stream_probe(int device, long long bytes, int reps, int nontemporal, double *ms)
{
    if (bytes <= 0 || bytes % 16 || reps <= 0 || !ms) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(device));
        hipStream_t s;
        PC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        dbuf a(bytes), sink(sizeof(unsigned));
        fill_uniform(a.p, bytes / 8, true, 7u, s);
        unsigned *k = static_cast<unsigned *>(sink.p);
        stream_probe(a.p, bytes, nontemporal != 0, k, s);  // warm-up
        hipEvent_t e0, e1;
        PC_HIP(hipEventCreate(&e0));
        PC_HIP(hipEventCreate(&e1));
        PC_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) stream_probe(a.p, bytes, nontemporal != 0, k, s);
        PC_HIP(hipEventRecord(e1, s));
        PC_HIP(hipEventSynchronize(e1));
        float t = 0;
        PC_HIP(hipEventElapsedTime(&t, e0, e1));
        static_cast<void>(hipEventDestroy(e0));
        static_cast<void>(hipEventDestroy(e1));
        PC_HIP(hipStreamDestroy(s));
        *ms = (double)t / reps;
        return 0;
    });
}

int pisces_kinetic_precond(pisces_ctx *c, int nvec, const double *r, double shift, double scale,
                           double floor, int precision, double *t)
{
    if (!c || nvec <= 0 || !r || !t || !(floor > 0)) return PISCES_ERR_ARG;
    if (precision != PISCES_FP64 && precision != PISCES_BF16X2) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        const size_t bytes = sizeof(double) * c->ngp * (size_t)nvec;
        dbuf R(bytes), Tb(bytes), W(sizeof(uint16_t) * c->ngp * (size_t)nvec);
        dbuf tasks(sizeof(FdmTask) * nvec);
        std::vector<FdmTask> h(nvec);
        for (int v = 0; v < nvec; ++v) {
            FdmTask &f = h[v];
            for (int a = 0; a < 3; ++a) f.ax[a] = c->ax[a].m;
            f.src = R.as<double>() + c->ngp * v;
            f.src32 = 0;
            f.dst32 = 0;
            f.dst = Tb.as<double>() + c->ngp * v;
            f.work = W.as<uint16_t>() + c->ngp * v;
            f.shift = shift;
            f.scale = scale;
            f.floor = floor;
        }
        PC_HIP(hipMemcpyAsync(R.p, r, bytes, hipMemcpyHostToDevice, c->stream));
        PC_HIP(hipMemcpyAsync(tasks.p, h.data(), sizeof(FdmTask) * nvec, hipMemcpyHostToDevice,
                              c->stream));
        precond_launch(c, h, tasks.as<FdmTask>(), precision);
        PC_HIP(hipMemcpyAsync(t, Tb.p, bytes, hipMemcpyDeviceToHost, c->stream));
        c->sync();
        return 0;
    });
}

int pisces_diag(pisces_ctx *c, double *diag)
{
    if (!c || !diag) return PISCES_ERR_ARG;
    if (!c->has_v) return PISCES_ERR_STATE;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        c->wdiag.ensure(sizeof(double) * c->ngp);
        diag_h(c->dV64.as<double>(), c->dT64[0].as<double>(), c->dT64[1].as<double>(),
               c->dT64[2].as<double>(), c->n[0], c->n[1], c->n[2], c->wdiag.as<double>(), c->stream);
        PC_HIP(hipMemcpyAsync(diag, c->wdiag.p, sizeof(double) * c->ngp, hipMemcpyDeviceToHost, c->stream));
        c->sync();
        return 0;
    });
}

int pisces_piab(pisces_ctx *c, double *wf)
{
    if (!c || !wf) return PISCES_ERR_ARG;
    // DVR::ParticleInAnDBoxWf (src/DVR.cpp:1725-1805): x0 = first grid point,
    // L = last - first; the lowest state of the central half box.
    const double pi = 3.14159265358979;
    double x0[3], Lb[3];
    for (int a = 0; a < 3; ++a) {
        x0[a] = c->x[a][0];
        Lb[a] = c->x[a][c->n[a] - 1] - x0[a];
    }
    for (int k = 0; k < c->n[2]; ++k)
        for (int j = 0; j < c->n[1]; ++j)
            for (int i = 0; i < c->n[0]; ++i) {
                const int ii[3] = {i, j, k};
                double w = 1.0;
                for (int a = 0; a < 3; ++a) {
                    const double q = c->x[a][ii[a]];
                    if (q < x0[a] / 2 || q > -x0[a] / 2) w = 0.0;
                    else w *= std::sin(pi / Lb[a] * 2.0 * (q - x0[a] / 2.0));
                }
                wf[i + (long long)c->n[0] * (j + (long long)c->n[1] * k)] = w;
            }
    return PISCES_OK;
}

int pisces_solve(pisces_ctx *c, int nroots, int maxsub, int maxmacro, int ptol, int corrflag,
                 int precision, double *vecs, double *evals, pisces_stats *st)
{
    if (!c || !vecs || !evals) return PISCES_ERR_ARG;
    return guarded([&] {
        PC_HIP(hipSetDevice(c->device));
        if (precision == PISCES_FP64)
            return solve_fp64(c, nroots, maxsub, maxmacro, ptol, corrflag, vecs, evals, st);
        PC_REQUIRE(precision == PISCES_FP32 || precision == PISCES_BF16X2, PISCES_ERR_ARG,
                   "unknown precision");
        return solve_mixed(c, nroots, maxsub, maxmacro, ptol, corrflag, precision, vecs, evals, st);
    });
}

unsigned pisces_debug_hx_spin_timeouts(void)
{
    unsigned v = 0;
    guarded([&] {
        v = hx_fused_spin_timeouts();
        return 0;
    });
    return v;
}

// test hook: the polarization-matrix inverse behind PolType 3 on the host,
// with its method chosen (no device work)
int pisces_debug_sym_inverse(int n, const double *a, int method, double *inv)
{
    return guarded([&] {
        PC_REQUIRE(n > 0 && a && inv && method >= 0 && method <= 2, PISCES_ERR_ARG,
                   "pisces_debug_sym_inverse: bad arguments");
        std::vector<double> A(a, a + (size_t)n * n);
        PC_REQUIRE(sym_inverse(n, A, n, inv, method), PISCES_ERR_NUMERIC,
                   "matrix singular (or not positive definite for method 1)");
        return PISCES_OK;
    });
}

// test hook: the H*X kernels of one context (0 VALU, 1 matrix cores, 2 matrix
// cores with every form in its multi-launch shape, 3 matrix cores with the
// split-BF16 H*X and preconditioner in their one-launch XCD-ordered form)
int pisces_debug_set_hx_variant(pisces_ctx *c, int v)
{
    return guarded([&] {
        PC_REQUIRE(c && v >= 0 && v <= 3, PISCES_ERR_ARG, "pisces_debug_set_hx_variant: variant 0-3");
        PC_HIP(hipSetDevice(c->device));
        PC_REQUIRE(v != 3 || one_launch_available(), PISCES_ERR_UNSUPPORTED,
                   "the one-launch form needs a device with eight XCDs");
        c->hx_variant = v;
        return PISCES_OK;
    });
}

}  // extern "C"
