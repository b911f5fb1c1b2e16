// potential.hip -- electron-water potential on the grid, and the FP64 induced-
// dipole solve of the water model, for gfx950.
//
// Grid build: DVR::ComputePotential with sampling 1 (src/DVR.cpp:661-859) calls
// Potential::Evaluate (src/Potential.cpp:128-182) -> EvaluateDPPGTOP (:844-956)
// at every grid point.  Here one thread evaluates one grid point; the site,
// charge, dipole, repulsive-core and polarizable-site tables of the cluster are
// staged once per workgroup in LDS (at most a few KB) and every thread walks
// them.  The per-site distance table of the reference (Rx, Ry, Rz, R2, R,
// Rminus3; src/Potential.cpp:136-146) is recomputed per use instead of stored,
// with the same expressions, so the FP64 result agrees with the reference to
// rounding.  The work is N_grid x N_site FP64 arithmetic with one reciprocal
// square root per site (1/R to ~1 ulp; R, R^-3, R^-4 follow by products, exact
// divisions only inside the damping radii) and one exp per repulsive core:
// compute-bound on the FP64 VALU.  The per-cluster tables are read as scalar
// loads through the constant address space.
#include "common.hpp"
#include "gmem.hpp"

namespace pisces {

namespace {

constexpr int kPotThreads = 256;
constexpr int kMaxSites = 1024;  // LDS budget check (4 sites per water)

struct SiteLds {
    double x, y, z;
};

__device__ inline double reff(double r, double r0)
{
    // effective-distance damping, src/Potential.cpp:1191-1196
    if (r < r0) {
        const double t = r / r0;
        return r0 * (0.5 + t * t * t * (1.0 - 0.5 * t));
    }
    return r;
}

// |r| from r^2 and 1/|r| (0 at r = 0, where 1/|r| is infinite)
__device__ __forceinline__ double norm_of(double r2, double ri) { return r2 > 0.0 ? r2 * ri : 0.0; }

__device__ void potential_point(const PotParams &P, const PotDev &D, const SiteLds *__restrict__ S,
                                int nx, int ny, int nz, const double *__restrict__ gx,
                                const double *__restrict__ gy, const double *__restrict__ gz,
                                double *__restrict__ v, float *__restrict__ v32, long long g)
{
    const int i = (int)(g % nx);
    const int j = (int)((g / nx) % ny);
    const int k = (int)(g / ((long long)nx * ny));
    const double r0 = gx[i], r1 = gy[j], r2 = gz[k];
    const cptr<double> q = cst(D.q), dm = cst(D.d), rp = cst(D.rep), ionx = cst(D.ion_xyz),
                       ionq = cst(D.ion_q);
    const cptr<int> iq = cst(D.iq), id = cst(D.id), irep = cst(D.irep), ipps = cst(D.ipps);

    // point charges, src/Potential.cpp:1168-1210
    double vpc = 0.0;
    for (int c = 0; c < D.nq; ++c) {
        const SiteLds s = S[iq[c]];
        const double rx = r0 - s.x, ry = r1 - s.y, rz = r2 - s.z;
        const double R2 = rx * rx + ry * ry + rz * rz;
        const double ri = rsq64(R2);
        if (P.damp_type == 1) {
            vpc += -q[c] * ri * (1.0 - exp(-P.qdamp * R2));
        } else {
            const double R = norm_of(R2, ri);
            vpc += (R < P.qdamp) ? -q[c] / reff(R, P.qdamp) : -q[c] * ri;
        }
    }
    // point dipoles, src/Potential.cpp:1222-1270
    double vind = 0.0;
    for (int c = 0; c < D.nd; ++c) {
        const SiteLds s = S[id[c]];
        const double rx = r0 - s.x, ry = r1 - s.y, rz = r2 - s.z;
        const double R2 = rx * rx + ry * ry + rz * rz;
        const double ri = rsq64(R2);
        const double dot = rx * dm[3 * c] + ry * dm[3 * c + 1] + rz * dm[3 * c + 2];
        double rm3 = ri * ri * ri;
        if (P.damp_type == 1) {
            const double dd = 1.0 - exp(-P.ddamp * R2);
            vind += -dot * rm3 * dd * dd;
        } else {
            const double R = norm_of(R2, ri);
            if (R < P.ddamp) {
                const double re = reff(R, P.ddamp);
                rm3 = 1.0 / (re * re * re);
            }
            vind += -dot * rm3;
        }
    }
    // repulsive cores, src/Potential.cpp:1282-1315
    double vrep = 0.0;
    if (P.rep_type > 0) {
        for (int c = 0; c < D.nrep; ++c) {
            const SiteLds s = S[irep[c]];
            const double rx = r0 - s.x, ry = r1 - s.y, rz = r2 - s.z;
            const double R2 = rx * rx + ry * ry + rz * rz;
            if (P.rep_type == 1) vrep += rp[2 * c + 1] * exp(-rp[2 * c] * R2);
            else vrep += rp[2 * c + 1] * exp(-rp[2 * c] * norm_of(R2, rsq64(R2)));
        }
    }
    // polarization of non-interacting sites, src/Potential.cpp:1331-1374, :875-887
    double vpol = 0.0;
    if (P.pol_type == 1) {
        double sp = 0.0;
        for (int c = 0; c < D.npps; ++c) {
            const SiteLds s = S[ipps[c]];
            const double rx = r0 - s.x, ry = r1 - s.y, rz = r2 - s.z;
            const double rsq = rx * rx + ry * ry + rz * rz;
            const double ri = rsq64(rsq), ri2 = ri * ri;
            if (P.damp_type == 1) {
                const double dd = 1.0 - exp(-P.pdamp * rsq);
                sp += P.alpha * dd * dd * (ri2 * ri2);
            } else {
                const double R = norm_of(rsq, ri);
                if (R < P.pdamp) {
                    const double re = reff(R, P.pdamp);
                    const double re2 = re * re;
                    sp += P.alpha * rsq / (re2 * re2 * re2);
                } else {
                    sp += P.alpha * (ri2 * ri2);
                }
            }
        }
        vpol = P.adiabatic ? P.eps - sqrt(P.eps * (P.eps + sp)) : -0.5 * sp;
    }
    // ions: erf-damped charges (Bloomfield form, src/NaCl.cpp:300-307)
    double vion = 0.0;
    for (int c = 0; c < D.nion; ++c) {
        const double rx = r0 - ionx[3 * c], ry = r1 - ionx[3 * c + 1], rz = r2 - ionx[3 * c + 2];
        const double R2 = rx * rx + ry * ry + rz * rz;
        const double ri = rsq64(R2);
        vion += -ionq[c] * erf(norm_of(R2, ri) / P.ion_r) * ri;
    }
    const double vt = vpc + vind + vrep + vpol + vion;
    if (v) v[g] = vt;
    if (v32) v32[g] = (float)vt;
}

__global__ void __launch_bounds__(kPotThreads)
potential_grid_kernel(PotParams P, PotDev D, int nx, int ny, int nz, const double *__restrict__ gx,
                      const double *__restrict__ gy, const double *__restrict__ gz,
                      double *__restrict__ v, float *__restrict__ v32)
{
    __shared__ SiteLds S[kMaxSites];
    for (int s = threadIdx.x; s < D.nsites; s += blockDim.x)
        S[s] = SiteLds{D.sites[3 * s], D.sites[3 * s + 1], D.sites[3 * s + 2]};
    __syncthreads();
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)nx * ny * nz) return;
    potential_point(P, D, S, nx, ny, nz, gx, gy, gz, v, v32, g);
}

// many clusters in one launch: blockIdx.y selects the task
__global__ void __launch_bounds__(kPotThreads)
potential_grid_batched_kernel(PotParams P, const PotTask *__restrict__ tasks)
{
    const PotTask &t = tasks[blockIdx.y];
    const long long ng = (long long)t.nx * t.ny * t.nz;
    if ((long long)blockIdx.x * blockDim.x >= ng) return;
    __shared__ SiteLds S[kMaxSites];
    for (int s = threadIdx.x; s < t.D.nsites; s += blockDim.x)
        S[s] = SiteLds{t.D.sites[3 * s], t.D.sites[3 * s + 1], t.D.sites[3 * s + 2]};
    __syncthreads();
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    potential_point(P, t.D, S, t.nx, t.ny, t.nz, t.gx, t.gy, t.gz, t.v, t.v32, g);
}

// --------------------------------------------------------------------------
// Induced dipoles of the rigid DPP waters, FP64, one workgroup.
// WaterCluster::CalcInducedDipoles (src/DPP.cpp:760-1065): field of the
// permanent charges of the other waters at the polarizable sites (O, H1, H2),
// then the 0.5-mixed fixed-point iteration mu <- 0.5 mu + 0.5 alpha (E_pc +
// T mu) until the mean change drops below 1e-12.  The damped dipole tensors
// T_ij (src/DPP.cpp:2377-2410) do not change during the iteration and are
// computed once into d_tensor (6 doubles per ordered pair).
constexpr int kScfThreads = 256;
constexpr int kScfSharedSites = kScfThreads / 2;  // clusters solved with the dipoles in LDS
constexpr double kA2B = 1.0 / 0.52917720859;

__device__ inline void site_params(int s, double &q, double &sig, double &alpha)
{
    // src/Water.cpp:233-236 (chargeselector 2) with src/ttm_data.h:14-15, 51-52
    const double a3 = kA2B * kA2B * kA2B;
    const double polO = 1.22 * a3, polH = 0.28 * a3;
    const double qH = 0.5742185854046506;
    if (s == 0) { q = 0.0; sig = 0.0; alpha = polO; }
    else if (s < 3) { q = qH; sig = polH; alpha = polH; }
    else { q = -2.0 * qH; sig = polO; alpha = 0.0; }
}

__device__ inline double damp_cd(double r, double ai, double aj, double a)
{
    const double A = pow(ai * aj, 1.0 / 6.0);
    const double x = a * pow(r / A, 3);
    return 1.0 - exp(-x);
}
__device__ inline double damp_dd(double r, double ai, double aj, double a)
{
    const double A = pow(ai * aj, 1.0 / 6.0);
    const double x = a * pow(r / A, 3);
    return 1.0 - (1 + x) * exp(-x);
}

__device__ void water_scf_body(int nw, const double *__restrict__ sites, double *__restrict__ mu_out,
                               double *__restrict__ epc_out, double *__restrict__ tens,
                               int *__restrict__ info)
{
    const int ns = 3 * nw;
    const double dampCD = 0.21, dampDD = 0.3;  // src/Water.cpp:243
    __shared__ double red[kScfThreads];
    __shared__ int s_done;
    // tensors of every ordered pair (i at row, dipole j)
    for (long long p = threadIdx.x; p < (long long)ns * ns; p += blockDim.x) {
        const int i = (int)(p / ns), j = (int)(p % ns);
        double *t = tens + 6 * p;
        if (i == j) { for (int c = 0; c < 6; ++c) t[c] = 0.0; continue; }
        double q, sg, ai, aj;
        site_params(i % 3, q, sg, ai);
        site_params(j % 3, q, sg, aj);
        const double *xi = sites + 12 * (i / 3) + 3 * (i % 3);
        const double *xj = sites + 12 * (j / 3) + 3 * (j % 3);
        const double rx = xi[0] - xj[0], ry = xi[1] - xj[1], rz = xi[2] - xj[2];
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double r5 = pow(r2, -2.5);
        const double d1 = damp_dd(sqrt(r2), ai, aj, dampDD);
        const double d2 = damp_cd(sqrt(r2), ai, aj, dampDD);
        t[0] = (d1 * 3. * rx * rx - d2 * r2) * r5;
        t[1] = (d1 * 3. * ry * ry - d2 * r2) * r5;
        t[2] = (d1 * 3. * rz * rz - d2 * r2) * r5;
        t[3] = d1 * 3. * rx * ry * r5;
        t[4] = d1 * 3. * rx * rz * r5;
        t[5] = d1 * 3. * ry * rz * r5;
    }
    // permanent-charge field (src/DPP.cpp:830-865); each thread owns sites i
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        double q, sg, ai;
        site_params(i % 3, q, sg, ai);
        const double *xi = sites + 12 * (i / 3) + 3 * (i % 3);
        double f0 = 0, f1 = 0, f2 = 0;
        for (int w = 0; w < nw; ++w)
            for (int s = 0; s < 4; ++s) {
                double qq, sig, a;
                site_params(s, qq, sig, a);
                if (qq == 0 || i / 3 == w) continue;
                const double *pc = sites + 12 * w + 3 * s;
                const double rx = xi[0] - pc[0], ry = xi[1] - pc[1], rz = xi[2] - pc[2];
                const double t = qq * pow(rx * rx + ry * ry + rz * rz, -1.5);
                const double dist = sqrt(rx * rx + ry * ry + rz * rz);
                const double dmp = damp_cd(dist, ai, sig, dampCD);
                f0 += t * rx * dmp;
                f1 += t * ry * dmp;
                f2 += t * rz * dmp;
            }
        epc_out[3 * i] = f0;
        epc_out[3 * i + 1] = f1;
        epc_out[3 * i + 2] = f2;
        mu_out[3 * i] = 0.5 * ai * f0;
        mu_out[3 * i + 1] = 0.5 * ai * f1;
        mu_out[3 * i + 2] = 0.5 * ai * f2;
    }
    __syncthreads();
    const int maxit = 5 * (ns + 10);
    int it = 0, conv = 0;
    if (ns <= kScfSharedSites) {
        // ns <= 128: tps = floor(256 / ns) >= 2 threads per site, each summing
        // the partners j = part, part + tps, ...; the dipoles live in LDS and
        // the tps partial fields of a site are added there
        __shared__ double smu[3 * kScfSharedSites];
        __shared__ double pe[kScfThreads][3];
        const int tps = kScfThreads / ns;
        const int i = threadIdx.x / tps, part = threadIdx.x - i * tps;
        const bool own = i < ns;
        for (int idx = threadIdx.x; idx < 3 * ns; idx += blockDim.x) smu[idx] = mu_out[idx];
        double q, sg, ai = 0.0;
        if (own) site_params(i % 3, q, sg, ai);
        __syncthreads();
        for (it = 0; it < maxit; ++it) {
            double e0 = 0, e1 = 0, e2 = 0;
            if (own) {
                const double *t = tens + 6LL * i * ns;
                for (int j = part; j < ns; j += tps) {
                    if (j == i) continue;
                    const double m0 = smu[3 * j], m1 = smu[3 * j + 1], m2 = smu[3 * j + 2];
                    const double *tj = t + 6 * j;
                    e0 += tj[0] * m0 + tj[3] * m1 + tj[4] * m2;
                    e1 += tj[3] * m0 + tj[1] * m1 + tj[5] * m2;
                    e2 += tj[4] * m0 + tj[5] * m1 + tj[2] * m2;
                }
            }
            pe[threadIdx.x][0] = e0;
            pe[threadIdx.x][1] = e1;
            pe[threadIdx.x][2] = e2;
            __syncthreads();  // partial fields written, old dipoles read
            double part_d = 0.0, n0 = 0, n1 = 0, n2 = 0;
            if (own && part == 0) {
                for (int p2 = 1; p2 < tps; ++p2) {
                    e0 += pe[threadIdx.x + p2][0];
                    e1 += pe[threadIdx.x + p2][1];
                    e2 += pe[threadIdx.x + p2][2];
                }
                const double o0 = smu[3 * i], o1 = smu[3 * i + 1], o2 = smu[3 * i + 2];
                n0 = 0.5 * o0 + 0.5 * ai * (epc_out[3 * i] + e0);
                n1 = 0.5 * o1 + 0.5 * ai * (epc_out[3 * i + 1] + e1);
                n2 = 0.5 * o2 + 0.5 * ai * (epc_out[3 * i + 2] + e2);
                const double d0 = o0 - n0, d1 = o1 - n1, d2 = o2 - n2;
                part_d = d0 * d0 + d1 * d1 + d2 * d2;
            }
            red[threadIdx.x] = part_d;
            __syncthreads();  // every owner has read its old dipole
            if (own && part == 0) {
                smu[3 * i] = n0;
                smu[3 * i + 1] = n1;
                smu[3 * i + 2] = n2;
            }
            for (int w = blockDim.x / 2; w > 0; w >>= 1) {
                __syncthreads();
                if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            }
            if (threadIdx.x == 0) s_done = (sqrt(red[0]) / 3 / ns < 1e-12) ? 1 : 0;
            __syncthreads();
            if (s_done) { conv = 1; break; }
        }
        for (int idx = threadIdx.x; idx < 3 * ns; idx += blockDim.x) mu_out[idx] = smu[idx];
        if (threadIdx.x == 0) {
            info[0] = conv;
            info[1] = conv ? it + 1 : it;
        }
        return;
    }
    // larger clusters: mu_out holds the current dipoles; a private next value
    // per owned site is kept in registers (at most ceil(ns/256) sites per thread)
    constexpr int kPer = 4;
    for (it = 0; it < maxit; ++it) {
        double nxt[kPer][3];
        double part = 0.0;
        int slot = 0;
        for (int i = threadIdx.x; i < ns; i += blockDim.x, ++slot) {
            double q, sg, ai;
            site_params(i % 3, q, sg, ai);
            double e0 = 0, e1 = 0, e2 = 0;
            const double *t = tens + 6LL * i * ns;
            for (int j = 0; j < ns; ++j) {
                if (j == i) continue;
                const double m0 = mu_out[3 * j], m1 = mu_out[3 * j + 1], m2 = mu_out[3 * j + 2];
                const double *tj = t + 6 * j;
                e0 += tj[0] * m0 + tj[3] * m1 + tj[4] * m2;
                e1 += tj[3] * m0 + tj[1] * m1 + tj[5] * m2;
                e2 += tj[4] * m0 + tj[5] * m1 + tj[2] * m2;
            }
            const double o0 = mu_out[3 * i], o1 = mu_out[3 * i + 1], o2 = mu_out[3 * i + 2];
            nxt[slot][0] = 0.5 * o0 + 0.5 * ai * (epc_out[3 * i] + e0);
            nxt[slot][1] = 0.5 * o1 + 0.5 * ai * (epc_out[3 * i + 1] + e1);
            nxt[slot][2] = 0.5 * o2 + 0.5 * ai * (epc_out[3 * i + 2] + e2);
            const double d0 = o0 - nxt[slot][0], d1 = o1 - nxt[slot][1], d2 = o2 - nxt[slot][2];
            part += d0 * d0 + d1 * d1 + d2 * d2;
        }
        red[threadIdx.x] = part;
        __syncthreads();  // every thread has read the old dipoles
        slot = 0;
        for (int i = threadIdx.x; i < ns; i += blockDim.x, ++slot) {
            mu_out[3 * i] = nxt[slot][0];
            mu_out[3 * i + 1] = nxt[slot][1];
            mu_out[3 * i + 2] = nxt[slot][2];
        }
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            __syncthreads();
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        }
        if (threadIdx.x == 0) s_done = (sqrt(red[0]) / 3 / ns < 1e-12) ? 1 : 0;
        __syncthreads();
        if (s_done) { conv = 1; break; }
    }
    if (threadIdx.x == 0) {
        info[0] = conv;
        info[1] = conv ? it + 1 : it;
    }
}

__global__ void __launch_bounds__(kScfThreads)
water_scf_kernel(int nw, const double *__restrict__ sites, double *__restrict__ mu_out,
                 double *__restrict__ epc_out, double *__restrict__ tens, int *__restrict__ info)
{
    water_scf_body(nw, sites, mu_out, epc_out, tens, info);
}

// one cluster per workgroup
__global__ void __launch_bounds__(kScfThreads) water_scf_batched_kernel(const ScfTask *__restrict__ tasks)
{
    const ScfTask &t = tasks[blockIdx.x];
    water_scf_body(t.nw, t.sites, t.mu, t.epc, t.tens, t.info);
}

}  // namespace

void potential_grid_launch(const PotParams &P, const PotDev &D, int nx, int ny, int nz,
                           const double *d_x, const double *d_y, const double *d_z,
                           double *d_v, float *d_v32, hipStream_t s)
{
    PC_REQUIRE(D.nsites <= kMaxSites, PISCES_ERR_UNSUPPORTED, "too many sites for the LDS table");
    const long long ng = (long long)nx * ny * nz;
    const unsigned blocks = (unsigned)((ng + kPotThreads - 1) / kPotThreads);
    hipLaunchKernelGGL(potential_grid_kernel, dim3(blocks), dim3(kPotThreads), 0, s, P, D, nx,
                       ny, nz, d_x, d_y, d_z, d_v, d_v32);
    PC_HIP(hipGetLastError());
}

void potential_grid_batched_launch(const PotParams &P, const PotTask *d_tasks, int ntasks,
                                   long long max_ng, hipStream_t s)
{
    if (ntasks <= 0) return;
    const unsigned blocks = (unsigned)((max_ng + kPotThreads - 1) / kPotThreads);
    hipLaunchKernelGGL(potential_grid_batched_kernel, dim3(blocks, ntasks), dim3(kPotThreads), 0, s,
                       P, d_tasks);
    PC_HIP(hipGetLastError());
}

void water_scf_batched_launch(const ScfTask *d_tasks, int ntasks, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(water_scf_batched_kernel, dim3(ntasks), dim3(kScfThreads), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

int water_scf_max_waters() { return 4 * kScfThreads / 3; }
int potential_max_sites() { return kMaxSites; }

void water_scf_launch(int nw, const double *d_sites, double *d_mu, double *d_epc,
                      double *d_tensor, int *d_info, hipStream_t s)
{
    PC_REQUIRE(nw > 0 && 3 * nw <= 4 * kScfThreads, PISCES_ERR_UNSUPPORTED,
               "too many waters for the dipole solve");
    hipLaunchKernelGGL(water_scf_kernel, dim3(1), dim3(kScfThreads), 0, s, nw, d_sites, d_mu,
                       d_epc, d_tensor, d_info);
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
