// grad.hip -- the electron side of the analytic gradient on the grid.
//
// DVR::ComputeGradient (src/DVR.cpp:1099-1176) sums, over the grid points r_g
// with weights w_g = psi_g^2, the derivative of the potential with respect to
// every site position (Potential::EvaluateDPPGTOPGradient,
// src/Gradients.cpp:79-238; Potential 3, DampType 2, Polarization 3):
//   * the point charges, -q (dReff/dR) / Reff^2 along R / |R| (R = r - site);
//   * the Slater repulsive cores, -c zeta e^{-zeta R} along R / |R|;
//   * the electron's field E_i = g(R_i) R_i at the polarizable sites, whose
//     derivative dE_i/dR_i is dotted with the induced dipole mu_i of
//     mu = InvA (E + Epc) at that point (DerivElecField);
// and accumulates tmu = sum w mu and mcm = sum w mu mu^T (MuCrossMu), which the
// caller's water-side derivatives turn into the rest of the gradient
// (FinalGradient).  Here:
//   grad_points_kernel: 16-point tiles like the PolType 3 potential kernel --
//     E rows in LDS, mu = E InvA on the FP64 matrix cores, InvA streamed from
//     L2 -- with the site terms and the dipole terms summed per workgroup over
//     its tiles and written as one row of partial sums;
//   grad_s_kernel: S = sum w (E + Epc)(E + Epc)^T over 16 x 16 blocks of the
//     upper triangle on the matrix cores, E rows built once per tile in LDS,
//     split over tile ranges and block groups;
//   colsum_kernel: the partial rows summed in a fixed order;
//   matmul_kernel: mcm = InvA S InvA (two small products).
#include "common.hpp"
#include "context.hpp"
#include "gmem.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace pisces {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kGradNT = 512;   // threads per workgroup of the point kernel
constexpr int kGradCBW = 5;    // column blocks of 16 per wave: 8 waves x 5 x 16 >= 576
constexpr int kLdsMax = 160 * 1024 - 4096;

struct GradDev {
    int nsites, np, n3p, pad;
    double qdamp, pdamp;
    const double *site4;  // per site: x, y, z, charge
    const double *srep;   // per site: zeta1, c1, zeta2, c2 (c = 0: unused)
    const double *rpol;   // 3 np: polarizable-site positions
    const double *epc;    // n3p: permanent-charge field (0 past 3 np)
    const double *inva;   // n3p x n3p, row-major, symmetric, zero-padded
};

// effective distance, src/Potential.cpp:1191-1196, and dReff/dR
__device__ inline double reff(double r, double r0, double &dre)
{
    if (r < r0) {
        const double t = r / r0;
        dre = 3.0 * t * t - 2.0 * t * t * t;
        return r0 * (0.5 + t * t * t * (1.0 - 0.5 * t));
    }
    dre = 1.0;
    return r;
}

// g(R) of the electron's field at a polarizable site and dg/dR
// (src/Gradients.cpp:131-142, 187-198); R2 > 0
__device__ inline void field_g(double R, double pdamp, double &g, double &dg)
{
    double dre;
    const double re = reff(R, pdamp, dre);
    const double re2 = re * re;
    g = 1.0 / (re2 * re);
    dg = dre * (-3.0 / (re2 * re2));
}

__device__ inline void point_of(long long p, int nx, int ny, const double *gx, const double *gy,
                                const double *gz, double r[3])
{
    const int ix = (int)(p % nx), iy = (int)((p / nx) % ny);
    const int iz = (int)(p / ((long long)nx * ny));
    r[0] = gx[ix];
    r[1] = gy[iy];
    r[2] = gz[iz];
}

// Partial row of a workgroup: [3 nsites site terms | n3p tmu | 3 n3p dipole terms]
__global__ void __launch_bounds__(kGradNT)
grad_points_kernel(GradDev D, int nx, int ny, int nz, const double *__restrict__ gx,
                   const double *__restrict__ gy, const double *__restrict__ gz,
                   const double *__restrict__ psi, double *__restrict__ part)
{
    constexpr int NW = kGradNT / 64;
    extern __shared__ double Es[];  // 16 rows of n3p (+1)
    __shared__ double pt[16][4];    // point x, y, z and its weight psi^2
    const int n3p = D.n3p, ldE = n3p + 1, np = D.np, n3 = 3 * np, ns = D.nsites;
    const long long ng = (long long)nx * ny * nz;
    const long long ntiles = (ng + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ncb = n3p / 16;
    const gptr<const double> A = glb(D.inva), rp = glb(D.rpol), ep = glb(D.epc);
    const gptr<const double> s4 = glb(D.site4), sr = glb(D.srep);

    // site terms: thread t owns sites t, t + kGradNT
    double gs[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double tmu[kGradCBW], gp[kGradCBW][3];
#pragma unroll
    for (int u = 0; u < kGradCBW; ++u) {
        tmu[u] = 0.0;
        gp[u][0] = gp[u][1] = gp[u][2] = 0.0;
    }
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const long long g0 = 16 * t;
        if (threadIdx.x < 16) {
            const long long g = g0 + threadIdx.x;
            double r[3] = {0.0, 0.0, 0.0};
            double w = 0.0;
            if (g < ng) {
                point_of(g, nx, ny, gx, gy, gz, r);
                w = psi[g] * psi[g];
            }
            pt[threadIdx.x][0] = r[0];
            pt[threadIdx.x][1] = r[1];
            pt[threadIdx.x][2] = r[2];
            pt[threadIdx.x][3] = w;
        }
        __syncthreads();
        // E rows (electron field + permanent-charge field); rows past the grid
        // carry weight 0
        for (int idx = threadIdx.x; idx < 16 * np; idx += kGradNT) {
            const int p = idx / np, i = idx - p * np;
            const double rx = pt[p][0] - rp[3 * i], ry = pt[p][1] - rp[3 * i + 1],
                         rz = pt[p][2] - rp[3 * i + 2];
            const double R2 = rx * rx + ry * ry + rz * rz;
            double g = 0.0, dg;
            if (R2 > 0.0) field_g(sqrt(R2), D.pdamp, g, dg);
            double *row = Es + p * ldE + 3 * i;
            row[0] = g * rx + ep[3 * i];
            row[1] = g * ry + ep[3 * i + 1];
            row[2] = g * rz + ep[3 * i + 2];
        }
        for (int idx = threadIdx.x; idx < 16 * (n3p - n3); idx += kGradNT) {
            const int p = idx / (n3p - n3);
            Es[p * ldE + n3 + (idx - p * (n3p - n3))] = 0.0;
        }
        // charges and repulsive cores of the owned sites
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int s = threadIdx.x + k * kGradNT;
            if (s >= ns) break;
            const double sx = s4[4 * s], sy = s4[4 * s + 1], sz = s4[4 * s + 2], q = s4[4 * s + 3];
            const double z1 = sr[4 * s], c1 = sr[4 * s + 1], z2 = sr[4 * s + 2], c2 = sr[4 * s + 3];
            for (int p = 0; p < 16; ++p) {
                const double w = pt[p][3];
                const double rx = pt[p][0] - sx, ry = pt[p][1] - sy, rz = pt[p][2] - sz;
                const double R2 = rx * rx + ry * ry + rz * rz;
                if (w == 0.0 || R2 == 0.0) continue;
                const double R = sqrt(R2);
                double f = 0.0;  // dV/dR along R / R
                if (q != 0.0) {
                    double dre;
                    const double re = reff(R, D.qdamp, dre);
                    f -= q * dre / (re * re);
                }
                f -= c1 * (-z1) * exp(-z1 * R) + c2 * (-z2) * exp(-z2 * R);
                const double fw = w * f / R;
                gs[k][0] += fw * rx;
                gs[k][1] += fw * ry;
                gs[k][2] += fw * rz;
            }
        }
        __syncthreads();
        // mu = E InvA on the matrix cores, column blocks of 16 per wave; the
        // lane holds mu of point lq + 4 r, column 16 cb + l16
        const double *ea = Es + l16 * ldE + lq;
#pragma unroll
        for (int u = 0; u < kGradCBW; ++u) {
            const int cb = wave + NW * u;
            if (cb >= ncb) break;
            f64x4 o = {0.0, 0.0, 0.0, 0.0};
            const gptr<const double> bc = A + (long long)lq * n3p + 16 * cb + l16;
#pragma unroll 8
            for (int k0 = 0; k0 < n3p; k0 += 4)
                o = __builtin_amdgcn_mfma_f64_16x16x4f64(ea[k0], bc[(long long)k0 * n3p], o, 0, 0, 0);
            const int c = 16 * cb + l16;
            if (c < n3) {
                const int i = c / 3, b = c - 3 * i;
                const double px = rp[3 * i], py = rp[3 * i + 1], pz = rp[3 * i + 2];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int p = lq + 4 * r;
                    const double w = pt[p][3], mu = o[r];
                    tmu[u] += w * mu;
                    const double R3[3] = {pt[p][0] - px, pt[p][1] - py, pt[p][2] - pz};
                    const double R2 = R3[0] * R3[0] + R3[1] * R3[1] + R3[2] * R3[2];
                    if (w == 0.0 || R2 == 0.0) continue;
                    const double R = sqrt(R2);
                    double g, dg;
                    field_g(R, D.pdamp, g, dg);
                    // -(dE_i/dR_i)^T mu along a: (dg R_a R_b / R + g delta_ab) mu_b
                    const double h = w * mu * dg * R3[b] / R;
                    gp[u][0] += h * R3[0];
                    gp[u][1] += h * R3[1];
                    gp[u][2] += h * R3[2];
                    gp[u][b] += w * mu * g;
                }
            }
        }
        __syncthreads();  // Es and pt are rewritten by the next tile
    }
    // the workgroup's partial row
    double *row = part + (size_t)blockIdx.x * (3 * ns + 4 * n3p);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int s = threadIdx.x + k * kGradNT;
        if (s < ns)
            for (int a = 0; a < 3; ++a) row[3 * s + a] = gs[k][a];
    }
#pragma unroll
    for (int u = 0; u < kGradCBW; ++u) {
        const int cb = wave + NW * u;
        if (cb >= ncb) break;
        // the four lane groups lq hold the same column: sum them
        double v[4] = {tmu[u], gp[u][0], gp[u][1], gp[u][2]};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[e] += __shfl_xor(v[e], 16, 64);
            v[e] += __shfl_xor(v[e], 32, 64);
        }
        if (lq == 0) {
            const int c = 16 * cb + l16;
            row[3 * ns + c] = v[0];
            for (int a = 0; a < 3; ++a) row[3 * ns + n3p + 3 * c + a] = v[1 + a];
        }
    }
}

// S = sum_p psi_p^2 (E_p + Epc)(E_p + Epc)^T over its upper 16 x 16 blocks
// (bi <= bj, numbered row by row over the upper triangle) on the FP64 matrix
// cores.  blockIdx.x: a range of 16-point tiles; blockIdx.y: a group of
// NW * kGradBPW blocks, block q of the group taken by wave q % NW.  Per tile the
// workgroup builds the E rows once in LDS (as grad_points_kernel does), and each
// wave adds, for each of its blocks, A = E[points][16 bi ..] (16 x 4 per
// k-step) times B = psi^2 E[points][16 bj ..] into a register accumulator:
// four v_mfma_f64_16x16x4f64 per block and tile.  Partial: part[x][blk][16 x
// 16] row-major.
constexpr int kGradBPW = 8;  // blocks per wave (4 accumulator registers each)

__global__ void __launch_bounds__(kGradNT)
grad_s_kernel(GradDev D, int nx, int ny, int nz, const double *__restrict__ gx,
              const double *__restrict__ gy, const double *__restrict__ gz,
              const double *__restrict__ psi, long long tiles_per_wg, double *__restrict__ part)
{
    constexpr int NW = kGradNT / 64;
    extern __shared__ double Es[];  // 16 rows of n3p (+1)
    __shared__ double pt[16][4];    // point x, y, z and its weight psi^2
    const int n3p = D.n3p, ldE = n3p + 1, np = D.np, n3 = 3 * np;
    const int ncb = n3p / 16, nblk = ncb * (ncb + 1) / 2;
    const long long ng = (long long)nx * ny * nz;
    const long long ntiles = (ng + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const gptr<const double> rp = glb(D.rpol), ep = glb(D.epc);
    // this wave's blocks: their A and B column offsets in an E row
    int oa[kGradBPW], ob[kGradBPW];
    int nown = 0;
#pragma unroll
    for (int u = 0; u < kGradBPW; ++u) {
        int blk = blockIdx.y * NW * kGradBPW + wave + NW * u;
        oa[u] = ob[u] = 0;
        if (blk < nblk) {
            int bi = 0;
            while (blk >= ncb - bi) {
                blk -= ncb - bi;
                ++bi;
            }
            oa[u] = 16 * bi;
            ob[u] = 16 * (bi + blk);
            nown = u + 1;
        }
    }
    f64x4 acc[kGradBPW];
#pragma unroll
    for (int u = 0; u < kGradBPW; ++u) acc[u] = f64x4{0.0, 0.0, 0.0, 0.0};
    const long long t0 = (long long)blockIdx.x * tiles_per_wg;
    const long long t1 = t0 + tiles_per_wg < ntiles ? t0 + tiles_per_wg : ntiles;
    for (long long t = t0; t < t1; ++t) {
        if (threadIdx.x < 16) {
            const long long g = 16 * t + threadIdx.x;
            double r[3] = {0.0, 0.0, 0.0};
            double w = 0.0;
            if (g < ng) {
                point_of(g, nx, ny, gx, gy, gz, r);
                w = psi[g] * psi[g];
            }
            pt[threadIdx.x][0] = r[0];
            pt[threadIdx.x][1] = r[1];
            pt[threadIdx.x][2] = r[2];
            pt[threadIdx.x][3] = w;
        }
        __syncthreads();
        // E rows (electron field + permanent-charge field), zero past 3 np;
        // rows past the grid carry weight 0
        for (int idx = threadIdx.x; idx < 16 * np; idx += kGradNT) {
            const int p = idx / np, i = idx - p * np;
            const double rx = pt[p][0] - rp[3 * i], ry = pt[p][1] - rp[3 * i + 1],
                         rz = pt[p][2] - rp[3 * i + 2];
            const double R2 = rx * rx + ry * ry + rz * rz;
            double g = 0.0, dg;
            if (R2 > 0.0) field_g(sqrt(R2), D.pdamp, g, dg);
            double *row = Es + p * ldE + 3 * i;
            row[0] = g * rx + ep[3 * i];
            row[1] = g * ry + ep[3 * i + 1];
            row[2] = g * rz + ep[3 * i + 2];
        }
        for (int idx = threadIdx.x; idx < 16 * (n3p - n3); idx += kGradNT) {
            const int p = idx / (n3p - n3);
            Es[p * ldE + n3 + (idx - p * (n3p - n3))] = 0.0;
        }
        __syncthreads();
        // k-step s: points 4 s + lq; A[c1][k] = E[4 s + k][16 bi + c1],
        // B[k][c2] = w_{4 s + k} E[4 s + k][16 bj + c2]
        double w4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) w4[q] = pt[4 * q + lq][3];
        const double *er = Es + lq * ldE + l16;
#pragma unroll
        for (int u = 0; u < kGradBPW; ++u) {
            if (u < nown) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double *e = er + 4 * q * ldE;
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(e[oa[u]], w4[q] * e[ob[u]],
                                                                 acc[u], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // Es and pt are rewritten by the next tile
    }
#pragma unroll
    for (int u = 0; u < kGradBPW; ++u) {
        const int blk = blockIdx.y * NW * kGradBPW + wave + NW * u;
        if (u < nown) {
            double *out = part + ((size_t)blockIdx.x * nblk + blk) * 256;
#pragma unroll
            for (int r = 0; r < 4; ++r) out[(lq + 4 * r) * 16 + l16] = acc[u][r];
        }
    }
}

// out[c] = sum over rows of part[row][c], rows in order
__global__ void colsum_kernel(const double *__restrict__ part, long long rows, long long cols,
                              double *__restrict__ out)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    double s = 0.0;
    for (long long r = 0; r < rows; ++r) s += part[r * cols + c];
    out[c] = s;
}

// The symmetric S (n3p x n3p) from its upper 16 x 16 blocks.
__global__ void unpack_kernel(const double *__restrict__ blocks, int n3p, double *__restrict__ S)
{
    const int ncb = n3p / 16;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n3p * n3p) return;
    int r = (int)(e / n3p), c = (int)(e % n3p);
    if (r > c) {
        const int t = r;
        r = c;
        c = t;
    }
    const int bi = r / 16, bj = c / 16;
    const int blk = bi * ncb - bi * (bi - 1) / 2 + (bj - bi);
    S[e] = blocks[(size_t)blk * 256 + (r % 16) * 16 + (c % 16)];
}

// C = A B for n x n row-major matrices, one thread per element
__global__ void matmul_kernel(const double *__restrict__ A, const double *__restrict__ B, int n,
                              double *__restrict__ C)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n) return;
    const int r = (int)(e / n), c = (int)(e % n);
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += A[(size_t)r * n + k] * B[(size_t)k * n + c];
    C[e] = s;
}

}  // namespace

void gradient_terms(pisces_ctx *c, const double *psi_host, double *gel, double *tmu, double *mcm)
{
    GradState &G = c->gs;
    PC_REQUIRE(G.valid, PISCES_ERR_STATE,
               "the gradient needs a PolType 3 potential (Potential 3, DampType 2, Polarization 3) "
               "built on this context");
    const int ns = G.nsites, np = G.np, n3p = G.n3p, n3 = 3 * np, ncb = n3p / 16;
    const long long ng = c->ngp;
    static std::once_flag once;
    std::call_once(once, [] {
        PC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(grad_points_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
        PC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(grad_s_kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
    });
    const size_t lds = (size_t)16 * (n3p + 1) * sizeof(double);
    PC_REQUIRE(lds <= (size_t)kLdsMax && n3p <= 8 * kGradCBW * 16, PISCES_ERR_UNSUPPORTED,
               "too many polarizable sites for the gradient");
    PC_REQUIRE(ns <= 2 * kGradNT, PISCES_ERR_UNSUPPORTED, "too many sites for the gradient");
    GradDev D{ns, np, n3p, 0, G.qdamp, G.pdamp, G.site4.as<double>(), G.srep.as<double>(),
              G.rpol.as<double>(), G.epc.as<double>(), G.inva.as<double>()};
    const long long ntiles = (ng + 15) / 16;
    const int wg = (int)std::min<long long>(ntiles, 1024);
    const long long cols = 3LL * ns + 4LL * n3p;
    const int nblk = ncb * (ncb + 1) / 2;
    // S: groups of 8 x kGradBPW blocks; about 256 workgroups (one per CU at
    // its register use) over the tiles, so that the partials stay small
    const int ngrp = (nblk + 8 * kGradBPW - 1) / (8 * kGradBPW);
    const long long tpw = std::max<long long>(1, (ntiles * ngrp + 255) / 256);
    const int nk = (int)((ntiles + tpw - 1) / tpw);
    dbuf dpsi(sizeof(double) * ng), dpart(sizeof(double) * (size_t)wg * cols),
        dsp(sizeof(double) * (size_t)nk * nblk * 256), dtot(sizeof(double) * cols),
        dblk(sizeof(double) * (size_t)nblk * 256), dS(sizeof(double) * (size_t)n3p * n3p),
        dT(sizeof(double) * (size_t)n3p * n3p), dM(sizeof(double) * (size_t)n3p * n3p);
    PC_HIP(hipMemcpyAsync(dpsi.p, psi_host, sizeof(double) * ng, hipMemcpyHostToDevice, c->stream));
    const double *gxp = c->dx[0].as<double>(), *gyp = c->dx[1].as<double>(), *gzp = c->dx[2].as<double>();
    hipLaunchKernelGGL(grad_points_kernel, dim3(wg), dim3(kGradNT), lds, c->stream, D, c->n[0],
                       c->n[1], c->n[2], gxp, gyp, gzp, dpsi.as<double>(), dpart.as<double>());
    hipLaunchKernelGGL(grad_s_kernel, dim3(nk, ngrp), dim3(kGradNT), lds, c->stream, D, c->n[0],
                       c->n[1], c->n[2], gxp, gyp, gzp, dpsi.as<double>(), tpw, dsp.as<double>());
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0, c->stream,
                       dpart.as<double>(), (long long)wg, cols, dtot.as<double>());
    const long long sc = (long long)nblk * 256;
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((sc + 255) / 256)), dim3(256), 0, c->stream,
                       dsp.as<double>(), (long long)nk, sc, dblk.as<double>());
    const long long nn = (long long)n3p * n3p;
    const unsigned gb = (unsigned)((nn + 255) / 256);
    hipLaunchKernelGGL(unpack_kernel, dim3(gb), dim3(256), 0, c->stream, dblk.as<double>(), n3p,
                       dS.as<double>());
    // mcm = InvA S InvA
    hipLaunchKernelGGL(matmul_kernel, dim3(gb), dim3(256), 0, c->stream, dS.as<double>(),
                       G.inva.as<double>(), n3p, dT.as<double>());
    hipLaunchKernelGGL(matmul_kernel, dim3(gb), dim3(256), 0, c->stream, G.inva.as<double>(),
                       dT.as<double>(), n3p, dM.as<double>());
    PC_HIP(hipGetLastError());
    std::vector<double> tot(cols), M((size_t)nn);
    PC_HIP(hipMemcpyAsync(tot.data(), dtot.p, sizeof(double) * cols, hipMemcpyDeviceToHost, c->stream));
    PC_HIP(hipMemcpyAsync(M.data(), dM.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    c->sync();
    // site terms, then the dipole terms of each polarizable site summed over
    // its three field components
    for (int k = 0; k < 3 * ns; ++k) gel[k] = tot[k];
    const double *gp = tot.data() + 3 * ns + n3p;
    for (int i = 0; i < np; ++i)
        for (int a = 0; a < 3; ++a)
            gel[3 * G.ip[i] + a] += gp[3 * (3 * i) + a] + gp[3 * (3 * i + 1) + a] + gp[3 * (3 * i + 2) + a];
    for (int k = 0; k < n3; ++k) tmu[k] = tot[3 * ns + k];
    for (int r = 0; r < n3; ++r)
        for (int k = 0; k < n3; ++k) mcm[(size_t)r * n3 + k] = M[(size_t)r * n3p + k];
}

}  // namespace pisces
