// pol3.hip -- the self-consistent polarization potential (PolType 3) on the grid.
//
// Potential::SelfConsistentPolarizability (src/Potential.cpp:1495-1592) at a
// grid point r: E_i = g(R_i) R_i + Epc_i at the polarizable sites i (R_i =
// r - site i, g the damped R^-3), mu = InvA E, Vpol = -E.mu / 2 -
// VpolWaterWater.  Over a tile of grid points that is the product E InvA (points
// x 3np times 3np x 3np) followed by a row-wise dot with E: FP64 matrix-core
// work of 2 (3np)^2 flops per point.  A workgroup takes 16 RB points (RB = 1
// by default): it builds their E rows in LDS (odd pitch: the 16 rows that one
// matrix-core operand reads fall on distinct banks), then each of its 8 or 16
// waves takes 16-column blocks of InvA, streams them from L2 (InvA is shared by all
// workgroups: at most 2.6 MB) as the B operand of v_mfma_f64_16x16x4 against
// the 16-point row blocks of E, and folds each 16 x 16 result into the
// per-point dot at once.
// InvA is symmetric, so only its upper block triangle is multiplied: (3np)^2
// (1 + 16 / 3np) flops per point.
#include "common.hpp"
#include "gmem.hpp"

#include <cstdlib>
#include <mutex>

namespace pisces {

namespace {

constexpr int kLdsMax = 160 * 1024 - 16 * 64 * (int)sizeof(double);  // E rows + red (16 waves)

typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ inline double reff3(double r, double r0)
{
    // effective-distance damping, src/Potential.cpp:1191-1196
    if (r < r0) {
        const double t = r / r0;
        return r0 * (0.5 + t * t * t * (1.0 - 0.5 * t));
    }
    return r;
}

// RB 16-point row blocks per workgroup (kP3Pts = 16 RB grid points), NT threads
template <int RB, int NT>
__global__ void __launch_bounds__(NT)
pol3_grid_kernel(Pol3Dev D, int nx, int ny, int nz, const double *__restrict__ gx,
                 const double *__restrict__ gy, const double *__restrict__ gz,
                 double *__restrict__ v)
{
    constexpr int kP3Pts = 16 * RB, kP3Threads = NT, NW = NT / 64;
    extern __shared__ double Es[];  // kP3Pts rows of n3p (+1) values
    __shared__ double red[kP3Threads / 64][kP3Pts];
    const int n3p = D.n3p, ldE = n3p + 1, np = D.np, n3 = 3 * np;
    const long long ng = (long long)nx * ny * nz;
    const long long g0 = (long long)blockIdx.x * kP3Pts;
    const gptr<const double> rp = glb(D.rpol), ep = glb(D.epc);

    // E rows: the electron's field at the sites (src/Potential.cpp:1531-1562,
    // DampType 1: R^-3 (1 - exp(-p R^3)); 2: Reff^-3 inside the damping radius)
    // plus the permanent-charge field; rows past the grid stay 0
    for (int idx = threadIdx.x; idx < kP3Pts * np; idx += kP3Threads) {
        const int p = idx / np, i = idx - p * np;
        const long long g = g0 + p;
        double e0 = 0.0, e1 = 0.0, e2 = 0.0;
        if (g < ng) {
            const int ix = (int)(g % nx), iy = (int)((g / nx) % ny);
            const int iz = (int)(g / ((long long)nx * ny));
            const double rx = gx[ix] - rp[3 * i], ry = gy[iy] - rp[3 * i + 1],
                         rz = gz[iz] - rp[3 * i + 2];
            const double R2 = rx * rx + ry * ry + rz * rz;
            double gij = 0.0;  // R = 0: the field's limit, 0
            if (R2 > 0.0) {
                const double ri = rsq64(R2), R = R2 * ri;
                if (D.damp_type == 1) {
                    gij = ri * ri * ri * (1.0 - exp(-D.pdamp * R * R2));
                } else if (R < D.pdamp) {
                    const double re = reff3(R, D.pdamp);
                    gij = 1.0 / (re * re * re);
                } else {
                    gij = ri * ri * ri;
                }
            }
            e0 = gij * rx + ep[3 * i];
            e1 = gij * ry + ep[3 * i + 1];
            e2 = gij * rz + ep[3 * i + 2];
        }
        double *row = Es + p * ldE + 3 * i;
        row[0] = e0;
        row[1] = e1;
        row[2] = e2;
    }
    const int padw = n3p - n3;
    for (int idx = threadIdx.x; idx < kP3Pts * padw; idx += kP3Threads) {
        const int p = idx / padw;
        Es[p * ldE + n3 + (idx - p * padw)] = 0.0;
    }
    __syncthreads();

    // E^T InvA E over the upper block triangle of the symmetric InvA: column
    // block cb takes rows k < 16 cb twice and its diagonal 16 x 16 tile once.
    // Block cb costs cb + 1 row blocks, so the waves take the blocks of each
    // group of four in alternating order to even out their work.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const gptr<const double> A = glb(D.inva);
    double sr[RB][4];
#pragma unroll
    for (int b = 0; b < RB; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) sr[b][r] = 0.0;
    const double *ea = Es + l16 * ldE + lq;  // row block b: + 16 b ldE
    const int ncb = n3p / 16;
    for (int j = 0; NW * j < ncb; ++j) {
        const int cb = (j & 1) ? NW * j + NW - 1 - wave : NW * j + wave;
        if (cb >= ncb) continue;
        f64x4 o[RB], d[RB];
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            o[b] = f64x4{0.0, 0.0, 0.0, 0.0};
            d[b] = f64x4{0.0, 0.0, 0.0, 0.0};
        }
        const gptr<const double> bc = A + (long long)lq * n3p + 16 * cb + l16;
        const int kd = 16 * cb;
#pragma unroll 8
        for (int k0 = 0; k0 < kd; k0 += 4) {
            const double bv = bc[(long long)k0 * n3p];  // InvA[k0 + lq][16 cb + l16]
#pragma unroll
            for (int b = 0; b < RB; ++b)
                o[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ea[16 * b * ldE + k0], bv, o[b], 0, 0, 0);
        }
#pragma unroll
        for (int k0 = kd; k0 < kd + 16; k0 += 4) {
            const double bv = bc[(long long)k0 * n3p];
#pragma unroll
            for (int b = 0; b < RB; ++b)
                d[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ea[16 * b * ldE + k0], bv, d[b], 0, 0, 0);
        }
        // o[b][r], d[b][r]: point 16 b + lq + 4 r, column 16 cb + l16
        const int c = 16 * cb + l16;
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                sr[b][r] += (2.0 * o[b][r] + d[b][r]) * Es[(16 * b + lq + 4 * r) * ldE + c];
    }
    // sum over the 16 columns a lane group holds, then over the waves
#pragma unroll
    for (int off = 1; off < 16; off <<= 1)
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) sr[b][r] += __shfl_xor(sr[b][r], off, 64);
    if (l16 == 0)
#pragma unroll
        for (int b = 0; b < RB; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[wave][16 * b + lq + 4 * r] = sr[b][r];
    __syncthreads();
    if (threadIdx.x < kP3Pts) {
        const long long g = g0 + threadIdx.x;
        if (g < ng) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kP3Threads / 64; ++w) s += red[w][threadIdx.x];
            v[g] += -0.5 * s - D.vpolww;
        }
    }
}

}  // namespace

void pol3_grid_launch(const Pol3Dev &D, int nx, int ny, int nz, const double *d_x,
                      const double *d_y, const double *d_z, double *d_v, hipStream_t s)
{
    PC_REQUIRE(D.np > 0 && D.n3p == pol3_padded(D.np) && D.n3p <= kPol3MaxN,
               PISCES_ERR_UNSUPPORTED, "too many polarizable sites");
    PC_REQUIRE(D.damp_type == 1 || D.damp_type == 2, PISCES_ERR_ARG, "damping type");
    static std::once_flag once;
    std::call_once(once, [] {
        const void *ks[] = {reinterpret_cast<const void *>(pol3_grid_kernel<1, 256>),
                            reinterpret_cast<const void *>(pol3_grid_kernel<2, 256>),
                            reinterpret_cast<const void *>(pol3_grid_kernel<4, 256>),
                            reinterpret_cast<const void *>(pol3_grid_kernel<1, 512>),
                            reinterpret_cast<const void *>(pol3_grid_kernel<1, 1024>)};
        for (const void *k : ks)
            PC_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, kLdsMax));
    });
    // RB 16-point row blocks per workgroup (PISCES_POL3_RB = 1, 2 or 4).  The E
    // rows in LDS bound how many workgroups a CU holds, and that decides the
    // speed: 16-point tiles (1) measured 1.27x / 1.16x faster than 32-point
    // ones at 216 / 405 polarizable dimensions, 64-point tiles slower still,
    // although each InvA fragment then feeds fewer matrix-core products.
    const char *rb_e = std::getenv("PISCES_POL3_RB");  // read per build: not a hot path
    const int rb_env = rb_e ? std::atoi(rb_e) : 0;
    // threads per workgroup (PISCES_POL3_NT = 256, 512 or 1024): the waves of a
    // workgroup share its column blocks, and more resident waves hide the L2
    // latency of the InvA operand.  Measured at 16-point tiles: 256 / 512 /
    // 1024 threads 0.578 / 0.419 / 0.487 ms at 216 dimensions (14 blocks),
    // 8.74 / 4.66 / 4.44 ms at 405 (26 blocks): 1024 from 24 blocks on.
    const char *nt_e = std::getenv("PISCES_POL3_NT");
    const int nt_v = nt_e ? std::atoi(nt_e) : (D.n3p / 16 >= 24 ? 1024 : 512);
    const int nt = (nt_v == 256 || nt_v == 1024) ? nt_v : 512;
    const size_t row = (size_t)(D.n3p + 1) * sizeof(double);
    int rb = (rb_env == 1 || rb_env == 2 || rb_env == 4) ? rb_env : 1;
    while (rb > 1 && 16 * rb * row > (size_t)kLdsMax) rb /= 2;
    const long long ng = (long long)nx * ny * nz;
    const int pts = 16 * rb;
    const unsigned blocks = (unsigned)((ng + pts - 1) / pts);
    const size_t lds = (size_t)pts * row;
    if (rb == 1 && nt == 1024)
        hipLaunchKernelGGL((pol3_grid_kernel<1, 1024>), dim3(blocks), dim3(1024), lds, s, D, nx, ny,
                           nz, d_x, d_y, d_z, d_v);
    else if (rb == 1 && nt == 512)
        hipLaunchKernelGGL((pol3_grid_kernel<1, 512>), dim3(blocks), dim3(512), lds, s, D, nx, ny,
                           nz, d_x, d_y, d_z, d_v);
    else if (rb == 4)
        hipLaunchKernelGGL((pol3_grid_kernel<4, 256>), dim3(blocks), dim3(256), lds, s, D, nx, ny,
                           nz, d_x, d_y, d_z, d_v);
    else if (rb == 2)
        hipLaunchKernelGGL((pol3_grid_kernel<2, 256>), dim3(blocks), dim3(256), lds, s, D, nx, ny,
                           nz, d_x, d_y, d_z, d_v);
    else
        hipLaunchKernelGGL((pol3_grid_kernel<1, 256>), dim3(blocks), dim3(256), lds, s, D, nx, ny,
                           nz, d_x, d_y, d_z, d_v);
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
