// hx.hip -- H*X on a 3-D sine-DVR grid for gfx950: dispatch, and the
// LDS-tiled VALU variant kept for cross-checking the matrix-core kernels.
//
//   Y = V o X + (Tx (x) I (x) I + I (x) Ty (x) I + I (x) I (x) Tz) X
//
// restating DVR::MatrixTimesVector (src/mtx.cpp:22-108) for a block of vectors.
// Vectors are x fastest: g = i + nx*(j + ny*k).  Two passes over each vector,
// both plane kernels that stage one 2-D plane of the vector in LDS and do the
// contractions of the two axes that lie in it:
//   pass A, plane (i,j) at fixed k:  Y  = V o X + contraction along y (Ty)
//   pass B, plane (i,k) at fixed j:  Y += contraction along x (Tx) + along z (Tz)
// The matrix-core kernels of both passes live in plane.hip.
#include "common.hpp"

#include <mutex>

namespace pisces {

namespace {

constexpr int kThreads = 256;

// Plane geometry shared by both passes.
struct PlaneGeo {
    int ni, nc;        // rows i (contiguous in memory) and columns c of the plane
    long long cs;      // global stride between columns
    long long base;    // global offset of element (0,0) of the plane (incl. vector)
    long long vbase;   // offset of (0,0) in V (no vector offset)
    const void *Tc;    // contraction matrix along c (Ty in pass A, Tz in pass B)
    const void *Ti;    // contraction matrix along i (Tx in pass B; null in pass A)
    int npc, npi;      // row pitch of Tc and Ti (padded dimension)
};

template <typename T, int PASS>
__device__ inline bool plane_geo(const HxJob &J, int p, PlaneGeo &g)
{
    g.npi = J.ax[0].np;
    const int nplanes = (PASS == 0) ? J.nz : J.ny;
    if (p >= nplanes * J.nvec) return false;
    const int v = p / nplanes, pl = p - v * nplanes;
    g.ni = J.nx;
    if (PASS == 0) {
        g.nc = J.ny;
        g.cs = J.nx;
        g.vbase = (long long)J.nx * J.ny * pl;
        g.Tc = sizeof(T) == 8 ? (const void *)J.ax[1].t64 : (const void *)J.ax[1].t32;
        g.Ti = nullptr;
        g.npc = J.ax[1].np;
    } else {
        g.nc = J.nz;
        g.cs = (long long)J.nx * J.ny;
        g.vbase = (long long)J.nx * pl;
        g.Tc = sizeof(T) == 8 ? (const void *)J.ax[2].t64 : (const void *)J.ax[2].t32;
        g.Ti = sizeof(T) == 8 ? (const void *)J.ax[0].t64 : (const void *)J.ax[0].t32;
        g.npc = J.ax[2].np;
    }
    g.base = (long long)v * J.ld + g.vbase;
    return true;
}

// ---------------------------------------------------------------- VALU form
// Each thread owns a 4(i) x 4(c) output block; T is read through the L1/L2
// (it is tiny and shared by every workgroup), the plane from LDS.
template <typename T, int PASS>
__global__ void __launch_bounds__(kThreads) hx_plane_valu(const HxJob *__restrict__ jobs)
{
    const HxJob &J = jobs[blockIdx.y];
    PlaneGeo g;
    if (!plane_geo<T, PASS>(J, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    T *P = reinterpret_cast<T *>(smem);
    const int nip = (g.ni + 3) & ~3, ncp = (g.nc + 3) & ~3;
    const int ld = nip;  // LDS row length
    const T *X = static_cast<const T *>(J.X);
    T *Y = static_cast<T *>(J.Y);
    for (int idx = threadIdx.x; idx < ncp * nip; idx += kThreads) {
        const int c = idx / nip, i = idx - c * nip;
        P[c * ld + i] = (i < g.ni && c < g.nc) ? X[g.base + i + c * g.cs] : T(0);
    }
    __syncthreads();
    const T *Tc = static_cast<const T *>(g.Tc);
    const T *Ti = static_cast<const T *>(g.Ti);
    const T *V = static_cast<const T *>(J.V);
    const int nbi = nip / 4, nbc = ncp / 4;
    for (int b = threadIdx.x; b < nbi * nbc; b += kThreads) {
        const int bi = b % nbi, bc = b / nbi;
        const int i0 = 4 * bi, c0 = 4 * bc;
        T acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[a][r] = T(0);
        // along c: out[i][c] += sum_c' Tc[c][c'] P[c'][i]   (Tc symmetric)
        for (int cp = 0; cp < g.nc; ++cp) {
            T pv[4], tv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) pv[a] = P[cp * ld + i0 + a];
#pragma unroll
            for (int r = 0; r < 4; ++r) tv[r] = Tc[(long long)cp * g.npc + c0 + r];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][r] += tv[r] * pv[a];
        }
        if (PASS == 1) {
            // along i: out[i][c] += sum_i' Ti[i][i'] P[c][i']
            for (int ip = 0; ip < g.ni; ++ip) {
                T pv[4], tv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) pv[r] = P[(c0 + r) * ld + ip];
#pragma unroll
                for (int a = 0; a < 4; ++a) tv[a] = Ti[(long long)ip * g.npi + i0 + a];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[a][r] += tv[a] * pv[r];
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = c0 + r;
            if (c >= g.nc) continue;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int i = i0 + a;
                if (i >= g.ni) continue;
                const long long gi = g.base + i + c * g.cs;
                if (PASS == 0) Y[gi] = V[g.vbase + i + c * g.cs] * P[c * ld + i] + acc[a][r];
                else Y[gi] = Y[gi] + acc[a][r];
            }
        }
    }
}

}  // namespace

void hx_launch(const HxJob *d_jobs, int njobs, int max_planes_a, int max_planes_b,
               int max_n, int precision, hipStream_t s, int variant)
{
    if (njobs <= 0) return;
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    if (variant != 0 || precision == PISCES_BF16X2) {
        hx_mfma_launch(d_jobs, njobs, max_planes_a, max_planes_b, max_n, precision, s);
        return;
    }
    static std::once_flag once;
    std::call_once(once, [] {
        for (const void *k : {reinterpret_cast<const void *>(hx_plane_valu<float, 0>),
                              reinterpret_cast<const void *>(hx_plane_valu<float, 1>),
                              reinterpret_cast<const void *>(hx_plane_valu<double, 0>),
                              reinterpret_cast<const void *>(hx_plane_valu<double, 1>)})
            static_cast<void>(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                  160 * 1024));
    });
    dim3 blk(kThreads);
    dim3 ga(max_planes_a, njobs), gb(max_planes_b, njobs);
    const int np4 = (max_n + 3) & ~3;
    if (precision == PISCES_FP32) {
        const size_t lds = (size_t)np4 * np4 * sizeof(float);
        hipLaunchKernelGGL((hx_plane_valu<float, 0>), ga, blk, lds, s, d_jobs);
        hipLaunchKernelGGL((hx_plane_valu<float, 1>), gb, blk, lds, s, d_jobs);
    } else {
        const size_t lds = (size_t)np4 * np4 * sizeof(double);
        hipLaunchKernelGGL((hx_plane_valu<double, 0>), ga, blk, lds, s, d_jobs);
        hipLaunchKernelGGL((hx_plane_valu<double, 1>), gb, blk, lds, s, d_jobs);
    }
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
