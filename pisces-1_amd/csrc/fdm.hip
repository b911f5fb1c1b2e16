// fdm.hip -- kinetic (fast-diagonalisation) preconditioner for gfx950.
//
// The sine-DVR kinetic operator is a Kronecker sum T = Tx (+) Ty (+) Tz, and
// each 1-D matrix has a dense eigenbasis T_a = Q_a diag(l_a) Q_a^T.  Hence
//   (T - theta)^{-1} = (Qx (x) Qy (x) Qz) diag(1 / (lx_i + ly_j + lz_k - theta))
//                      (Qx (x) Qy (x) Qz)^T
// is applied exactly with six 1-D contractions.  Used as the Davidson
// correction t = -(T - theta)^{-1} r (an alternative to the diagonal
// correction r / (theta - diag H) of the reference, src/Davidson.cpp:557-566,
// which barely reduces residuals when the kinetic term dominates diag H).
// Three plane passes, each staging one plane of the vector in LDS:
//   pass 0, plane (i,j) at k: Qx^T along x, then Qy^T along y   (R -> T)
//   pass 1, plane (i,k) at j: Qz^T along z, scale, Qz along z   (T -> T)
//   pass 2, plane (i,j) at k: Qy along y, then Qx along x       (T -> T)
// Passes 1 and 2 work in place: every workgroup reads its whole plane before
// writing it, and planes are disjoint.
#include "fdm.hpp"

namespace pisces {

namespace {

constexpr int kT = 256;

// out[c][n] = sum_i M(n, i) P[c][i], M(n, i) = Q[i + nq*n] (trans: Q^T) or
// Q[n + nq*i]; P has rows c < nc, columns i < ni, pitch ld.
__device__ void contract_rows(const double *__restrict__ P, double *__restrict__ O, int ld, int ni,
                              int nc, const double *__restrict__ Q, bool trans)
{
    const int nip = (ni + 3) & ~3, ncp = (nc + 3) & ~3;
    const int nbi = nip / 4, nbc = ncp / 4;
    for (int b = threadIdx.x; b < nbi * nbc; b += kT) {
        const int n0 = 4 * (b % nbi), c0 = 4 * (b / nbi);
        double acc[4][4] = {};
        for (int i = 0; i < ni; ++i) {
            double p[4], m[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) p[r] = P[(c0 + r) * ld + i];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int nn = n0 + a;
                m[a] = (nn < ni) ? (trans ? Q[i + (long long)ni * nn] : Q[nn + (long long)ni * i]) : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][r] += m[a] * p[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int a = 0; a < 4; ++a) O[(c0 + r) * ld + n0 + a] = acc[a][r];
    }
}

// out[n][i] = sum_c M(n, c) P[c][i]  (contraction along the plane's columns)
__device__ void contract_cols(const double *__restrict__ P, double *__restrict__ O, int ld, int ni,
                              int nc, const double *__restrict__ Q, bool trans)
{
    const int nip = (ni + 3) & ~3, ncp = (nc + 3) & ~3;
    const int nbi = nip / 4, nbc = ncp / 4;
    for (int b = threadIdx.x; b < nbi * nbc; b += kT) {
        const int i0 = 4 * (b % nbi), n0 = 4 * (b / nbi);
        double acc[4][4] = {};
        for (int c = 0; c < nc; ++c) {
            double p[4], m[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) p[a] = P[c * ld + i0 + a];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int nn = n0 + r;
                m[r] = (nn < nc) ? (trans ? Q[c + (long long)nc * nn] : Q[nn + (long long)nc * c]) : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][r] += m[r] * p[a];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int a = 0; a < 4; ++a) O[(n0 + r) * ld + i0 + a] = acc[a][r];
    }
}

template <int PASS>
__global__ void __launch_bounds__(kT) fdm_kernel(const FdmTask *__restrict__ tasks)
{
    const FdmTask &t = tasks[blockIdx.y];
    const int nx = t.n[0], ny = t.n[1], nz = t.n[2];
    const int nplanes = (PASS == 1) ? ny : nz;
    const int pl = blockIdx.x;
    if (pl >= nplanes) return;
    extern __shared__ __align__(16) unsigned char smem[];
    double *P = reinterpret_cast<double *>(smem);
    const int nc = (PASS == 1) ? nz : ny;
    const int nip = (nx + 3) & ~3, ncp = (nc + 3) & ~3;
    const int ld = nip;
    double *Q2 = P + (size_t)ncp * ld;
    const long long cs = (PASS == 1) ? (long long)nx * ny : nx;
    const long long base = (PASS == 1) ? (long long)nx * pl : (long long)nx * ny * pl;
    const double *src = (PASS == 0) ? t.src : t.dst;
    for (int idx = threadIdx.x; idx < ncp * nip; idx += kT) {
        const int c = idx / nip, i = idx - c * nip;
        P[c * ld + i] = (i < nx && c < nc) ? src[base + i + c * cs] : 0.0;
        Q2[c * ld + i] = 0.0;
    }
    __syncthreads();
    if (PASS == 0) {
        contract_rows(P, Q2, ld, nx, ny, t.Q[0], true);   // Qx^T along x
        __syncthreads();
        contract_cols(Q2, P, ld, nx, ny, t.Q[1], true);   // Qy^T along y
    } else if (PASS == 1) {
        contract_cols(P, Q2, ld, nx, nz, t.Q[2], true);   // Qz^T along z
        __syncthreads();
        const double ly = t.lam[1][pl];
        for (int idx = threadIdx.x; idx < nz * nx; idx += kT) {
            const int c = idx / nx, i = idx - c * nx;
            double den = t.lam[0][i] + ly + t.lam[2][c] - t.shift;
            if (fabs(den) < t.floor) den = den < 0 ? -t.floor : t.floor;
            Q2[c * ld + i] *= t.scale / den;
        }
        __syncthreads();
        contract_cols(Q2, P, ld, nx, nz, t.Q[2], false);  // Qz along z
    } else {
        contract_cols(P, Q2, ld, nx, ny, t.Q[1], false);  // Qy along y
        __syncthreads();
        contract_rows(Q2, P, ld, nx, ny, t.Q[0], false);  // Qx along x
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nc * nx; idx += kT) {
        const int c = idx / nx, i = idx - c * nx;
        t.dst[base + i + c * cs] = P[c * ld + i];
    }
}

}  // namespace

void fdm_launch(const FdmTask *d_tasks, int ntasks, int max_n, hipStream_t s)
{
    if (ntasks <= 0) return;
    PC_REQUIRE(max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    const int np = (max_n + 3) & ~3;
    const size_t lds = 2 * (size_t)np * np * sizeof(double);
    static bool once = false;
    if (!once) {
        static_cast<void>(hipFuncSetAttribute(reinterpret_cast<const void *>(fdm_kernel<0>),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        static_cast<void>(hipFuncSetAttribute(reinterpret_cast<const void *>(fdm_kernel<1>),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        static_cast<void>(hipFuncSetAttribute(reinterpret_cast<const void *>(fdm_kernel<2>),
                                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        once = true;
    }
    PC_REQUIRE(lds <= 160 * 1024, PISCES_ERR_UNSUPPORTED, "grid too large for the LDS planes");
    hipLaunchKernelGGL(fdm_kernel<0>, dim3(max_n, ntasks), dim3(kT), lds, s, d_tasks);
    hipLaunchKernelGGL(fdm_kernel<1>, dim3(max_n, ntasks), dim3(kT), lds, s, d_tasks);
    hipLaunchKernelGGL(fdm_kernel<2>, dim3(max_n, ntasks), dim3(kT), lds, s, d_tasks);
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
