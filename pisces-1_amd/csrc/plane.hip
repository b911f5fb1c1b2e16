// plane.hip -- matrix-core plane kernels for gfx950: H*X (FP64, FP32 and
// split-BF16 forms) and the kinetic fast-diagonalisation preconditioner
// (FP64 and split-BF16 forms).
//
// Both operators are sums or products of dense 1-D matrices applied along the
// axes of the x-fastest grid (g = i + nx*(j + ny*k)).  A workgroup (4 waves)
// stages one 2-D plane of one vector in LDS -- rows c, columns i, both padded
// to multiples of 32 -- and contracts it along one or both of its axes:
//   along c:  D(c, i) += sum_k Mc[c][k] P[k][i]   (A = Mc from HBM/L2, B = plane)
//   along i:  D(c, i) += sum_k P[c][k] Mi[k][i]   (A = plane, B = Mi)
// on 16 x 16 output tiles, wave w taking tile rows w, w+4, ... and two
// column tiles at a time (two independent accumulator chains).
//
// Instruction operand layouts (lane l, l16 = l % 16, lq = l / 16):
//   f64 16x16x4 : A[l16][lq], B[lq][l16], D element r = D[lq + 4r][l16]
//   f32 16x16x4 : A[l16][lq], B[lq][l16], D element r = D[4lq + r][l16]
//   bf16 16x16x32 (f32 acc): A[l16][8lq + e], B[8lq + e][l16], e < 8,
//                 D element r = D[4lq + r][l16]
// The split-BF16 form stores every operand as hi = bf16(x), lo = bf16(x - hi)
// and accumulates hi*hi + hi*lo + lo*hi in FP32 (about 16 significant bits).
// Its B operand needs 8 consecutive k per lane, so the plane is kept in LDS
// transposed (i-major) for contractions along c and as is (c-major) for
// contractions along i, and the matrix Mi is read as Mi^T (row-major); every
// operand fetch is then one 16-byte load.
//
// H*X: pass A, plane (i,j) at k:  Y  = V o X + Ty along j
//      pass B, plane (i,k) at j:  Y += Tx along i + Tz along k
// Preconditioner t = scale (T - shift)^{-1} r through the eigenbases
// (T_a = Q_a diag(lam_a) Q_a^T), three passes:
//      0, plane (i,j) at k:  W = Qy^T (P Qx)                    (r -> W)
//      1, plane (i,k) at j:  W = Qz (S o (Qz^T P)),  S = scale / (lam_x + lam_y + lam_z - shift)
//      2, plane (i,j) at k:  t = (Qy P) Qx^T                     (W -> t)
// with |lam_x + lam_y + lam_z - shift| floored at `floor`.
#include "common.hpp"
#include "fdm.hpp"

#include <mutex>

namespace pisces {

namespace {

constexpr int kPT = 256;

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ uint16_t bf16_rne(float x)
{
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ void split2(float x, uint16_t &h, uint16_t &l)
{
    h = bf16_rne(x);
    l = bf16_rne(x - __uint_as_float((unsigned int)h << 16));
}
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(p));
}
__device__ __forceinline__ f32x4 mma_bf(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mma(double a, double b, f64x4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma(float a, float b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T> struct Acc;
template <> struct Acc<double> {
    typedef f64x4 type;
    __device__ static inline int row(int lq, int r) { return lq + 4 * r; }
};
template <> struct Acc<float> {
    typedef f32x4 type;
    __device__ static inline int row(int lq, int r) { return 4 * lq + r; }
};

// Plane geometry.  kind 0: plane (i,j) at fixed k (rows c = j); kind 1: plane
// (i,k) at fixed j (rows c = k).
struct Geo {
    int ni, nc, npi, npc;
    long long cs;     // global stride between rows
    long long vbase;  // offset of (0, 0) within one vector
    long long base;   // offset of (0, 0) including the vector offset
    int pl;           // index of the plane along its normal axis
};

__device__ __forceinline__ bool geo_of(int kind, int nx, int ny, int nz, int nvec, long long ld,
                                       int npi, int npc, int p, Geo &g)
{
    const int nplanes = kind == 0 ? nz : ny;
    if (p >= nplanes * nvec) return false;
    const int v = p / nplanes, pl = p - v * nplanes;
    g.ni = nx;
    g.npi = npi;
    g.npc = npc;
    g.pl = pl;
    if (kind == 0) {
        g.nc = ny;
        g.cs = nx;
        g.vbase = (long long)nx * ny * pl;
    } else {
        g.nc = nz;
        g.cs = (long long)nx * ny;
        g.vbase = (long long)nx * pl;
    }
    g.base = (long long)v * ld + g.vbase;
    return true;
}

// ------------------------------------------------------------------ staging
// Plane values as FP32 for items of 2 rows x 4 columns.
template <typename S>
__device__ __forceinline__ void load_item(const S *__restrict__ src, const Geo &g, int c, int i,
                                          bool vec, float (&x)[2][4])
{
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int cc = c + r;
        if (cc >= g.nc) {
#pragma unroll
            for (int q = 0; q < 4; ++q) x[r][q] = 0.f;
            continue;
        }
        const S *row = src + g.base + cc * g.cs;
        if (vec && i + 4 <= g.ni) {
            if (sizeof(S) == 4) {
                const float4 v = *reinterpret_cast<const float4 *>(row + i);
                x[r][0] = v.x; x[r][1] = v.y; x[r][2] = v.z; x[r][3] = v.w;
            } else {
                const double2 a = *reinterpret_cast<const double2 *>(row + i);
                const double2 b = *reinterpret_cast<const double2 *>(row + i + 2);
                x[r][0] = (float)a.x; x[r][1] = (float)a.y; x[r][2] = (float)b.x; x[r][3] = (float)b.y;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) x[r][q] = (i + q < g.ni) ? (float)row[i + q] : 0.f;
        }
    }
}

// Split-BF16 plane into LDS: c-major (rows c, pitch npi + 8) and/or i-major
// (rows i, pitch npc + 8); null pointers skip a layout.
template <typename S>
__device__ void stage_bf2(const S *__restrict__ src, const Geo &g, bool vec, uint16_t *Pch,
                          uint16_t *Pcl, uint16_t *Pth, uint16_t *Ptl)
{
    const int nq = g.npi / 4, items = (g.npc / 2) * nq;
    const int pc = g.npi + 8, pt = g.npc + 8;
    for (int it = threadIdx.x; it < items; it += kPT) {
        const int c = 2 * (it / nq), i = 4 * (it - (it / nq) * nq);
        float x[2][4];
        load_item(src, g, c, i, vec, x);
        uint16_t h[2][4], l[2][4];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) split2(x[r][q], h[r][q], l[r][q]);
        if (Pch) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                u32x2 vh, vl;
                vh[0] = h[r][0] | ((unsigned)h[r][1] << 16);
                vh[1] = h[r][2] | ((unsigned)h[r][3] << 16);
                vl[0] = l[r][0] | ((unsigned)l[r][1] << 16);
                vl[1] = l[r][2] | ((unsigned)l[r][3] << 16);
                *reinterpret_cast<u32x2 *>(Pch + (c + r) * pc + i) = vh;
                *reinterpret_cast<u32x2 *>(Pcl + (c + r) * pc + i) = vl;
            }
        }
        if (Pth) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                *reinterpret_cast<unsigned *>(Pth + (i + q) * pt + c) = h[0][q] | ((unsigned)h[1][q] << 16);
                *reinterpret_cast<unsigned *>(Ptl + (i + q) * pt + c) = l[0][q] | ((unsigned)l[1][q] << 16);
            }
        }
    }
}

// FP32/FP64 plane into LDS, c-major, pitch npi + pad.
template <typename T, typename S>
__device__ void stage_f(const S *__restrict__ src, const Geo &g, bool vec, T *P, int pitch)
{
    const int nq = g.npi / 4, items = (g.npc / 2) * nq;
    for (int it = threadIdx.x; it < items; it += kPT) {
        const int c = 2 * (it / nq), i = 4 * (it - (it / nq) * nq);
        if (sizeof(S) == 8 && sizeof(T) == 8) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int cc = c + r;
                const S *row = src + g.base + (long long)cc * g.cs;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    P[cc * pitch + i + q] = (cc < g.nc && i + q < g.ni) ? (T)row[i + q] : T(0);
            }
        } else {
            float x[2][4];
            load_item(src, g, c, i, vec, x);
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) P[(c + r) * pitch + i + q] = (T)x[r][q];
        }
    }
}

__device__ __forceinline__ bool vec_ok(const void *p, const Geo &g, int elems_per_16b)
{
    return ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && (g.ni % 4 == 0) && (g.cs % 4 == 0) &&
           (g.base % elems_per_16b == 0) && (elems_per_16b == 4 || g.base % 2 == 0);
}

// f64/f32 pitch: np + 2 doubles (odd number of 16-byte units per row) or np + 4 floats
template <typename T> __host__ __device__ inline int fpitch(int np) { return sizeof(T) == 8 ? np + 2 : np + 4; }

// ------------------------------------------------------------------ H*X, split-BF16 form
template <int PASS, int NPMAX>
__global__ void __launch_bounds__(kPT) hx_bf2_kernel(const HxJob *__restrict__ jobs)
{
    const HxJob &J = jobs[blockIdx.y];
    const AxisMats &Mi = J.ax[0];
    const AxisMats &Mc = J.ax[PASS == 0 ? 1 : 2];
    Geo g;
    if (!geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pt = g.npc + 8, pc = g.npi + 8;
    uint16_t *Pth = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Ptl = Pth + g.npi * pt;
    uint16_t *Pch = Ptl + g.npi * pt;
    uint16_t *Pcl = Pch + g.npc * pc;
    const float *X = static_cast<const float *>(J.X);
    float *Y = static_cast<float *>(J.Y);
    const bool vec = vec_ok(X, g, 4);
    stage_bf2(X, g, vec, PASS == 1 ? Pch : nullptr, Pcl, Pth, Ptl);
    __syncthreads();
    const float *V = static_cast<const float *>(J.V);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16, ksc = g.npc / 32, ksi = g.npi / 32;
    constexpr int KS = NPMAX / 32;
    for (int tc = wave; tc < ntc; tc += 4) {
        const int c0 = 16 * tc;
        bf16x8 ach[KS], acl[KS], aih[KS], ail[KS];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if (s >= ksc) break;
            ach[s] = ld8(Mc.th + (c0 + l16) * g.npc + 32 * s + 8 * lq);
            acl[s] = ld8(Mc.tl + (c0 + l16) * g.npc + 32 * s + 8 * lq);
        }
        if (PASS == 1) {
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (s >= ksi) break;
                aih[s] = ld8(Pch + (c0 + l16) * pc + 32 * s + 8 * lq);
                ail[s] = ld8(Pcl + (c0 + l16) * pc + 32 * s + 8 * lq);
            }
        }
        for (int ti = 0; ti < nti; ti += 2) {
            const int i0 = 16 * ti, i1 = i0 + 16;
            f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                if (s >= ksc) break;
                const bf16x8 b0h = ld8(Pth + (i0 + l16) * pt + 32 * s + 8 * lq);
                const bf16x8 b0l = ld8(Ptl + (i0 + l16) * pt + 32 * s + 8 * lq);
                const bf16x8 b1h = ld8(Pth + (i1 + l16) * pt + 32 * s + 8 * lq);
                const bf16x8 b1l = ld8(Ptl + (i1 + l16) * pt + 32 * s + 8 * lq);
                a0 = mma_bf(ach[s], b0h, a0);
                a1 = mma_bf(ach[s], b1h, a1);
                a0 = mma_bf(ach[s], b0l, a0);
                a1 = mma_bf(ach[s], b1l, a1);
                a0 = mma_bf(acl[s], b0h, a0);
                a1 = mma_bf(acl[s], b1h, a1);
            }
            if (PASS == 1) {
#pragma unroll
                for (int s = 0; s < KS; ++s) {
                    if (s >= ksi) break;
                    // B[k][n] = Tx[k][i0 + n] = Tx[i0 + n][k] (symmetric)
                    const bf16x8 b0h = ld8(Mi.th + (i0 + l16) * g.npi + 32 * s + 8 * lq);
                    const bf16x8 b0l = ld8(Mi.tl + (i0 + l16) * g.npi + 32 * s + 8 * lq);
                    const bf16x8 b1h = ld8(Mi.th + (i1 + l16) * g.npi + 32 * s + 8 * lq);
                    const bf16x8 b1l = ld8(Mi.tl + (i1 + l16) * g.npi + 32 * s + 8 * lq);
                    a0 = mma_bf(aih[s], b0h, a0);
                    a1 = mma_bf(aih[s], b1h, a1);
                    a0 = mma_bf(aih[s], b0l, a0);
                    a1 = mma_bf(aih[s], b1l, a1);
                    a0 = mma_bf(ail[s], b0h, a0);
                    a1 = mma_bf(ail[s], b1h, a1);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = c0 + 4 * lq + r;
                if (c >= g.nc) continue;
                const long long rowg = g.base + (long long)c * g.cs;
                const int ia = i0 + l16, ib = i1 + l16;
                if (PASS == 0) {
                    const long long rowv = g.vbase + (long long)c * g.cs;
                    if (ia < g.ni) Y[rowg + ia] = V[rowv + ia] * X[rowg + ia] + a0[r];
                    if (ib < g.ni) Y[rowg + ib] = V[rowv + ib] * X[rowg + ib] + a1[r];
                } else {
                    if (ia < g.ni) Y[rowg + ia] += a0[r];
                    if (ib < g.ni) Y[rowg + ib] += a1[r];
                }
            }
        }
    }
}

// ------------------------------------------------------------------ H*X, FP64 / FP32 forms
template <typename T> __device__ inline const T *tmat(const AxisMats &m);
template <> __device__ inline const double *tmat<double>(const AxisMats &m) { return m.t64; }
template <> __device__ inline const float *tmat<float>(const AxisMats &m) { return m.t32; }

template <typename T, int PASS>
__global__ void __launch_bounds__(kPT) hx_f_kernel(const HxJob *__restrict__ jobs)
{
    typedef typename Acc<T>::type acc_t;
    const HxJob &J = jobs[blockIdx.y];
    const AxisMats &Mi = J.ax[0];
    const AxisMats &Mc = J.ax[PASS == 0 ? 1 : 2];
    Geo g;
    if (!geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pitch = fpitch<T>(g.npi);
    T *P = reinterpret_cast<T *>(smem);
    const T *X = static_cast<const T *>(J.X);
    T *Y = static_cast<T *>(J.Y);
    stage_f<T, T>(X, g, vec_ok(X, g, 16 / (int)sizeof(T)), P, pitch);
    __syncthreads();
    const T *V = static_cast<const T *>(J.V);
    const T *Tc = tmat<T>(Mc), *Ti = tmat<T>(Mi);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
    for (int tc = wave; tc < ntc; tc += 4) {
        const int c0 = 16 * tc;
        for (int ti = 0; ti < nti; ti += 2) {
            const int i0 = 16 * ti, i1 = i0 + 16;
            acc_t a0 = {}, a1 = {};
            // along c: A[m][k] = Tc[c0 + m][k], B[k][n] = P[k][i + n]
            const T *tcr = Tc + (c0 + l16) * g.npc + lq;
#pragma unroll 4
            for (int k0 = 0; k0 < g.npc; k0 += 4) {
                const T a = tcr[k0];
                const T *pr = P + (k0 + lq) * pitch + l16;
                a0 = mma(a, pr[i0], a0);
                a1 = mma(a, pr[i1], a1);
            }
            if (PASS == 1) {
                // along i: A[m][k] = P[c0 + m][k], B[k][n] = Ti[k][i + n]
                const T *prow = P + (c0 + l16) * pitch + lq;
#pragma unroll 4
                for (int k0 = 0; k0 < g.npi; k0 += 4) {
                    const T a = prow[k0];
                    const T *tr = Ti + (k0 + lq) * g.npi + l16;
                    a0 = mma(a, tr[i0], a0);
                    a1 = mma(a, tr[i1], a1);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = c0 + Acc<T>::row(lq, r);
                if (c >= g.nc) continue;
                const long long rowg = g.base + (long long)c * g.cs;
                const int ia = i0 + l16, ib = i1 + l16;
                if (PASS == 0) {
                    const long long rowv = g.vbase + (long long)c * g.cs;
                    if (ia < g.ni) Y[rowg + ia] = V[rowv + ia] * P[c * pitch + ia] + a0[r];
                    if (ib < g.ni) Y[rowg + ib] = V[rowv + ib] * P[c * pitch + ib] + a1[r];
                } else {
                    if (ia < g.ni) Y[rowg + ia] += a0[r];
                    if (ib < g.ni) Y[rowg + ib] += a1[r];
                }
            }
        }
    }
}

// ------------------------------------------------------------------ preconditioner, FP64
// Contract the LDS plane P (c-major) into the wave's accumulator tiles:
// along c with Mc (row-major), or along i with Mi (row-major).
template <int NCT, int NIT>
__device__ __forceinline__ void f64_contract(const double *P, int pitch, const Geo &g,
                                             const double *M, bool along_c, f64x4 (&acc)[NCT][NIT])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
        const int c0 = 16 * tc;
#pragma unroll
        for (int v = 0; v < NIT; v += 2) {
            if (v >= nti) break;
            const int i0 = 16 * v, i1 = i0 + 16;
            f64x4 a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
            if (along_c) {
                const double *mr = M + (c0 + l16) * g.npc + lq;
#pragma unroll 4
                for (int k0 = 0; k0 < g.npc; k0 += 4) {
                    const double a = mr[k0];
                    const double *pr = P + (k0 + lq) * pitch + l16;
                    a0 = mma(a, pr[i0], a0);
                    a1 = mma(a, pr[i1], a1);
                }
            } else {
                const double *prow = P + (c0 + l16) * pitch + lq;
#pragma unroll 4
                for (int k0 = 0; k0 < g.npi; k0 += 4) {
                    const double a = prow[k0];
                    const double *mr = M + (k0 + lq) * g.npi + l16;
                    a0 = mma(a, mr[i0], a0);
                    a1 = mma(a, mr[i1], a1);
                }
            }
            acc[u][v] = a0;
            acc[u][v + 1] = a1;
        }
    }
}

// Write the wave's tiles back into the LDS plane (c-major), optionally scaled
// per element by the preconditioner denominators.
template <int NCT, int NIT>
__device__ __forceinline__ void f64_store_lds(double *P, int pitch, const Geo &g,
                                              const f64x4 (&acc)[NCT][NIT], const FdmTask *t)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            if (v >= nti) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = 16 * tc + lq + 4 * r, i = 16 * v + l16;
                double x = acc[u][v][r];
                if (t) {
                    // plane (i = a, c = z index) at fixed j = b
                    double den = 1.0;
                    if (i < g.ni && c < g.nc) {
                        den = t->ax[0].lam[i] + t->ax[1].lam[g.pl] + t->ax[2].lam[c] - t->shift;
                        if (fabs(den) < t->floor) den = den < 0 ? -t->floor : t->floor;
                    }
                    x *= t->scale / den;
                }
                P[c * pitch + i] = x;
            }
        }
    }
}

template <typename D, int NCT, int NIT>
__device__ __forceinline__ void f64_store_global(D *dst, const Geo &g, const f64x4 (&acc)[NCT][NIT])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            if (v >= nti) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = 16 * tc + lq + 4 * r, i = 16 * v + l16;
                if (c < g.nc && i < g.ni) dst[g.base + (long long)c * g.cs + i] = (D)acc[u][v][r];
            }
        }
    }
}

template <int PASS, int NPMAX>
__global__ void __launch_bounds__(kPT) fdm_f64_kernel(const FdmTask *__restrict__ tasks)
{
    constexpr int NIT = NPMAX / 16, NCT = (NPMAX / 16 + 3) / 4;
    const FdmTask &t = tasks[blockIdx.y];
    Geo g;
    const int kind = PASS == 1 ? 1 : 0;
    const AxisMats &Mx = t.ax[0], &My = t.ax[1], &Mz = t.ax[2];
    const AxisMats &Mc = PASS == 1 ? Mz : My;
    if (!geo_of(kind, Mx.n, My.n, Mz.n, 1, 0, Mx.np, Mc.np, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pitch = fpitch<double>(g.npi);
    double *P = reinterpret_cast<double *>(smem);
    const double *src = static_cast<const double *>(PASS == 0 ? t.src : t.dst);
    stage_f<double, double>(src, g, false, P, pitch);
    __syncthreads();
    f64x4 acc[NCT][NIT];
    if (PASS == 0) f64_contract(P, pitch, g, Mx.q64, false, acc);         // P Qx
    else if (PASS == 1) f64_contract(P, pitch, g, Mz.qt64, true, acc);    // Qz^T P
    else f64_contract(P, pitch, g, My.q64, true, acc);                    // Qy P
    __syncthreads();
    f64_store_lds(P, pitch, g, acc, PASS == 1 ? &t : nullptr);
    __syncthreads();
    if (PASS == 0) f64_contract(P, pitch, g, My.qt64, true, acc);         // Qy^T (.)
    else if (PASS == 1) f64_contract(P, pitch, g, Mz.q64, true, acc);     // Qz (.)
    else f64_contract(P, pitch, g, Mx.qt64, false, acc);                  // (.) Qx^T
    f64_store_global(static_cast<double *>(t.dst), g, acc);
}

// ------------------------------------------------------------------ preconditioner, split-BF16
// Contract a split-BF16 LDS plane: along c (B from the i-major copy Pt) with
// the A-form matrix Mc (row-major), or along i (A from the c-major copy Pc)
// with MiT = Mi^T (row-major).
template <int NCT, int NIT>
__device__ __forceinline__ void bf2_contract(const uint16_t *Ph, const uint16_t *Pl, const Geo &g,
                                             const uint16_t *Mh, const uint16_t *Ml, bool along_c,
                                             f32x4 (&acc)[NCT][NIT])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
    const int pt = g.npc + 8, pc = g.npi + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
        const int c0 = 16 * tc;
#pragma unroll
        for (int v = 0; v < NIT; v += 2) {
            if (v >= nti) break;
            const int i0 = 16 * v, i1 = i0 + 16;
            f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
            if (along_c) {
                for (int k0 = 0; k0 < g.npc; k0 += 32) {
                    const bf16x8 ah = ld8(Mh + (c0 + l16) * g.npc + k0 + 8 * lq);
                    const bf16x8 al = ld8(Ml + (c0 + l16) * g.npc + k0 + 8 * lq);
                    const bf16x8 b0h = ld8(Ph + (i0 + l16) * pt + k0 + 8 * lq);
                    const bf16x8 b0l = ld8(Pl + (i0 + l16) * pt + k0 + 8 * lq);
                    const bf16x8 b1h = ld8(Ph + (i1 + l16) * pt + k0 + 8 * lq);
                    const bf16x8 b1l = ld8(Pl + (i1 + l16) * pt + k0 + 8 * lq);
                    a0 = mma_bf(ah, b0h, a0);
                    a1 = mma_bf(ah, b1h, a1);
                    a0 = mma_bf(ah, b0l, a0);
                    a1 = mma_bf(ah, b1l, a1);
                    a0 = mma_bf(al, b0h, a0);
                    a1 = mma_bf(al, b1h, a1);
                }
            } else {
                for (int k0 = 0; k0 < g.npi; k0 += 32) {
                    const bf16x8 ah = ld8(Ph + (c0 + l16) * pc + k0 + 8 * lq);
                    const bf16x8 al = ld8(Pl + (c0 + l16) * pc + k0 + 8 * lq);
                    const bf16x8 b0h = ld8(Mh + (i0 + l16) * g.npi + k0 + 8 * lq);
                    const bf16x8 b0l = ld8(Ml + (i0 + l16) * g.npi + k0 + 8 * lq);
                    const bf16x8 b1h = ld8(Mh + (i1 + l16) * g.npi + k0 + 8 * lq);
                    const bf16x8 b1l = ld8(Ml + (i1 + l16) * g.npi + k0 + 8 * lq);
                    a0 = mma_bf(ah, b0h, a0);
                    a1 = mma_bf(ah, b1h, a1);
                    a0 = mma_bf(ah, b0l, a0);
                    a1 = mma_bf(ah, b1l, a1);
                    a0 = mma_bf(al, b0h, a0);
                    a1 = mma_bf(al, b1h, a1);
                }
            }
            acc[u][v] = a0;
            acc[u][v + 1] = a1;
        }
    }
}

// Write the tiles (optionally scaled) back into LDS as split BF16, i-major
// (for a following contraction along c) or c-major (along i).
template <int NCT, int NIT>
__device__ __forceinline__ void bf2_store_lds(uint16_t *Ph, uint16_t *Pl, const Geo &g,
                                              const f32x4 (&acc)[NCT][NIT], bool imajor,
                                              const FdmTask *t)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
    const int pt = g.npc + 8, pc = g.npi + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            if (v >= nti) break;
            const int i = 16 * v + l16, cb = 16 * tc + 4 * lq;
            uint16_t h[4], l[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float x = acc[u][v][r];
                if (t) {
                    const int c = cb + r;
                    double den = 1.0;
                    if (i < g.ni && c < g.nc) {
                        den = t->ax[0].lam[i] + t->ax[1].lam[g.pl] + t->ax[2].lam[c] - t->shift;
                        if (fabs(den) < t->floor) den = den < 0 ? -t->floor : t->floor;
                    }
                    x = (float)((double)x * (t->scale / den));
                }
                split2(x, h[r], l[r]);
            }
            if (imajor) {
                u32x2 vh, vl;
                vh[0] = h[0] | ((unsigned)h[1] << 16);
                vh[1] = h[2] | ((unsigned)h[3] << 16);
                vl[0] = l[0] | ((unsigned)l[1] << 16);
                vl[1] = l[2] | ((unsigned)l[3] << 16);
                *reinterpret_cast<u32x2 *>(Ph + i * pt + cb) = vh;
                *reinterpret_cast<u32x2 *>(Pl + i * pt + cb) = vl;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    Ph[(cb + r) * pc + i] = h[r];
                    Pl[(cb + r) * pc + i] = l[r];
                }
            }
        }
    }
}

template <typename D, int NCT, int NIT>
__device__ __forceinline__ void bf2_store_global(D *dst, const Geo &g, const f32x4 (&acc)[NCT][NIT])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            if (v >= nti) break;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = 16 * tc + 4 * lq + r, i = 16 * v + l16;
                if (c < g.nc && i < g.ni) dst[g.base + (long long)c * g.cs + i] = (D)acc[u][v][r];
            }
        }
    }
}

template <int PASS, int NPMAX>
__global__ void __launch_bounds__(kPT) fdm_bf2_kernel(const FdmTask *__restrict__ tasks)
{
    constexpr int NIT = NPMAX / 16, NCT = (NPMAX / 16 + 3) / 4;
    const FdmTask &t = tasks[blockIdx.y];
    Geo g;
    const int kind = PASS == 1 ? 1 : 0;
    const AxisMats &Mx = t.ax[0], &My = t.ax[1], &Mz = t.ax[2];
    const AxisMats &Mc = PASS == 1 ? Mz : My;
    if (!geo_of(kind, Mx.n, My.n, Mz.n, 1, 0, Mx.np, Mc.np, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pt = g.npc + 8, pc = g.npi + 8;
    uint16_t *Pth = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Ptl = Pth + g.npi * pt;
    uint16_t *Pch = Ptl + g.npi * pt;
    uint16_t *Pcl = Pch + g.npc * pc;
    f32x4 acc[NCT][NIT];
    if (PASS == 0) {
        if (t.src32) {
            const float *src = static_cast<const float *>(t.src);
            stage_bf2(src, g, vec_ok(src, g, 4), Pch, Pcl, nullptr, nullptr);
        } else {
            const double *src = static_cast<const double *>(t.src);
            stage_bf2(src, g, vec_ok(src, g, 2), Pch, Pcl, nullptr, nullptr);
        }
        __syncthreads();
        bf2_contract(Pch, Pcl, g, Mx.qth, Mx.qtl, false, acc);  // P Qx   (B-form: Qx^T)
        bf2_store_lds(Pth, Ptl, g, acc, true, nullptr);
        __syncthreads();
        bf2_contract(Pth, Ptl, g, My.qth, My.qtl, true, acc);   // Qy^T (.)
        bf2_store_global(t.work, g, acc);
    } else if (PASS == 1) {
        stage_bf2(t.work, g, vec_ok(t.work, g, 4), nullptr, nullptr, Pth, Ptl);
        __syncthreads();
        bf2_contract(Pth, Ptl, g, Mz.qth, Mz.qtl, true, acc);   // Qz^T P
        __syncthreads();
        bf2_store_lds(Pth, Ptl, g, acc, true, &t);               // scaled
        __syncthreads();
        bf2_contract(Pth, Ptl, g, Mz.qh, Mz.ql, true, acc);     // Qz (.)
        bf2_store_global(t.work, g, acc);
    } else {
        stage_bf2(t.work, g, vec_ok(t.work, g, 4), nullptr, nullptr, Pth, Ptl);
        __syncthreads();
        bf2_contract(Pth, Ptl, g, My.qh, My.ql, true, acc);     // Qy P
        bf2_store_lds(Pch, Pcl, g, acc, false, nullptr);
        __syncthreads();
        bf2_contract(Pch, Pcl, g, Mx.qh, Mx.ql, false, acc);    // (.) Qx^T  (B-form: Qx)
        if (t.dst32) bf2_store_global(static_cast<float *>(t.dst), g, acc);
        else bf2_store_global(static_cast<double *>(t.dst), g, acc);
    }
}

// ------------------------------------------------------------------ launch helpers
template <typename K>
void allow_lds(K kernel)
{
    static_cast<void>(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
}

size_t bf2_lds(int np) { return (size_t)2 * 2 * np * (np + 8) * sizeof(uint16_t); }
template <typename T> size_t f_lds(int np) { return (size_t)np * fpitch<T>(np) * sizeof(T); }

template <int NPMAX>
void init_np()
{
    static std::once_flag once;
    std::call_once(once, [] {
        allow_lds(hx_bf2_kernel<0, NPMAX>);
        allow_lds(hx_bf2_kernel<1, NPMAX>);
        allow_lds(fdm_f64_kernel<0, NPMAX>);
        allow_lds(fdm_f64_kernel<1, NPMAX>);
        allow_lds(fdm_f64_kernel<2, NPMAX>);
        allow_lds(fdm_bf2_kernel<0, NPMAX>);
        allow_lds(fdm_bf2_kernel<1, NPMAX>);
        allow_lds(fdm_bf2_kernel<2, NPMAX>);
    });
}

void init_f()
{
    static std::once_flag once;
    std::call_once(once, [] {
        allow_lds(hx_f_kernel<double, 0>);
        allow_lds(hx_f_kernel<double, 1>);
        allow_lds(hx_f_kernel<float, 0>);
        allow_lds(hx_f_kernel<float, 1>);
    });
}

template <int NPMAX>
void hx_bf2_launch(const HxJob *d, int n, int pa, int pb, int np, hipStream_t s)
{
    init_np<NPMAX>();
    // pass A uses only the i-major copy (first half of the layout)
    hipLaunchKernelGGL((hx_bf2_kernel<0, NPMAX>), dim3(pa, n), dim3(kPT), bf2_lds(np) / 2, s, d);
    hipLaunchKernelGGL((hx_bf2_kernel<1, NPMAX>), dim3(pb, n), dim3(kPT), bf2_lds(np), s, d);
}

template <int NPMAX>
void fdm_np_launch(const FdmTask *d, int n, int nplanes, int np, int precision, hipStream_t s)
{
    init_np<NPMAX>();
    if (precision == PISCES_FP64) {
        const size_t lds = f_lds<double>(np);
        hipLaunchKernelGGL((fdm_f64_kernel<0, NPMAX>), dim3(nplanes, n), dim3(kPT), lds, s, d);
        hipLaunchKernelGGL((fdm_f64_kernel<1, NPMAX>), dim3(nplanes, n), dim3(kPT), lds, s, d);
        hipLaunchKernelGGL((fdm_f64_kernel<2, NPMAX>), dim3(nplanes, n), dim3(kPT), lds, s, d);
    } else {
        const size_t lds = bf2_lds(np);
        hipLaunchKernelGGL((fdm_bf2_kernel<0, NPMAX>), dim3(nplanes, n), dim3(kPT), lds, s, d);
        // pass 1 uses only the i-major copy (first half of the layout)
        hipLaunchKernelGGL((fdm_bf2_kernel<1, NPMAX>), dim3(nplanes, n), dim3(kPT), lds / 2, s, d);
        hipLaunchKernelGGL((fdm_bf2_kernel<2, NPMAX>), dim3(nplanes, n), dim3(kPT), lds, s, d);
    }
}

}  // namespace

void hx_mfma_launch(const HxJob *d_jobs, int njobs, int max_planes_a, int max_planes_b, int max_n,
                    int precision, hipStream_t s)
{
    if (njobs <= 0) return;
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    const int np = axis_pad(max_n);
    if (precision == PISCES_BF16X2) {
        if (np <= 32) hx_bf2_launch<32>(d_jobs, njobs, max_planes_a, max_planes_b, np, s);
        else if (np <= 64) hx_bf2_launch<64>(d_jobs, njobs, max_planes_a, max_planes_b, np, s);
        else hx_bf2_launch<128>(d_jobs, njobs, max_planes_a, max_planes_b, np, s);
    } else if (precision == PISCES_FP32) {
        init_f();
        const size_t lds = f_lds<float>(np);
        hipLaunchKernelGGL((hx_f_kernel<float, 0>), dim3(max_planes_a, njobs), dim3(kPT), lds, s, d_jobs);
        hipLaunchKernelGGL((hx_f_kernel<float, 1>), dim3(max_planes_b, njobs), dim3(kPT), lds, s, d_jobs);
    } else {
        init_f();
        const size_t lds = f_lds<double>(np);
        hipLaunchKernelGGL((hx_f_kernel<double, 0>), dim3(max_planes_a, njobs), dim3(kPT), lds, s, d_jobs);
        hipLaunchKernelGGL((hx_f_kernel<double, 1>), dim3(max_planes_b, njobs), dim3(kPT), lds, s, d_jobs);
    }
    PC_HIP(hipGetLastError());
}

void fdm_launch(const FdmTask *d_tasks, int ntasks, int max_n, int precision, hipStream_t s)
{
    if (ntasks <= 0) return;
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    const int np = axis_pad(max_n);
    // planes per pass: the largest of ny (pass 1) and nz (passes 0, 2) <= max_n
    if (np <= 32) fdm_np_launch<32>(d_tasks, ntasks, max_n, np, precision, s);
    else if (np <= 64) fdm_np_launch<64>(d_tasks, ntasks, max_n, np, precision, s);
    else fdm_np_launch<128>(d_tasks, ntasks, max_n, np, precision, s);
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
