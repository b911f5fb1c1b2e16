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
#include "gmem.hpp"

#include <algorithm>
#include <cstdlib>
#include <mutex>

namespace pisces {

namespace {

constexpr int kPT = 256;

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
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
__device__ __forceinline__ bf16x8 ld8(gptr<const uint16_t> p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<gptr<const u32x4>>(p));
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
__device__ __forceinline__ void load_item(const S *src_, const Geo &g, int c, int i, bool vec,
                                          float (&x)[2][4])
{
    const gptr<const S> src = glb(src_) + g.base;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int cc = c + r;
        const bool okc = cc < g.nc;
        const gptr<const S> row = src + (okc ? (long long)cc * g.cs : 0);
        if (vec && i + 4 <= g.ni) {
            float y[4];
            if (sizeof(S) == 4) {
                const f32x4 v = *reinterpret_cast<gptr<const f32x4>>(row + i);
                y[0] = v[0]; y[1] = v[1]; y[2] = v[2]; y[3] = v[3];
            } else {
                const f64x2 a = *reinterpret_cast<gptr<const f64x2>>(row + i);
                const f64x2 b = *reinterpret_cast<gptr<const f64x2>>(row + i + 2);
                y[0] = (float)a[0]; y[1] = (float)a[1]; y[2] = (float)b[0]; y[3] = (float)b[1];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) x[r][q] = okc ? y[q] : 0.f;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool ok = okc && i + q < g.ni;
                const float y = (float)row[ok ? i + q : 0];
                x[r][q] = ok ? y : 0.f;
            }
        }
    }
}

// FP32/FP64 plane into LDS, c-major, pitch npi + pad.
template <typename T, typename S>
__device__ void stage_f(const S *src, const Geo &g, bool vec, T *P, int pitch)
{
    const int nq = g.npi / 4, items = (g.npc / 2) * nq;
    for (int it = threadIdx.x; it < items; it += kPT) {
        const int c = 2 * (it / nq), i = 4 * (it - (it / nq) * nq);
        if (sizeof(S) == 8 && sizeof(T) == 8) {
            const gptr<const S> sg = glb(src) + g.base;
            T y[2][4];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int cc = c + r;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool ok = cc < g.nc && i + q < g.ni;
                    const T v = (T)sg[ok ? (long long)cc * g.cs + i + q : 0];
                    y[r][q] = ok ? v : T(0);
                }
            }
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int q = 0; q < 4; ++q) P[(c + r) * pitch + i + q] = y[r][q];
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

// ------------------------------------------------------------------ owned tiles (split-BF16 forms)
// Wave w owns the 16-row tiles tc = w + 4u (u < NCT) of a plane and all of its
// 16-column tiles v < NIT; in each, lane (l16, lq) holds rows 4 lq + r (r < 4)
// of column l16 -- the accumulator layout of the f32-accumulating MFMAs.
// Global loads, LDS staging and global stores follow this ownership, so a
// wave combines its results with values it loaded itself, and a plane's rows
// move HBM -> registers -> LDS without a layout change.  A workgroup walks the
// planes p = blockIdx.x, blockIdx.x + gridDim.x, ... of one job; the loads of
// the next plane are issued before the current one is contracted.
// The middle pass of H*X (pass B) and of the preconditioner (pass 1) reads
// what the pass before wrote: it walks its grid in reverse (workgroups are
// dispatched in order of their linear index), starting on the data the
// previous pass left last in the L2 and the memory-side cache.
__device__ __forceinline__ int plane_x(bool rev) { return rev ? gridDim.x - 1 - blockIdx.x : blockIdx.x; }
__device__ __forceinline__ int plane_y(bool rev) { return rev ? gridDim.y - 1 - blockIdx.y : blockIdx.y; }

template <int NPMAX> struct Own {
    static constexpr int NIT = NPMAX / 16;            // column tiles
    static constexpr int NCT = (NPMAX / 16 + 3) / 4;  // row tiles per wave
    static constexpr int KS = NPMAX / 32;             // k-steps of the bf16 MFMA
};

// Owned plane values as FP32 (zero outside the n_c x n_i plane).  Every lane
// loads unconditionally -- out-of-plane elements from the plane's (0, 0) --
// so the loads issue back to back, and the value is selected afterwards.
template <typename S, int NCT, int NIT>
__device__ __forceinline__ void own_load(const S *src_, long long base, const Geo &g,
                                         float (&x)[NCT][NIT][4])
{
    const gptr<const S> src = glb(src_) + base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int u = 0; u < NCT; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = 16 * (wave + 4 * u) + 4 * lq + r;
#pragma unroll
            for (int v = 0; v < NIT; ++v) {
                const int i = 16 * v + l16;
                const bool ok = c < g.nc && i < g.ni;
                const float y = (float)src[ok ? (long long)c * g.cs + i : 0];
                x[u][v][r] = ok ? y : 0.f;
            }
        }
}

// Exchange a dword with the neighbouring lane (DPP quad_perm 1,0,3,2).
__device__ __forceinline__ unsigned swap_adj(unsigned x)
{
    return (unsigned)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
}

// Split one owned tile column (rows cb .. cb+3 of column i) into the LDS
// plane: i-major (Pt, pitch pt) and/or c-major (Pc, pitch pc); a null pointer
// skips a layout.  For the c-major rows the lanes of a column pair (i even,
// i + 1) meet in one dword: the even lane writes rows cb, cb+1, the odd lane
// rows cb+2, cb+3.  Every lane of the wave must take part.
__device__ __forceinline__ void put_tile(const float (&x)[4], int cb, int i, int pt, int pc,
                                         uint16_t *Pth, uint16_t *Ptl, uint16_t *Pch, uint16_t *Pcl)
{
    uint16_t h[4], l[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) split2(x[r], h[r], l[r]);
    const unsigned h01 = h[0] | ((unsigned)h[1] << 16), h23 = h[2] | ((unsigned)h[3] << 16);
    const unsigned l01 = l[0] | ((unsigned)l[1] << 16), l23 = l[2] | ((unsigned)l[3] << 16);
    if (Pth) {
        u32x2 vh, vl;
        vh[0] = h01; vh[1] = h23;
        vl[0] = l01; vl[1] = l23;
        *reinterpret_cast<u32x2 *>(Pth + i * pt + cb) = vh;
        *reinterpret_cast<u32x2 *>(Ptl + i * pt + cb) = vl;
    }
    if (Pch) {
        const bool odd = i & 1;
        const unsigned hr = swap_adj(odd ? h01 : h23), lr = swap_adj(odd ? l01 : l23);
        const unsigned hm = odd ? h23 : h01, lm = odd ? l23 : l01;
        const unsigned hlo = odd ? hr : hm, hhi = odd ? hm : hr;  // columns i & ~1, i | 1
        const unsigned llo = odd ? lr : lm, lhi = odd ? lm : lr;
        const int row = cb + (odd ? 2 : 0), col = i & ~1;
        *reinterpret_cast<unsigned *>(Pch + row * pc + col) = (hlo & 0xffffu) | (hhi << 16);
        *reinterpret_cast<unsigned *>(Pch + (row + 1) * pc + col) = (hlo >> 16) | (hhi & 0xffff0000u);
        *reinterpret_cast<unsigned *>(Pcl + row * pc + col) = (llo & 0xffffu) | (lhi << 16);
        *reinterpret_cast<unsigned *>(Pcl + (row + 1) * pc + col) = (llo >> 16) | (lhi & 0xffff0000u);
    }
}

template <int NCT, int NIT>
__device__ __forceinline__ void own_stage(const float (&x)[NCT][NIT][4], const Geo &g,
                                          uint16_t *Pth, uint16_t *Ptl, uint16_t *Pch, uint16_t *Pcl)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16, pt = g.npc + 8, pc = g.npi + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            if (v >= nti) break;
            put_tile(x[u][v], 16 * tc + 4 * lq, 16 * v + l16, pt, pc, Pth, Ptl, Pch, Pcl);
        }
    }
}

__device__ __forceinline__ bf16x8 bf_zero()
{
    const u32x4 z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, z);
}

// A-form fragments of a row-major split matrix (leading dimension ld) for the
// wave's own tile rows: rows 16 tc + l16, columns 32 s + 8 lq .. + 7.
template <int NCT, int KS>
__device__ __forceinline__ void afrag(const uint16_t *Mh, const uint16_t *Ml, int ld, int ntc, int ks,
                                      bf16x8 (&ah)[NCT][KS], bf16x8 (&al)[NCT][KS])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int u = 0; u < NCT; ++u)
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int tc = wave + 4 * u;
            if (tc < ntc && s < ks) {
                ah[u][s] = ld8(glb(Mh) + (16 * tc + l16) * ld + 32 * s + 8 * lq);
                al[u][s] = ld8(glb(Ml) + (16 * tc + l16) * ld + 32 * s + 8 * lq);
            } else {
                ah[u][s] = bf_zero();
                al[u][s] = bf_zero();
            }
        }
}

// D(c, i) += sum_k A(c, k) B(k, i) on the column-tile pair (i0, i0 + 16) of
// one row tile, split-BF16 (hi*hi + hi*lo + lo*hi).  A: held fragments.
// B(k, i) = Bm[i][k] with Bm row-major of pitch ld: the i-major LDS plane for
// a contraction along c, or the B-form matrix for one along i.
template <int KS, typename BP>
__device__ __forceinline__ void mm_pair(const bf16x8 (&ah)[KS], const bf16x8 (&al)[KS], BP Bh, BP Bl,
                                        int ld, int ks, int i0, f32x4 &a0, f32x4 &a1)
{
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
    const int o0 = (i0 + l16) * ld + 8 * lq, o1 = o0 + 16 * ld;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s >= ks) break;
        const bf16x8 b0h = ld8(Bh + o0 + 32 * s), b0l = ld8(Bl + o0 + 32 * s);
        const bf16x8 b1h = ld8(Bh + o1 + 32 * s), b1l = ld8(Bl + o1 + 32 * s);
        a0 = mma_bf(ah[s], b0h, a0);
        a1 = mma_bf(ah[s], b1h, a1);
        a0 = mma_bf(ah[s], b0l, a0);
        a1 = mma_bf(ah[s], b1l, a1);
        a0 = mma_bf(al[s], b0h, a0);
        a1 = mma_bf(al[s], b1h, a1);
    }
}

// B-form fragments of a row-major split matrix (leading dimension ld) for all
// column tiles: rows 16 v + l16, columns 32 s + 8 lq .. + 7.  Held in
// registers for a workgroup's lifetime where they fit (grids up to 64).
template <int NIT, int KS>
__device__ __forceinline__ void bfrag(const uint16_t *Mh, const uint16_t *Ml, int ld, int nti, int ks,
                                      bf16x8 (&bh)[NIT][KS], bf16x8 (&bl)[NIT][KS])
{
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int v = 0; v < NIT; ++v)
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if (v < nti && s < ks) {
                bh[v][s] = ld8(glb(Mh) + (16 * v + l16) * ld + 32 * s + 8 * lq);
                bl[v][s] = ld8(glb(Ml) + (16 * v + l16) * ld + 32 * s + 8 * lq);
            } else {
                bh[v][s] = bf_zero();
                bl[v][s] = bf_zero();
            }
        }
}

// mm_pair with both operands held in registers.
template <int KS>
__device__ __forceinline__ void mm_pair_held(const bf16x8 (&ah)[KS], const bf16x8 (&al)[KS],
                                             const bf16x8 (&b0h)[KS], const bf16x8 (&b0l)[KS],
                                             const bf16x8 (&b1h)[KS], const bf16x8 (&b1l)[KS], int ks,
                                             f32x4 &a0, f32x4 &a1)
{
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s >= ks) break;
        a0 = mma_bf(ah[s], b0h[s], a0);
        a1 = mma_bf(ah[s], b1h[s], a1);
        a0 = mma_bf(ah[s], b0l[s], a0);
        a1 = mma_bf(ah[s], b1l[s], a1);
        a0 = mma_bf(al[s], b0h[s], a0);
        a1 = mma_bf(al[s], b1h[s], a1);
    }
}

// A-form fragments of the wave's own rows of the c-major LDS plane.
template <int KS>
__device__ __forceinline__ void afrag_lds(const uint16_t *Pch, const uint16_t *Pcl, int pc, int c0, int ks,
                                          bf16x8 (&ah)[KS], bf16x8 (&al)[KS])
{
    const int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s >= ks) break;
        ah[s] = ld8(Pch + (c0 + l16) * pc + 32 * s + 8 * lq);
        al[s] = ld8(Pcl + (c0 + l16) * pc + 32 * s + 8 * lq);
    }
}

// ------------------------------------------------------------------ H*X, split-BF16 form
// Pass A (plane (i,j) at k): Y = V o X + Ty X along c = j; the accumulators
// start from V o X.  Pass B (plane (i,k) at j): Y += Tz X along c = k + X Tx
// along i; the wave's Y values of the next plane are loaded as soon as those
// of the current plane have been used.  LDS: the i-major plane (pass A), plus
// the c-major plane (pass B).
template <int PASS, int NPMAX, bool MULTI>
__global__ void __launch_bounds__(kPT) hx_bf2_kernel(const HxJob *__restrict__ jobs)
{
    constexpr int NCT = Own<NPMAX>::NCT, NIT = Own<NPMAX>::NIT, KS = Own<NPMAX>::KS;
    const HxJob &J = jobs[plane_y(PASS == 1)];
    const AxisMats &Mi = J.ax[0];
    const AxisMats &Mc = J.ax[PASS == 0 ? 1 : 2];
    int p = plane_x(PASS == 1);
    Geo g;
    if (!geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, p, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pt = g.npc + 8, pc = g.npi + 8;
    uint16_t *Pth = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Ptl = Pth + g.npi * pt;
    uint16_t *Pch = Ptl + g.npi * pt;
    uint16_t *Pcl = Pch + g.npc * pc;
    const float *X = static_cast<const float *>(J.X);
    const float *V = static_cast<const float *>(J.V);
    const gptr<float> Y = glb(static_cast<float *>(J.Y));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16, ksc = g.npc / 32, ksi = g.npi / 32;
    bf16x8 ach[NCT][KS], acl[NCT][KS];
    afrag(Mc.th, Mc.tl, g.npc, ntc, ksc, ach, acl);
    // pass B, grids up to 64: Tx in B form held for the workgroup's lifetime
    constexpr bool HOLD = PASS == 1 && NPMAX <= 64;
    constexpr int NB = HOLD ? NIT : 1, KB = HOLD ? KS : 1;
    bf16x8 bth[NB][KB], btl[NB][KB];
    if constexpr (HOLD) bfrag(Mi.th, Mi.tl, g.npi, nti, ksi, bth, btl);
    float xb[NCT][NIT][4], wb[NCT][NIT][4];  // X; V (pass A) or Y (pass B)
    own_load(X, g.base, g, xb);
    own_load(PASS == 0 ? V : static_cast<const float *>(J.Y), PASS == 0 ? g.vbase : g.base, g, wb);
    for (;;) {
        own_stage(xb, g, Pth, Ptl, PASS == 1 ? Pch : nullptr, Pcl);
        if (PASS == 0) {
#pragma unroll
            for (int u = 0; u < NCT; ++u)
#pragma unroll
                for (int v = 0; v < NIT; ++v)
#pragma unroll
                    for (int r = 0; r < 4; ++r) wb[u][v][r] *= xb[u][v][r];
        }
        __syncthreads();
        Geo gn;
        const bool more = MULTI && geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, p + gridDim.x, gn);
        if (more) own_load(X, gn.base, gn, xb);
#pragma unroll
        for (int u = 0; u < NCT; ++u) {
            const int tc = wave + 4 * u;
            if (tc >= ntc) break;
            const int c0 = 16 * tc;
            bf16x8 aih[KS], ail[KS];
            if (PASS == 1) afrag_lds(Pch, Pcl, pc, c0, ksi, aih, ail);
#pragma unroll
            for (int v = 0; v < NIT; v += 2) {
                if (v >= nti) break;
                f32x4 a0, a1;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    a0[r] = PASS == 0 ? wb[u][v][r] : 0.f;
                    a1[r] = PASS == 0 ? wb[u][v + 1][r] : 0.f;
                }
                mm_pair(ach[u], acl[u], Pth, Ptl, pt, ksc, 16 * v, a0, a1);
                // B(k, i) = Tx[k][i] = Tx[i][k] (symmetric)
                if constexpr (HOLD)
                    mm_pair_held(aih, ail, bth[v], btl[v], bth[v + 1], btl[v + 1], ksi, a0, a1);
                else if (PASS == 1)
                    mm_pair(aih, ail, glb(Mi.th), glb(Mi.tl), g.npi, ksi, 16 * v, a0, a1);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = c0 + 4 * lq + r;
                    if (c >= g.nc) continue;
                    const long long row = (long long)c * g.cs;
                    const int ia = 16 * v + l16, ib = ia + 16;
                    if (ia < g.ni) Y[g.base + row + ia] = PASS == 0 ? a0[r] : wb[u][v][r] + a0[r];
                    if (ib < g.ni) Y[g.base + row + ib] = PASS == 0 ? a1[r] : wb[u][v + 1][r] + a1[r];
                    if (more) {
                        const gptr<const float> W =
                            PASS == 0 ? glb(V) + gn.vbase + row : (gptr<const float>)Y + gn.base + row;
                        const float wa = W[ia < g.ni ? ia : 0], wv = W[ib < g.ni ? ib : 0];
                        wb[u][v][r] = ia < g.ni ? wa : 0.f;
                        wb[u][v + 1][r] = ib < g.ni ? wv : 0.f;
                    }
                }
            }
        }
        if (!more) break;
        __syncthreads();  // every wave is done with this plane's LDS copy
        g = gn;
        p += gridDim.x;
    }
}

// ------------------------------------------------------------------ H*X, FP64 form
// Owned tiles as above, in the f64 accumulator layout: lane (l16, lq) holds
// rows lq + 4 r (r < 4) of column l16 in each of its wave's tiles.  The plane
// is kept in LDS as FP64, c-major (pitch npi + 2 doubles); the along-c matrix
// (A operand, rows of the wave's tiles) is held in registers, and in pass B
// the along-i matrix Tx (B operand) is staged once per workgroup into LDS
// next to the plane.  Each k-step feeds one A value to the MFMAs of all the
// wave's column tiles (independent accumulator chains).
template <int NCT, int NIT>
__device__ __forceinline__ void own_load64(const double *src_, long long base, const Geo &g,
                                           double (&x)[NCT][NIT][4])
{
    const gptr<const double> src = glb(src_) + base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int u = 0; u < NCT; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = 16 * (wave + 4 * u) + lq + 4 * r;
#pragma unroll
            for (int v = 0; v < NIT; ++v) {
                const int i = 16 * v + l16;
                const bool ok = c < g.nc && i < g.ni;
                const double y = src[ok ? (long long)c * g.cs + i : 0];
                x[u][v][r] = ok ? y : 0.0;
            }
        }
}

template <int PASS, int NPMAX, bool MULTI>
__global__ void __launch_bounds__(kPT) hx_f64_kernel(const HxJob *__restrict__ jobs)
{
    constexpr int NCT = Own<NPMAX>::NCT, NIT = Own<NPMAX>::NIT, KF = NPMAX / 4;
    const HxJob &J = jobs[plane_y(PASS == 1)];
    const AxisMats &Mi = J.ax[0];
    const AxisMats &Mc = J.ax[PASS == 0 ? 1 : 2];
    int p = plane_x(PASS == 1);
    Geo g;
    if (!geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, p, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pitch = g.npi + 2;
    double *P = reinterpret_cast<double *>(smem);
    double *Tl = P + g.npc * pitch;  // pass B: Tx, rows k, pitch npi + 2
    const double *X = static_cast<const double *>(J.X);
    const double *V = static_cast<const double *>(J.V);
    const gptr<double> Y = glb(static_cast<double *>(J.Y));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16, ksc = g.npc / 4, ksi = g.npi / 4;
    // along-c matrix in A form: ac[u][s] = Tc[16 tc + l16][4 s + lq]
    double ac[NCT][KF];
    {
        const gptr<const double> Tc = glb(Mc.t64);
#pragma unroll
        for (int u = 0; u < NCT; ++u)
#pragma unroll
            for (int s = 0; s < KF; ++s) {
                const int tc = wave + 4 * u;
                const bool ok = tc < ntc && s < ksc;
                const double y = Tc[ok ? (16 * tc + l16) * g.npc + 4 * s + lq : 0];
                ac[u][s] = ok ? y : 0.0;
            }
    }
    if (PASS == 1) {
        const gptr<const double> Ti = glb(Mi.t64);
        for (int e = threadIdx.x; e < g.npi * g.npi; e += kPT) {
            const int k = e / g.npi, i = e - k * g.npi;
            Tl[k * pitch + i] = Ti[e];
        }
    }
    double xb[NCT][NIT][4], wb[NCT][NIT][4];  // X; V (pass A) or Y (pass B)
    own_load64<NCT, NIT>(X, g.base, g, xb);
    own_load64<NCT, NIT>(PASS == 0 ? V : static_cast<const double *>(J.Y), PASS == 0 ? g.vbase : g.base, g, wb);
    for (;;) {
#pragma unroll
        for (int u = 0; u < NCT; ++u) {
            const int tc = wave + 4 * u;
            if (tc >= ntc) break;
#pragma unroll
            for (int v = 0; v < NIT; ++v) {
                if (v >= nti) break;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    P[(16 * tc + lq + 4 * r) * pitch + 16 * v + l16] = xb[u][v][r];
                    if (PASS == 0) wb[u][v][r] *= xb[u][v][r];
                }
            }
        }
        __syncthreads();
        Geo gn;
        const bool more = MULTI && geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, p + gridDim.x, gn);
        if (more) own_load64<NCT, NIT>(X, gn.base, gn, xb);
#pragma unroll
        for (int u = 0; u < NCT; ++u) {
            const int tc = wave + 4 * u;
            if (tc >= ntc) break;
            const int c0 = 16 * tc;
            f64x4 acc[NIT];
#pragma unroll
            for (int v = 0; v < NIT; ++v)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[v][r] = PASS == 0 ? wb[u][v][r] : 0.0;
            // along c: D(c, i) += sum_k Tc[c][k] P[k][i]
#pragma unroll
            for (int s = 0; s < KF; ++s) {
                if (s >= ksc) break;
                const double *pr = P + (4 * s + lq) * pitch + l16;
#pragma unroll
                for (int v = 0; v < NIT; ++v) {
                    if (v >= nti) break;
                    acc[v] = mma(ac[u][s], pr[16 * v], acc[v]);
                }
            }
            // along i: D(c, i) += sum_k P[c][k] Tx[k][i]
            if (PASS == 1) {
#pragma unroll
                for (int s = 0; s < KF; ++s) {
                    if (s >= ksi) break;
                    const double a = P[(c0 + l16) * pitch + 4 * s + lq];
                    const double *tr = Tl + (4 * s + lq) * pitch + l16;
#pragma unroll
                    for (int v = 0; v < NIT; ++v) {
                        if (v >= nti) break;
                        acc[v] = mma(a, tr[16 * v], acc[v]);
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < NIT; ++v) {
                if (v >= nti) break;
                const int i = 16 * v + l16;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = c0 + lq + 4 * r;
                    if (c >= g.nc || i >= g.ni) continue;
                    const long long row = (long long)c * g.cs;
                    Y[g.base + row + i] = PASS == 0 ? acc[v][r] : wb[u][v][r] + acc[v][r];
                    if (more) {
                        const gptr<const double> W =
                            PASS == 0 ? glb(V) + gn.vbase + row : (gptr<const double>)Y + gn.base + row;
                        wb[u][v][r] = W[i];
                    }
                }
            }
        }
        if (!more) break;
        __syncthreads();  // every wave is done with this plane's LDS copy
        g = gn;
        p += gridDim.x;
    }
}

// ------------------------------------------------------------------ H*X, FP32 form

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
    const gptr<T> Y = glb(static_cast<T *>(J.Y));
    stage_f<T, T>(X, g, vec_ok(X, g, 16 / (int)sizeof(T)), P, pitch);
    __syncthreads();
    const gptr<const T> V = glb(static_cast<const T *>(J.V));
    const gptr<const T> Tc = glb(tmat<T>(Mc)), Ti = glb(tmat<T>(Mi));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16;
    for (int tc = wave; tc < ntc; tc += 4) {
        const int c0 = 16 * tc;
        for (int ti = 0; ti < nti; ti += 2) {
            const int i0 = 16 * ti, i1 = i0 + 16;
            acc_t a0 = {}, a1 = {};
            // along c: A[m][k] = Tc[c0 + m][k], B[k][n] = P[k][i + n]
            const gptr<const T> tcr = Tc + (c0 + l16) * g.npc + lq;
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
                    const gptr<const T> tr = Ti + (k0 + lq) * g.npi + l16;
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
                                             const double *M_, bool along_c, f64x4 (&acc)[NCT][NIT])
{
    const gptr<const double> M = glb(M_);
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
                const gptr<const double> mr = M + (c0 + l16) * g.npc + lq;
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
                    const gptr<const double> mr = M + (k0 + lq) * g.npi + l16;
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
__device__ __forceinline__ void f64_store_global(D *dst_, const Geo &g, const f64x4 (&acc)[NCT][NIT])
{
    const gptr<D> dst = glb(dst_);
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
// Owned tiles scaled by scale / (lam_x[i] + ly + lam_z[c]), ly = lam_y[j] -
// shift of plane (i = a, c = z index) at fixed j, then split into the i-major
// plane.  lx[v] = lam_x[16 v + l16] and lz[u][r] = lam_z[c] are held per
// workgroup; the factor is formed in FP32 (the scaled values are FP32 and go
// on as split BF16).
template <int NCT, int NIT>
__device__ __forceinline__ void put_scaled(const f32x4 (&acc)[NCT][NIT], const Geo &g, float ly,
                                           const float (&lx)[NIT], const float (&lz)[NCT][4],
                                           float scale, float floor, uint16_t *Pth, uint16_t *Ptl)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16, pt = g.npc + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            if (v >= nti) break;
            const int i = 16 * v + l16, cb = 16 * tc + 4 * lq;
            float x[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float den = lx[v] + ly + lz[u][r];
                if (fabsf(den) < floor) den = den < 0.f ? -floor : floor;
                x[r] = acc[u][v][r] * (scale / den);
            }
            put_tile(x, cb, i, pt, 0, Pth, Ptl, nullptr, nullptr);
        }
    }
}

template <int NCT, int NIT>
__device__ __forceinline__ void put_acc(const f32x4 (&acc)[NCT][NIT], const Geo &g, uint16_t *Pth,
                                        uint16_t *Ptl, uint16_t *Pch, uint16_t *Pcl)
{
    float x[NCT][NIT][4];
#pragma unroll
    for (int u = 0; u < NCT; ++u)
#pragma unroll
        for (int v = 0; v < NIT; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) x[u][v][r] = acc[u][v][r];
    own_stage(x, g, Pth, Ptl, Pch, Pcl);
}

template <typename D, int NCT, int NIT>
__device__ __forceinline__ void bf2_store_global(D *dst_, const Geo &g, const f32x4 (&acc)[NCT][NIT])
{
    const gptr<D> dst = glb(dst_);
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

// All owned tiles of one contraction: along c (A = held fragments of the
// along-c matrix, B = i-major plane) or along i (A = own rows of the c-major
// plane, B = B-form matrix from global memory).
template <int NCT, int NIT, int KS>
__device__ __forceinline__ void mm_c_all(const bf16x8 (&ah)[NCT][KS], const bf16x8 (&al)[NCT][KS],
                                         const uint16_t *Pth, const uint16_t *Ptl, const Geo &g,
                                         f32x4 (&acc)[NCT][NIT])
{
    const int wave = threadIdx.x >> 6;
    const int ntc = g.npc / 16, nti = g.npi / 16, ksc = g.npc / 32, pt = g.npc + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        if (wave + 4 * u >= ntc) break;
#pragma unroll
        for (int v = 0; v < NIT; v += 2) {
            if (v >= nti) break;
            f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
            mm_pair(ah[u], al[u], Pth, Ptl, pt, ksc, 16 * v, a0, a1);
            acc[u][v] = a0;
            acc[u][v + 1] = a1;
        }
    }
}

template <int NCT, int NIT, int KS>
__device__ __forceinline__ void mm_i_held(const uint16_t *Pch, const uint16_t *Pcl,
                                          const bf16x8 (&bh)[NIT][KS], const bf16x8 (&bl)[NIT][KS],
                                          const Geo &g, f32x4 (&acc)[NCT][NIT])
{
    const int wave = threadIdx.x >> 6;
    const int ntc = g.npc / 16, nti = g.npi / 16, ksi = g.npi / 32, pc = g.npi + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
        bf16x8 ah[KS], al[KS];
        afrag_lds(Pch, Pcl, pc, 16 * tc, ksi, ah, al);
#pragma unroll
        for (int v = 0; v < NIT; v += 2) {
            if (v >= nti) break;
            f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
            mm_pair_held(ah, al, bh[v], bl[v], bh[v + 1], bl[v + 1], ksi, a0, a1);
            acc[u][v] = a0;
            acc[u][v + 1] = a1;
        }
    }
}

template <int NCT, int NIT, int KS>
__device__ __forceinline__ void mm_i_all(const uint16_t *Pch, const uint16_t *Pcl, const uint16_t *Mh,
                                         const uint16_t *Ml, const Geo &g, f32x4 (&acc)[NCT][NIT])
{
    const int wave = threadIdx.x >> 6;
    const int ntc = g.npc / 16, nti = g.npi / 16, ksi = g.npi / 32, pc = g.npi + 8;
#pragma unroll
    for (int u = 0; u < NCT; ++u) {
        const int tc = wave + 4 * u;
        if (tc >= ntc) break;
        bf16x8 ah[KS], al[KS];
        afrag_lds(Pch, Pcl, pc, 16 * tc, ksi, ah, al);
#pragma unroll
        for (int v = 0; v < NIT; v += 2) {
            if (v >= nti) break;
            f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
            mm_pair(ah, al, glb(Mh), glb(Ml), g.npi, ksi, 16 * v, a0, a1);
            acc[u][v] = a0;
            acc[u][v + 1] = a1;
        }
    }
}

// Three passes, each walking planes like the H*X kernels:
//   0, plane (i,j) at k: r -> c-major plane; X Qx along i -> i-major plane;
//      Qy^T (.) along c -> work
//   1, plane (i,k) at j: work -> i-major plane A; Qz^T (.) along c, scaled ->
//      i-major plane B; Qz (.) along c -> work
//   2, plane (i,j) at k: work -> i-major plane; Qy (.) along c -> c-major
//      plane; (.) Qx^T along i -> t
// The c-major plane is only read by the wave that wrote its rows.
template <int PASS, int NPMAX, bool MULTI>
__global__ void __launch_bounds__(kPT) fdm_bf2_kernel(const FdmTask *__restrict__ tasks)
{
    constexpr int NCT = Own<NPMAX>::NCT, NIT = Own<NPMAX>::NIT, KS = Own<NPMAX>::KS;
    const FdmTask &t = tasks[plane_y(PASS == 1)];
    const int kind = PASS == 1 ? 1 : 0;
    const AxisMats &Mx = t.ax[0], &My = t.ax[1], &Mz = t.ax[2];
    const AxisMats &Mc = PASS == 1 ? Mz : My;
    int p = plane_x(PASS == 1);
    Geo g;
    if (!geo_of(kind, Mx.n, My.n, Mz.n, 1, 0, Mx.np, Mc.np, p, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const int pt = g.npc + 8, pc = g.npi + 8;
    uint16_t *Pth = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Ptl = Pth + g.npi * pt;
    uint16_t *P2h = Ptl + g.npi * pt;  // c-major plane (passes 0, 2) or i-major plane B (pass 1)
    uint16_t *P2l = P2h + (PASS == 1 ? g.npi * pt : g.npc * pc);
    const int ntc = g.npc / 16, ksc = g.npc / 32;
    bf16x8 ah[NCT][KS], al[NCT][KS], bh[NCT][KS], bl[NCT][KS];
    if (PASS == 0) afrag(My.qth, My.qtl, g.npc, ntc, ksc, ah, al);       // Qy^T
    if (PASS == 1) {
        afrag(Mz.qth, Mz.qtl, g.npc, ntc, ksc, ah, al);                  // Qz^T
        afrag(Mz.qh, Mz.ql, g.npc, ntc, ksc, bh, bl);                    // Qz
    }
    if (PASS == 2) afrag(My.qh, My.ql, g.npc, ntc, ksc, ah, al);         // Qy
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int nti = g.npi / 16, ksi = g.npi / 32;
    // passes 0 and 2, grids up to 64: Qx in B form held for the workgroup's lifetime
    constexpr bool HOLD = PASS != 1 && NPMAX <= 64;
    constexpr int NB = HOLD ? NIT : 1, KB = HOLD ? KS : 1;
    bf16x8 qbh[NB][KB], qbl[NB][KB];
    if constexpr (HOLD) {
        if (PASS == 0) bfrag(Mx.qth, Mx.qtl, g.npi, nti, ksi, qbh, qbl);  // B form of Qx
        else bfrag(Mx.qh, Mx.ql, g.npi, nti, ksi, qbh, qbl);              // B form of Qx^T
    }
    // pass 1: eigenvalues of the owned rows and columns
    float lx[NIT], lz[NCT][4];
    if (PASS == 1) {
        const gptr<const double> Lx = glb(Mx.lam), Lz = glb(Mz.lam);
#pragma unroll
        for (int v = 0; v < NIT; ++v) {
            const int i = 16 * v + l16;
            const double y = Lx[i < g.ni ? i : 0];
            lx[v] = i < g.ni ? (float)y : 0.f;
        }
#pragma unroll
        for (int u = 0; u < NCT; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int c = 16 * (wave + 4 * u) + 4 * lq + r;
                const double y = Lz[c < g.nc ? c : 0];
                lz[u][r] = c < g.nc ? (float)y : 0.f;
            }
    }
    const void *src = PASS == 0 ? t.src : t.work;
    const bool s32 = PASS != 0 || t.src32;
    float xb[NCT][NIT][4];
    if (s32) own_load(static_cast<const float *>(src), g.base, g, xb);
    else own_load(static_cast<const double *>(src), g.base, g, xb);
    f32x4 acc[NCT][NIT];
    for (;;) {
        own_stage(xb, g, PASS == 0 ? nullptr : Pth, Ptl, PASS == 0 ? P2h : nullptr, P2l);
        __syncthreads();
        Geo gn;
        const bool more = MULTI && geo_of(kind, Mx.n, My.n, Mz.n, 1, 0, Mx.np, Mc.np, p + gridDim.x, gn);
        if (more) {
            if (s32) own_load(static_cast<const float *>(src), gn.base, gn, xb);
            else own_load(static_cast<const double *>(src), gn.base, gn, xb);
        }
        if (PASS == 0) {
            if constexpr (HOLD) mm_i_held<NCT, NIT, KS>(P2h, P2l, qbh, qbl, g, acc);  // X Qx
            else mm_i_all<NCT, NIT, KS>(P2h, P2l, Mx.qth, Mx.qtl, g, acc);
            put_acc(acc, g, Pth, Ptl, nullptr, nullptr);
            __syncthreads();
            mm_c_all(ah, al, Pth, Ptl, g, acc);                            // Qy^T (.)
            bf2_store_global(t.work, g, acc);
        } else if (PASS == 1) {
            mm_c_all(ah, al, Pth, Ptl, g, acc);                            // Qz^T X
            // rows of the scaled plane: i = a, c = z index, at j = g.pl
            put_scaled(acc, g, (float)(glb(My.lam)[g.pl] - t.shift), lx, lz, (float)t.scale,
                       (float)t.floor, P2h, P2l);
            __syncthreads();
            mm_c_all(bh, bl, P2h, P2l, g, acc);                            // Qz (.)
            bf2_store_global(t.work, g, acc);
        } else {
            mm_c_all(ah, al, Pth, Ptl, g, acc);                            // Qy X
            put_acc(acc, g, nullptr, nullptr, P2h, P2l);
            __syncthreads();
            if constexpr (HOLD) mm_i_held<NCT, NIT, KS>(P2h, P2l, qbh, qbl, g, acc);  // (.) Qx^T
            else mm_i_all<NCT, NIT, KS>(P2h, P2l, Mx.qh, Mx.ql, g, acc);
            if (t.dst32) bf2_store_global(static_cast<float *>(t.dst), g, acc);
            else bf2_store_global(static_cast<double *>(t.dst), g, acc);
        }
        if (!more) break;
        g = gn;
        p += gridDim.x;
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
        allow_lds(hx_bf2_kernel<0, NPMAX, false>);
        allow_lds(hx_bf2_kernel<1, NPMAX, false>);
        allow_lds(hx_bf2_kernel<0, NPMAX, true>);
        allow_lds(hx_bf2_kernel<1, NPMAX, true>);
        allow_lds(fdm_f64_kernel<0, NPMAX>);
        allow_lds(fdm_f64_kernel<1, NPMAX>);
        allow_lds(fdm_f64_kernel<2, NPMAX>);
        allow_lds(fdm_bf2_kernel<0, NPMAX, false>);
        allow_lds(fdm_bf2_kernel<1, NPMAX, false>);
        allow_lds(fdm_bf2_kernel<2, NPMAX, false>);
        allow_lds(fdm_bf2_kernel<0, NPMAX, true>);
        allow_lds(fdm_bf2_kernel<1, NPMAX, true>);
        allow_lds(fdm_bf2_kernel<2, NPMAX, true>);
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

// Planes per workgroup (measured at 64^3 with 8 - 160 vectors): the H*X
// kernels run fastest with one plane per workgroup, the split-BF16
// preconditioner with four.  PISCES_PPW_HX / PISCES_PPW_F64 / PISCES_PPW_FDM
// override.
int ppw_of(const char *env, int def)
{
    const char *e = std::getenv(env);
    const int v = e ? std::atoi(e) : 0;
    return v > 0 ? v : def;
}

template <int NPMAX>
void hx_bf2_launch(const HxJob *d, int n, int pa, int pb, int np, hipStream_t s)
{
    init_np<NPMAX>();
    static const int ppw = ppw_of("PISCES_PPW_HX", 1);
    const dim3 ga((pa + ppw - 1) / ppw, n), gb((pb + ppw - 1) / ppw, n);
    // pass A uses only the i-major plane (first half of the layout)
    if (ppw == 1) {
        hipLaunchKernelGGL((hx_bf2_kernel<0, NPMAX, false>), ga, dim3(kPT), bf2_lds(np) / 2, s, d);
        hipLaunchKernelGGL((hx_bf2_kernel<1, NPMAX, false>), gb, dim3(kPT), bf2_lds(np), s, d);
    } else {
        hipLaunchKernelGGL((hx_bf2_kernel<0, NPMAX, true>), ga, dim3(kPT), bf2_lds(np) / 2, s, d);
        hipLaunchKernelGGL((hx_bf2_kernel<1, NPMAX, true>), gb, dim3(kPT), bf2_lds(np), s, d);
    }
}

template <int NPMAX>
void hx_f64_launch(const HxJob *d, int n, int pa, int pb, int np, hipStream_t s)
{
    static std::once_flag once;
    std::call_once(once, [] {
        allow_lds(hx_f64_kernel<0, NPMAX, false>);
        allow_lds(hx_f64_kernel<1, NPMAX, false>);
        allow_lds(hx_f64_kernel<0, NPMAX, true>);
        allow_lds(hx_f64_kernel<1, NPMAX, true>);
    });
    static const int ppw = ppw_of("PISCES_PPW_F64", 1);
    const dim3 ga((pa + ppw - 1) / ppw, n), gb((pb + ppw - 1) / ppw, n);
    const size_t la = (size_t)np * (np + 2) * sizeof(double), lb = 2 * la;
    if (ppw == 1) {
        hipLaunchKernelGGL((hx_f64_kernel<0, NPMAX, false>), ga, dim3(kPT), la, s, d);
        hipLaunchKernelGGL((hx_f64_kernel<1, NPMAX, false>), gb, dim3(kPT), lb, s, d);
    } else {
        hipLaunchKernelGGL((hx_f64_kernel<0, NPMAX, true>), ga, dim3(kPT), la, s, d);
        hipLaunchKernelGGL((hx_f64_kernel<1, NPMAX, true>), gb, dim3(kPT), lb, s, d);
    }
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
        static const int ppw = ppw_of("PISCES_PPW_FDM", 4);
        const dim3 grid((nplanes + ppw - 1) / ppw, n);
        if (ppw == 1) {
            hipLaunchKernelGGL((fdm_bf2_kernel<0, NPMAX, false>), grid, dim3(kPT), lds, s, d);
            hipLaunchKernelGGL((fdm_bf2_kernel<1, NPMAX, false>), grid, dim3(kPT), lds, s, d);
            hipLaunchKernelGGL((fdm_bf2_kernel<2, NPMAX, false>), grid, dim3(kPT), lds, s, d);
        } else {
            hipLaunchKernelGGL((fdm_bf2_kernel<0, NPMAX, true>), grid, dim3(kPT), lds, s, d);
            hipLaunchKernelGGL((fdm_bf2_kernel<1, NPMAX, true>), grid, dim3(kPT), lds, s, d);
            hipLaunchKernelGGL((fdm_bf2_kernel<2, NPMAX, true>), grid, dim3(kPT), lds, s, d);
        }
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
    } else if (np <= 32) {
        hx_f64_launch<32>(d_jobs, njobs, max_planes_a, max_planes_b, np, s);
    } else if (np <= 64) {
        hx_f64_launch<64>(d_jobs, njobs, max_planes_a, max_planes_b, np, s);
    } else {
        // a 128-wide FP64 plane and Tx do not fit in LDS together
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
