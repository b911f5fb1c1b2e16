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

// Owned plane values as FP32 (zero outside the n_c x n_i plane) from FP64,
// FP32 or BF16 (uint16_t bit patterns) storage.  Every lane loads
// unconditionally -- out-of-plane elements from the plane's (0, 0) -- so the
// loads issue back to back, and the value is selected afterwards.
template <typename S> __device__ __forceinline__ float as_f32(S v) { return (float)v; }
template <> __device__ __forceinline__ float as_f32<uint16_t>(uint16_t v)
{
    return __uint_as_float((unsigned int)v << 16);
}
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
                const float y = as_f32(src[ok ? (long long)c * g.cs + i : 0]);
                x[u][v][r] = ok ? y : 0.f;
            }
        }
}

// Row padding (BF16 elements) of the split planes in LDS: a row of a 64-point
// plane spans 64 + kBf2Pad elements (16 and 24 measured no faster)
constexpr int kBf2Pad = 8;

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
    const int ntc = g.npc / 16, nti = g.npi / 16, pt = g.npc + kBf2Pad, pc = g.npi + kBf2Pad;
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
// Planes p, p + pstride, ... (MULTI) or plane p alone of job J; returns at
// once where p is past the job's planes.
template <int PASS, int NPMAX, bool MULTI>
__device__ __forceinline__ void hx_bf2_planes(const HxJob &J, int p, int pstride, unsigned char *smem)
{
    constexpr int NCT = Own<NPMAX>::NCT, NIT = Own<NPMAX>::NIT, KS = Own<NPMAX>::KS;
    const AxisMats &Mi = J.ax[0];
    const AxisMats &Mc = J.ax[PASS == 0 ? 1 : 2];
    Geo g;
    if (!geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, p, g)) return;
    const int pt = g.npc + kBf2Pad, pc = g.npi + kBf2Pad;
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
    // pass B, grids up to 64, several planes per workgroup: Tx in B form held
    // for the workgroup's lifetime (with one plane it is loaded where used:
    // the same loads, 64 fewer live registers)
    constexpr bool HOLD = PASS == 1 && NPMAX <= 64 && MULTI;
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
        const bool more = MULTI && geo_of(PASS, J.nx, J.ny, J.nz, J.nvec, J.ld, Mi.np, Mc.np, p + pstride, gn);
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
        p += pstride;
    }
}

template <int PASS, int NPMAX, bool MULTI>
__global__ void __launch_bounds__(kPT) hx_bf2_kernel(const HxJob *__restrict__ jobs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    hx_bf2_planes<PASS, NPMAX, MULTI>(jobs[plane_y(PASS == 1)], plane_x(PASS == 1), gridDim.x, smem);
}

// ------------------------------------------------------------------ H*X, split-BF16, one launch
// Both passes in one launch, ordered per XCD so that pass B of a vector reads
// the X and Y that pass A left in that XCD's L2 instead of HBM.  The vectors
// are split into eight contiguous runs of about equal work, one per XCD; the
// workgroups of an XCD (read from the hardware's XCC_ID) take tickets from
// that XCD's counter in order, `batch` consecutive tickets per counter
// increment (one atomic per ticket serialised the 64 workgroups of an XCD on
// the counter's cache line: ~0.5 us per draw), each batch worked through in
// order, and the ticket sequence of run s0, s1, ... is
//   A(s0) | A(s1) B(s0) | A(s2) B(s1) | ... | B(s_last)
// (A(s): the n_z (i,j) planes of pass A of vector s; B(s): its n_y (i,k)
// planes).  A workgroup that draws a plane of B(s) waits until every plane
// of A(s) is stored: each A plane, once its stores have completed (all
// memory counters drained), adds 1 to the vector's completion count, and the
// wait polls that count in the XCD's L2.  Every plane a ticket waits on has a
// smaller ticket: either earlier in the waiting workgroup's own batch (done
// already) or drawn earlier by a running workgroup, which works through its
// batch in order and can itself wait only on smaller tickets; the smallest
// unfinished ticket waits on nothing, so the wait always ends.  It is bounded
// all the same (a spin limit; a workgroup past it counts the event in
// hx_spin_timeouts and goes on).
// X is not written by the launch and Y of a vector only by its own A and B
// planes, so no stale copy of either can sit in a CU's L1: the L1 is written
// through and holds only lines that the CU read after their last writer
// finished.
// Plan layout (ints): [0, 8) ticket counters, [8, 16) tickets per XCD,
// [16, 25) first segment of each XCD, [25] number of vectors, then the
// vectors' completion counts, then the segments (kSegInts ints each:
// job, vector, pass, first ticket, planes, vector slot).  The counters arrive
// zeroed with the plan.
constexpr int kSegInts = 6;
constexpr int kXcds = 8;
__device__ unsigned int hx_spin_timeouts;

// The ticket and completion counters of an XCD are touched only by that XCD's
// workgroups, and read-modify-writes are performed in the L2 (the vector L1
// performs no atomics), which every CU of the XCD shares.  A count is read by
// a compare-and-swap that never matches (an atomic performed in the L2), not
// by a load: a workgroup-scope load may be served by the L1, and an atomic add
// of 0 is compiled as such a load.
__device__ __forceinline__ int xcd_add(int *p, int v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int xcd_read(int *p)
{
    int seen = -1;  // counts are never negative: the swap never happens
    __hip_atomic_compare_exchange_strong(p, &seen, -1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return seen;
}

template <int NPMAX>
__global__ void __launch_bounds__(kPT) hx_bf2_fused_kernel(const HxJob *__restrict__ jobs,
                                                           int *__restrict__ plan, int batch)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_task[4];  // job, vector, pass, plane; -1 job: no tickets left
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kXcds - 1);
    const int nvecs = plan[25];
    int *done = plan + 26;
    const int *segs = plan + 26 + nvecs;
    const int seg_end = plan[16 + xcc + 1];
    int cursor = plan[16 + xcc];
    int bt = 0, be = 0;  // thread 0: the next ticket of the workgroup's batch, its end
    for (;;) {
        if (threadIdx.x == 0) {
            if (bt >= be) {
                bt = xcd_add(plan + xcc, batch);
                be = bt + batch;
            }
            const int t = bt++;
            int job = -1;
            if (t < plan[8 + xcc]) {
                while (cursor < seg_end && t >= segs[kSegInts * cursor + 3] + segs[kSegInts * cursor + 4])
                    ++cursor;
                const int *sg = segs + kSegInts * cursor;
                job = sg[0];
                s_task[1] = sg[1];
                s_task[2] = sg[2];
                s_task[3] = t - sg[3];
                if (sg[2] == 1) {
                    // pass B: every (i,j) plane of this vector's pass A must be stored
                    const int need = jobs[sg[0]].nz;
                    unsigned spins = 0;
                    while (xcd_read(done + sg[5]) < need) {
                        __builtin_amdgcn_s_sleep(2);
                        if (++spins > (1u << 24)) {
                            atomicAdd(&hx_spin_timeouts, 1u);
                            break;
                        }
                    }
                }
            }
            s_task[0] = job;
        }
        __syncthreads();
        const int job = s_task[0];
        if (job < 0) break;
        const HxJob &J = jobs[job];
        const int pass = s_task[2];
        // plane index within the job: vector-major, as the two-launch kernels count
        if (pass == 0) hx_bf2_planes<0, NPMAX, false>(J, s_task[1] * J.nz + s_task[3], 1, smem);
        else hx_bf2_planes<1, NPMAX, false>(J, s_task[1] * J.ny + s_task[3], 1, smem);
        if (pass == 0) {
            // this plane's Y stores have reached the L2 before it is counted
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0) {
                const int *sg = segs + kSegInts * cursor;
                xcd_add(done + sg[5], 1);
            }
        }
        __syncthreads();  // s_task and the LDS planes are reused by the next ticket
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
                        if (t->absden) den = fabs(den);
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
                                           float scale, float floor, bool absden, uint16_t *Pth,
                                           uint16_t *Ptl)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int ntc = g.npc / 16, nti = g.npi / 16, pt = g.npc + kBf2Pad;
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
                if (absden) den = fabsf(den);
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

template <typename D> __device__ __forceinline__ D of_f32(float v) { return (D)v; }
template <> __device__ __forceinline__ uint16_t of_f32<uint16_t>(float v) { return bf16_rne(v); }

// Stores the owned tiles to FP64, FP32 or BF16 (uint16_t bit patterns, rounded
// to nearest even) storage.
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
                if (c < g.nc && i < g.ni) dst[g.base + (long long)c * g.cs + i] = of_f32<D>(acc[u][v][r]);
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
    const int ntc = g.npc / 16, nti = g.npi / 16, ksc = g.npc / 32, pt = g.npc + kBf2Pad;
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
    const int ntc = g.npc / 16, nti = g.npi / 16, ksi = g.npi / 32, pc = g.npi + kBf2Pad;
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
    const int ntc = g.npc / 16, nti = g.npi / 16, ksi = g.npi / 32, pc = g.npi + kBf2Pad;
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
// Planes p, p + pstride, ... (MULTI) or plane p alone of one pass of task t.
template <int PASS, int NPMAX, bool MULTI>
__device__ __forceinline__ void fdm_bf2_planes(const FdmTask &t, int p, int pstride, unsigned char *smem)
{
    constexpr int NCT = Own<NPMAX>::NCT, NIT = Own<NPMAX>::NIT, KS = Own<NPMAX>::KS;
    const int kind = PASS == 1 ? 1 : 0;
    const AxisMats &Mx = t.ax[0], &My = t.ax[1], &Mz = t.ax[2];
    const AxisMats &Mc = PASS == 1 ? Mz : My;
    Geo g;
    if (!geo_of(kind, Mx.n, My.n, Mz.n, 1, 0, Mx.np, Mc.np, p, g)) return;
    const int pt = g.npc + kBf2Pad, pc = g.npi + kBf2Pad;
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
    // passes 0 and 2, grids up to 64, several planes per workgroup: Qx in B form
    // held for the workgroup's lifetime (with one plane it is loaded where used)
    constexpr bool HOLD = PASS != 1 && NPMAX <= 64 && MULTI;
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
    // pass 0 reads r (FP64 or FP32); passes 1 and 2 the BF16 intermediate
    float xb[NCT][NIT][4];
    auto load = [&](const Geo &gg) {
        if constexpr (PASS != 0) own_load(static_cast<const uint16_t *>(t.work), gg.base, gg, xb);
        else if (t.src32) own_load(static_cast<const float *>(t.src), gg.base, gg, xb);
        else own_load(static_cast<const double *>(t.src), gg.base, gg, xb);
    };
    load(g);
    f32x4 acc[NCT][NIT];
    for (;;) {
        own_stage(xb, g, PASS == 0 ? nullptr : Pth, Ptl, PASS == 0 ? P2h : nullptr, P2l);
        __syncthreads();
        Geo gn;
        const bool more = MULTI && geo_of(kind, Mx.n, My.n, Mz.n, 1, 0, Mx.np, Mc.np, p + pstride, gn);
        if (more) load(gn);
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
                       (float)t.floor, t.absden != 0, P2h, P2l);
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
        p += pstride;
    }
}

template <int PASS, int NPMAX, bool MULTI>
__global__ void __launch_bounds__(kPT) fdm_bf2_kernel(const FdmTask *__restrict__ tasks)
{
    extern __shared__ __align__(16) unsigned char smem[];
    fdm_bf2_planes<PASS, NPMAX, MULTI>(tasks[plane_y(PASS == 1)], plane_x(PASS == 1), gridDim.x, smem);
}

// ------------------------------------------------------------------ preconditioner, one launch
// The three passes of the split-BF16 preconditioner in one launch, ordered per
// XCD as the one-launch H*X is: the vectors split into eight contiguous runs,
// one per XCD, whose ticket sequence is
//   P0(s0) | P0(s1) P1(s0) | P0(s2) P1(s1) P2(s0) | ... | P2(s_last)
// so that pass 1 reads the intermediate pass 0 left in the XCD's L2 and pass 2
// the one pass 1 left there.  A plane of P1(s) waits until every plane of
// P0(s) is stored, a plane of P2(s) until every plane of P1(s) is; the counts
// are kept per vector as for H*X, and every awaited plane was drawn earlier by
// a resident workgroup that is not waiting on a later one.
// Plan layout (ints): [0, 8) ticket counters, [8, 16) tickets per XCD,
// [16, 25) first segment of each XCD, [25] vectors, then per vector the
// completion counts of pass 0 and of pass 1, then the segments (task, 0,
// pass, first ticket, planes, vector slot).
template <int NPMAX>
__global__ void __launch_bounds__(kPT) fdm_bf2_fused_kernel(const FdmTask *__restrict__ tasks,
                                                            int *__restrict__ plan, int batch)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_task[3];  // task, pass, plane; -1 task: no tickets left
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kXcds - 1);
    const int nvecs = plan[25];
    int *done = plan + 26;  // done[2 v + pass]
    const int *segs = plan + 26 + 2 * nvecs;
    const int seg_end = plan[16 + xcc + 1];
    int cursor = plan[16 + xcc];
    int bt = 0, be = 0;  // thread 0: the next ticket of the workgroup's batch, its end
    for (;;) {
        if (threadIdx.x == 0) {
            if (bt >= be) {
                bt = xcd_add(plan + xcc, batch);
                be = bt + batch;
            }
            const int t = bt++;
            int task = -1;
            if (t < plan[8 + xcc]) {
                while (cursor < seg_end && t >= segs[kSegInts * cursor + 3] + segs[kSegInts * cursor + 4])
                    ++cursor;
                const int *sg = segs + kSegInts * cursor;
                task = sg[0];
                s_task[1] = sg[2];
                s_task[2] = t - sg[3];
                if (sg[2] > 0) {
                    // every plane of the previous pass of this vector must be stored
                    const FdmTask &T = tasks[sg[0]];
                    const int need = sg[2] == 1 ? T.ax[2].n : T.ax[1].n;
                    unsigned spins = 0;
                    while (xcd_read(done + 2 * sg[5] + sg[2] - 1) < need) {
                        __builtin_amdgcn_s_sleep(2);
                        if (++spins > (1u << 24)) {
                            atomicAdd(&hx_spin_timeouts, 1u);
                            break;
                        }
                    }
                }
            }
            s_task[0] = task;
        }
        __syncthreads();
        const int task = s_task[0];
        if (task < 0) break;
        const FdmTask &T = tasks[task];
        const int pass = s_task[1];
        if (pass == 0) fdm_bf2_planes<0, NPMAX, false>(T, s_task[2], 1, smem);
        else if (pass == 1) fdm_bf2_planes<1, NPMAX, false>(T, s_task[2], 1, smem);
        else fdm_bf2_planes<2, NPMAX, false>(T, s_task[2], 1, smem);
        if (pass < 2) {
            // this plane's stores have reached the L2 before it is counted
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (threadIdx.x == 0) {
                const int *sg = segs + kSegInts * cursor;
                xcd_add(done + 2 * sg[5] + pass, 1);
            }
        }
        __syncthreads();  // s_task and the LDS planes are reused by the next ticket
    }
}

// ------------------------------------------------------------------ launch helpers
// Raises a kernel's dynamic LDS limit to `bytes` (the 160 KB of a CU less the
// kernel's static LDS); a failure is reported, not left behind as the thread's
// last HIP error.
template <typename K>
void allow_lds(K kernel, size_t bytes = 160 * 1024)
{
    PC_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

size_t bf2_lds(int np) { return (size_t)2 * 2 * np * (np + kBf2Pad) * sizeof(uint16_t); }
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
// preconditioner with four (one for 32-, 96- and 128-point planes).  PISCES_PPW_HX /
// PISCES_PPW_F64 / PISCES_PPW_FDM override.
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
        // four planes per workgroup for 64-point planes, one otherwise (64
        // vectors: 96^3 0.64 against 1.13 ms, where the multi-plane form holds
        // one wave per SIMD; 32^3 0.025 against 0.038 ms)
        static const int ppw = ppw_of("PISCES_PPW_FDM", NPMAX == 64 ? 4 : 1);
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
        else if (np <= 96) hx_bf2_launch<96>(d_jobs, njobs, max_planes_a, max_planes_b, np, s);
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

int hx_fused_plan_ints(const HxJob *h_jobs, int njobs)
{
    int nv = 0;
    for (int j = 0; j < njobs; ++j) nv += h_jobs[j].nvec;
    return 26 + nv + kSegInts * 2 * nv;
}

void hx_fused_plan(const HxJob *h_jobs, int njobs, int *plan)
{
    // every vector (job-major) with its plane work; eight contiguous runs of
    // about equal work, one per XCD, so the vectors of a job mostly share an
    // XCD and its L2 copy of the job's potential
    struct Vec { int job, vec; double w; };
    std::vector<Vec> vs;
    double total = 0;
    for (int j = 0; j < njobs; ++j) {
        const HxJob &J = h_jobs[j];
        const double w = (double)(J.nz + J.ny) * axis_pad(J.nx) * axis_pad(std::max(J.ny, J.nz));
        for (int v = 0; v < J.nvec; ++v) {
            vs.push_back({j, v, w});
            total += w;
        }
    }
    const int nv = (int)vs.size();
    std::fill(plan, plan + 26 + nv, 0);
    plan[25] = nv;
    int *segs = plan + 26 + nv;
    int nseg = 0, k = 0;
    double acc = 0;
    for (int x = 0; x < kXcds; ++x) {
        plan[16 + x] = nseg;
        std::vector<int> run;  // vector indices of this XCD, in order
        const double upto = total * (x + 1) / kXcds;
        while (k < nv && (x == kXcds - 1 || acc + 0.5 * vs[k].w <= upto)) {
            acc += vs[k].w;
            run.push_back(k++);
        }
        int t = 0;
        auto add = [&](int vi, int pass) {
            const HxJob &J = h_jobs[vs[vi].job];
            int *sg = segs + kSegInts * nseg++;
            sg[0] = vs[vi].job;
            sg[1] = vs[vi].vec;
            sg[2] = pass;
            sg[3] = t;
            sg[4] = pass == 0 ? J.nz : J.ny;
            sg[5] = vi;
            t += sg[4];
        };
        // A(s0) | A(s1) B(s0) | ... | B(s_last)
        for (size_t r = 0; r < run.size(); ++r) {
            add(run[r], 0);
            if (r > 0) add(run[r - 1], 1);
        }
        if (!run.empty()) add(run.back(), 1);
        plan[8 + x] = t;
    }
    plan[16 + kXcds] = nseg;
}

namespace {
template <int NPMAX>
void hx_fused_np_launch(const HxJob *d_jobs, int *d_plan, int np, int max_tickets, hipStream_t s)
{
    static std::once_flag once;
    std::call_once(once, [] { allow_lds(hx_bf2_fused_kernel<NPMAX>, bf2_lds(NPMAX)); });
    // workgroups per XCD: two per CU (32 CUs) unless an XCD has fewer planes
    static const int wpx = ppw_of("PISCES_HXF_WG", 64);
    static const int batch = ppw_of("PISCES_HXF_BATCH", 4);
    const int w = std::max(1, std::min(wpx, max_tickets));
    hipLaunchKernelGGL((hx_bf2_fused_kernel<NPMAX>), dim3(kXcds * w), dim3(kPT), bf2_lds(np), s,
                       d_jobs, d_plan, batch);
}
}  // namespace

void hx_fused_launch(const HxJob *h_jobs, const HxJob *d_jobs, int njobs, const int *h_plan,
                     int *d_plan, int max_n, hipStream_t s)
{
    if (njobs <= 0) return;
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    int mt = 0;
    for (int x = 0; x < kXcds; ++x) mt = std::max(mt, h_plan[8 + x]);
    const int np = axis_pad(max_n);
    if (np <= 32) hx_fused_np_launch<32>(d_jobs, d_plan, np, mt, s);
    else if (np <= 64) hx_fused_np_launch<64>(d_jobs, d_plan, np, mt, s);
    else hx_fused_np_launch<128>(d_jobs, d_plan, np, mt, s);
    PC_HIP(hipGetLastError());
}

bool one_launch_available()
{
    // the ordering needs the eight XCDs of one device (SPX partitioning):
    // with fewer compute units the device is a partition of one XCD
    static const bool on = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return false;
        return cus >= 8 * 16;
    }();
    return on;
}

// The one-launch forms are opt-in (PISCES_HX_FUSED=1, PISCES_FDM_FUSED=1): at
// 64^3 x 64 vectors they take 0.29 ms (H*X) and 0.43 ms (preconditioner)
// against 0.086 and 0.129 ms for the multi-launch forms (0.49 and 0.67 ms with
// one ticket per counter increment).  A ticket's plane takes ~14 us from draw
// to counted store, and an XCD keeps 64 - 128 of them in flight, where the
// multi-launch grids keep thousands of planes going; these kernels are bound
// by that latency, not by the HBM traffic the L2 reuse saves.
bool hx_fused_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("PISCES_HX_FUSED");
        return e && e[0] == '1' && one_launch_available();
    }();
    return on;
}

unsigned hx_fused_spin_timeouts()
{
    // a stream of its own (released again), not the null stream
    unsigned v = 0;
    hipStream_t s;
    PC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const hipError_t e = hipMemcpyFromSymbolAsync(&v, HIP_SYMBOL(hx_spin_timeouts), sizeof(v), 0,
                                                  hipMemcpyDeviceToHost, s);
    const hipError_t w = hipStreamSynchronize(s);
    static_cast<void>(hipStreamDestroy(s));
    PC_HIP(e);
    PC_HIP(w);
    return v;
}

int fdm_fused_plan_ints(int ntasks) { return 26 + 2 * ntasks + kSegInts * 3 * ntasks; }

void fdm_fused_plan(const FdmTask *h_tasks, int ntasks, int *plan)
{
    // runs of about equal work per XCD (plane work ~ planes x plane size)
    std::vector<double> w(ntasks);
    double total = 0;
    for (int i = 0; i < ntasks; ++i) {
        const FdmTask &T = h_tasks[i];
        w[i] = (double)(2 * T.ax[2].n + T.ax[1].n) * axis_pad(T.ax[0].n) *
               axis_pad(std::max(T.ax[1].n, T.ax[2].n));
        total += w[i];
    }
    std::fill(plan, plan + 26 + 2 * ntasks, 0);
    plan[25] = ntasks;
    int *segs = plan + 26 + 2 * ntasks;
    int nseg = 0, k = 0;
    double acc = 0;
    for (int x = 0; x < kXcds; ++x) {
        plan[16 + x] = nseg;
        std::vector<int> run;
        const double upto = total * (x + 1) / kXcds;
        while (k < ntasks && (x == kXcds - 1 || acc + 0.5 * w[k] <= upto)) {
            acc += w[k];
            run.push_back(k++);
        }
        int t = 0;
        auto add = [&](int ti, int pass) {
            const FdmTask &T = h_tasks[ti];
            int *sg = segs + kSegInts * nseg++;
            sg[0] = ti;
            sg[1] = 0;
            sg[2] = pass;
            sg[3] = t;
            sg[4] = pass == 1 ? T.ax[1].n : T.ax[2].n;
            sg[5] = ti;
            t += sg[4];
        };
        // P0(s) | P0(s+1) P1(s) | P0(s+2) P1(s+1) P2(s) | ...
        const int S = (int)run.size();
        for (int st = 0; st < S + 2; ++st) {
            if (st < S) add(run[st], 0);
            if (st >= 1 && st - 1 < S) add(run[st - 1], 1);
            if (st >= 2 && st - 2 < S) add(run[st - 2], 2);
        }
        plan[8 + x] = t;
    }
    plan[16 + kXcds] = nseg;
}

namespace {
template <int NPMAX>
void fdm_fused_np_launch(const FdmTask *d_tasks, int *d_plan, int np, int max_tickets, hipStream_t s)
{
    static std::once_flag once;
    std::call_once(once, [] { allow_lds(fdm_bf2_fused_kernel<NPMAX>, bf2_lds(NPMAX)); });
    static const int wpx = ppw_of("PISCES_FDMF_WG", 64);
    static const int batch = ppw_of("PISCES_FDMF_BATCH", 4);
    const int w = std::max(1, std::min(wpx, max_tickets));
    hipLaunchKernelGGL((fdm_bf2_fused_kernel<NPMAX>), dim3(kXcds * w), dim3(kPT), bf2_lds(np), s,
                       d_tasks, d_plan, batch);
}
}  // namespace

void fdm_fused_launch(const FdmTask *d_tasks, int *d_plan, const int *h_plan, int max_n,
                      hipStream_t s)
{
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    int mt = 0;
    for (int x = 0; x < kXcds; ++x) mt = std::max(mt, h_plan[8 + x]);
    if (mt == 0) return;
    const int np = axis_pad(max_n);
    if (np <= 32) fdm_fused_np_launch<32>(d_tasks, d_plan, np, mt, s);
    else if (np <= 64) fdm_fused_np_launch<64>(d_tasks, d_plan, np, mt, s);
    else fdm_fused_np_launch<128>(d_tasks, d_plan, np, mt, s);
    PC_HIP(hipGetLastError());
}

bool fdm_fused_enabled()
{
    static const bool on = [] {
        const char *e = std::getenv("PISCES_FDM_FUSED");
        return e && e[0] == '1' && one_launch_available();
    }();
    return on;
}

void fdm_launch(const FdmTask *d_tasks, int ntasks, int max_n, int precision, hipStream_t s)
{
    if (ntasks <= 0) return;
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    const int np = axis_pad(max_n);
    // planes per pass: the largest of ny (pass 1) and nz (passes 0, 2) <= max_n
    if (np <= 32) fdm_np_launch<32>(d_tasks, ntasks, max_n, np, precision, s);
    else if (np <= 64) fdm_np_launch<64>(d_tasks, ntasks, max_n, np, precision, s);
    else if (np <= 96) fdm_np_launch<96>(d_tasks, ntasks, max_n, np, precision, s);
    else fdm_np_launch<128>(d_tasks, ntasks, max_n, np, precision, s);
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
