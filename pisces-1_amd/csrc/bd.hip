// bd.hip -- batched kernels of the block-Davidson engine (gfx950).
//
// One launch serves every live instance: blockIdx.y (or blockIdx.x for the
// one-workgroup-per-task kernels) picks the task.  Loads of FP32 or FP64 columns
// go through ld_any(); the flag is uniform per task, so the branch is uniform.
#include "bd.hpp"
#include "blockops.hpp"
#include "gmem.hpp"

#include <dlfcn.h>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <cstdlib>
#include <vector>
#include <mutex>

namespace pisces {

namespace {

constexpr int kT = 256;
// combinations and Ritz passes walk their grids in reverse (the Gram passes
// they follow walk forward)
#ifndef PISCES_COMB_REVERSE
#define PISCES_COMB_REVERSE 1
#endif
constexpr int kCombReverse = PISCES_COMB_REVERSE;

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline double ld_any(const void *p, long long i, int is32)
{
    return is32 ? (double)glb(static_cast<const float *>(p))[i] : glb(static_cast<const double *>(p))[i];
}
__device__ inline void st_any(void *p, long long i, int is32, double v)
{
    if (is32) glb(static_cast<float *>(p))[i] = (float)v;
    else glb(static_cast<double *>(p))[i] = v;
}
// Streamed outputs (combinations, Ritz vectors, their H-images, residuals)
// are stored with the nontemporal hint (stnt): on the served batch the
// combination family went from 4.93 to 5.13 TB/s and the Ritz/residual family
// from 5.69 to 5.92 TB/s.
__device__ inline void st_any_nt(void *p, long long i, int is32, double v)
{
    if (is32) stnt(glb(static_cast<float *>(p)) + i, (float)v);
    else stnt(glb(static_cast<double *>(p)) + i, v);
}

// ------------------------------------------------------------------ Gram
// A^T B of two tall column blocks (rows = grid points) on the FP64 matrix cores.
// The column pointers stay generic (flat loads): with global-address-space
// loads the compiler hoists every column's loads ahead of the MFMAs, and the
// 64-row variant drops from three to two waves per SIMD and runs slower.
// v_mfma_f64_16x16x4f64 takes A^T as a 16 x 4 operand (lane l: row m = l%16,
// k = l/16) and B as 4 x 16 (k = l/16, column n = l%16) and accumulates a
// 16 x 16 tile, element r of lane l holding D[l/16 + 4r][l%16].  The k index
// is a row of the blocks; any assignment of rows to k works as long as A and B
// use the same one, so lane (n, q) loads rows r0 + 4q .. r0 + 4q + 3 of column
// n (two 16-byte loads; the wave covers 16 rows x 16 columns, whole cache
// lines) and feeds them to four consecutive instructions.  All column tiles of
// A (ma <= 64) against the single tile of B (mb <= 16) are accumulated at once,
// so every element is read from memory once.  Workgroup = 4 waves on
// interleaved 16-row blocks of one chunk; the waves' tiles are summed in LDS
// into the chunk's partial.
typedef double gram_f64x4 __attribute__((ext_vector_type(4)));

// A columns are read by one workgroup each (nontemporal); the B tile is read
// by every workgroup of its rows (default policy, served from the L2)
template <bool NT, typename V> __device__ __forceinline__ V gram_ld(const V *p)
{
    if constexpr (NT) return ldnt(p);
    else return *p;
}

template <bool NT, typename T>
__device__ __forceinline__ void gram_load4(const T *col, long long r, long long N, bool vec,
                                           double (&v)[4])
{
    if (vec && r + 4 <= N) {
        if (sizeof(T) == 8) {
            const f64x2 a = gram_ld<NT>(reinterpret_cast<const f64x2 *>(col + r));
            const f64x2 b = gram_ld<NT>(reinterpret_cast<const f64x2 *>(col + r + 2));
            v[0] = a[0]; v[1] = a[1]; v[2] = b[0]; v[3] = b[1];
        } else {
            const f32x4 a = gram_ld<NT>(reinterpret_cast<const f32x4 *>(col + r));
            v[0] = a[0]; v[1] = a[1]; v[2] = a[2]; v[3] = a[3];
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (r + j < N) ? (double)col[r + j] : 0.0;
    }
}

template <typename TA, typename TB>
__device__ __forceinline__ void gram_chunk(const TA *A, long long lda, int ma,
                                           const TB *B, long long ldb, int mb,
                                           long long N, long long r_beg, long long r_end,
                                           bool vec, gram_f64x4 (&acc)[kBdMaxM / 16])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    const int nta = (ma + 15) / 16;
    const bool bval = n < mb;
    const TB *bcol = B + (long long)(bval ? n : 0) * ldb;
    for (long long r0 = r_beg + 16 * wave; r0 < r_end; r0 += 64) {
        const long long r = r0 + 4 * q;
        double bv[4];
        gram_load4<false>(bcol, r, N, vec, bv);
        if (!bval) bv[0] = bv[1] = bv[2] = bv[3] = 0.0;
#pragma unroll
        for (int t = 0; t < kBdMaxM / 16; ++t) {
            if (t >= nta) break;
            const int a = 16 * t + n;
            double av[4];
            gram_load4<true>(A + (long long)(a < ma ? a : 0) * lda, r, N, vec, av);
            if (a >= ma) av[0] = av[1] = av[2] = av[3] = 0.0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[j], bv[j], acc[t], 0, 0, 0);
        }
    }
}

// MFMA variant with long runs: per iteration a wave covers 64 consecutive rows,
// lane (n, q) holding rows r0 + 16q .. r0 + 16q + 15 of column n (four 16-byte
// loads per column for FP64), fed to 16 instructions; each column is then read
// in 512-byte runs like the VALU kernel.
template <bool NT, typename T>
__device__ __forceinline__ void gram_load16(const T *col, long long r, long long N, bool vec,
                                            double (&v)[16])
{
    if (vec && r + 16 <= N) {
        if (sizeof(T) == 8) {
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const f64x2 a = gram_ld<NT>(reinterpret_cast<const f64x2 *>(col + r + 2 * h));
                v[2 * h] = a[0];
                v[2 * h + 1] = a[1];
            }
        } else {
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const f32x4 a = gram_ld<NT>(reinterpret_cast<const f32x4 *>(col + r + 4 * h));
                v[4 * h] = a[0]; v[4 * h + 1] = a[1]; v[4 * h + 2] = a[2]; v[4 * h + 3] = a[3];
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = (r + j < N) ? (double)col[r + j] : 0.0;
    }
}

template <typename TA, typename TB>
__device__ __forceinline__ void gram_chunk64(const TA *A, long long lda, int ma,
                                             const TB *B, long long ldb, int mb,
                                             long long N, long long r_beg, long long r_end,
                                             bool vec, gram_f64x4 (&acc)[kBdMaxM / 16])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    const int nta = (ma + 15) / 16;
    const bool bval = n < mb;
    const TB *bcol = B + (long long)(bval ? n : 0) * ldb;
    for (long long r0 = r_beg + 64 * wave; r0 < r_end; r0 += 256) {
        const long long r = r0 + 16 * q;
        double bv[16];
        gram_load16<false>(bcol, r, N, vec, bv);
        if (!bval)
#pragma unroll
            for (int j = 0; j < 16; ++j) bv[j] = 0.0;
#pragma unroll
        for (int t = 0; t < kBdMaxM / 16; ++t) {
            if (t >= nta) break;
            const int a = 16 * t + n;
            double av[16];
            gram_load16<true>(A + (long long)(a < ma ? a : 0) * lda, r, N, vec, av);
            if (a >= ma)
#pragma unroll
                for (int j = 0; j < 16; ++j) av[j] = 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[j], bv[j], acc[t], 0, 0, 0);
        }
    }
}

// ------------------------------------------------------------------ Gram, VALU variant
// A^T B of two tall column blocks (rows = grid points).  Workgroup = one row
// chunk x one 8 x 8 tile of (A column, B column) pairs.  Each thread streams
// rows g = chunk start + tid + 256 s (consecutive lanes on consecutive rows,
// so every column load is one coalesced 256/512-byte wave access) and keeps
// the 64 products of its tile in registers; a butterfly reduce-scatter over
// the wave (63 shuffles) leaves tile entry `lane` on lane `lane`, and the four
// waves are summed through LDS into this chunk's partial.
template <typename TA, typename TB>
__device__ __forceinline__ void gram_rows(gptr<const TA> A, long long lda, int na,
                                          gptr<const TB> B, long long ldb, int nb,
                                          long long r0, long long r1, double (&acc)[kGTile * kGTile])
{
    if (na == kGTile && nb == kGTile) {
#pragma unroll 2
        for (long long g = r0 + threadIdx.x; g < r1; g += kT) {
            double a[kGTile], b[kGTile];
#pragma unroll
            for (int i = 0; i < kGTile; ++i) a[i] = (double)A[g + i * lda];
#pragma unroll
            for (int j = 0; j < kGTile; ++j) b[j] = (double)B[g + j * ldb];
#pragma unroll
            for (int i = 0; i < kGTile; ++i)
#pragma unroll
                for (int j = 0; j < kGTile; ++j) acc[i * kGTile + j] = fma(a[i], b[j], acc[i * kGTile + j]);
        }
    } else {
        // partial tile: columns past na / nb read column 0 and are weighted 0
        for (long long g = r0 + threadIdx.x; g < r1; g += kT) {
            double a[kGTile], b[kGTile];
#pragma unroll
            for (int i = 0; i < kGTile; ++i) a[i] = (i < na) ? (double)A[g + i * lda] : 0.0;
#pragma unroll
            for (int j = 0; j < kGTile; ++j) b[j] = (j < nb) ? (double)B[g + j * ldb] : 0.0;
#pragma unroll
            for (int i = 0; i < kGTile; ++i)
#pragma unroll
                for (int j = 0; j < kGTile; ++j) acc[i * kGTile + j] = fma(a[i], b[j], acc[i * kGTile + j]);
        }
    }
}

// Butterfly reduce-scatter of 64 values per lane over a 64-lane wave: after it,
// lane L holds the wave-wide sum of value L.  Each step halves the values a
// lane keeps: it sends the half its partner keeps and adds the half it gets.
template <int OFF>
__device__ __forceinline__ void rs_step(double (&v)[64], int lane)
{
    const bool up = (lane & OFF) != 0;
#pragma unroll
    for (int k = 0; k < OFF; ++k) {
        const double keep = up ? v[k + OFF] : v[k];
        const double give = up ? v[k] : v[k + OFF];
        v[k] = keep + __shfl_xor(give, OFF, 64);
    }
}

__device__ __forceinline__ double wave_reduce_scatter64(double (&v)[64])
{
    const int lane = threadIdx.x & 63;
    rs_step<32>(v, lane);
    rs_step<16>(v, lane);
    rs_step<8>(v, lane);
    rs_step<4>(v, lane);
    rs_step<2>(v, lane);
    rs_step<1>(v, lane);
    return v[0];
}

__global__ void __launch_bounds__(kT) gram_valu_kernel(const GramTask *__restrict__ tasks,
                                                       double *__restrict__ partial, int rows)
{
    const GramTask t = tasks[blockIdx.y];
    if (t.kind != 0) return;
    if (t.skip && *cst(t.skip)) return;
    const int ntile = t.cpb;  // tiles of this task
    if ((int)blockIdx.x >= t.nch * ntile) return;
    const int chunk = blockIdx.x / ntile, tile = blockIdx.x - chunk * ntile;
    const int nta = (t.ma + kGTile - 1) / kGTile;
    const int a0 = (tile % nta) * kGTile, b0 = (tile / nta) * kGTile;
    const int na = min(kGTile, t.ma - a0), nb = min(kGTile, t.mb - b0);
    const long long r0 = (long long)chunk * rows;
    const long long r1 = min(r0 + rows, t.N);
    double acc[kGTile * kGTile];
#pragma unroll
    for (int i = 0; i < kGTile * kGTile; ++i) acc[i] = 0.0;
    const gptr<const float> Bf = glb(static_cast<const float *>(t.B)) + b0 * t.ldb;
    const gptr<const double> Bd = glb(static_cast<const double *>(t.B)) + b0 * t.ldb;
    if (t.a32) {
        const gptr<const float> A = glb(static_cast<const float *>(t.A)) + a0 * t.lda;
        if (t.b32) gram_rows(A, t.lda, na, Bf, t.ldb, nb, r0, r1, acc);
        else gram_rows(A, t.lda, na, Bd, t.ldb, nb, r0, r1, acc);
    } else {
        const gptr<const double> A = glb(static_cast<const double *>(t.A)) + a0 * t.lda;
        if (t.b32) gram_rows(A, t.lda, na, Bf, t.ldb, nb, r0, r1, acc);
        else gram_rows(A, t.lda, na, Bd, t.ldb, nb, r0, r1, acc);
    }
    const double w = wave_reduce_scatter64(acc);
    __shared__ double red[kT / 64][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    red[wv][lane] = w;
    __syncthreads();
    if (threadIdx.x < 64) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < kT / 64; ++q) sum += red[q][threadIdx.x];
        const int i = threadIdx.x / kGTile, j = threadIdx.x - i * kGTile;
        if (i < na && j < nb) {
            const long long o = (long long)(a0 + i) + (long long)t.ma * (b0 + j);
            partial[t.part_off + o * t.nch + chunk] = sum;
        }
    }
}

template <int LONG>
__global__ void __launch_bounds__(kT) gram_kernel(const GramTask *__restrict__ tasks,
                                                  double *__restrict__ partial, int rows)
{
    const GramTask &t = tasks[blockIdx.y];
    if (t.kind != 0 || (int)blockIdx.x >= t.nch) return;
    if (t.skip && *cst(t.skip)) return;
    const long long r_beg = (long long)blockIdx.x * rows;
    const long long r_end = min(r_beg + rows, t.N);
    // 16-byte loads need every column start aligned: uniform per task
    const int ea = t.a32 ? 4 : 8, eb = t.b32 ? 4 : 8;
    const bool vec = ((reinterpret_cast<uintptr_t>(t.A) | reinterpret_cast<uintptr_t>(t.B)) & 15) == 0 &&
                     ((t.lda * ea) & 15) == 0 && ((t.ldb * eb) & 15) == 0;
    gram_f64x4 acc[kBdMaxM / 16];
#pragma unroll
    for (int i = 0; i < kBdMaxM / 16; ++i) acc[i] = gram_f64x4{0.0, 0.0, 0.0, 0.0};
#define PC_GRAM_CALL(TA_, TB_)                                                                  \
    do {                                                                                        \
        const TA_ *A = static_cast<const TA_ *>(t.A);                                           \
        const TB_ *B = static_cast<const TB_ *>(t.B);                                           \
        if (LONG) gram_chunk64(A, t.lda, t.ma, B, t.ldb, t.mb, t.N, r_beg, r_end, vec, acc);    \
        else gram_chunk(A, t.lda, t.ma, B, t.ldb, t.mb, t.N, r_beg, r_end, vec, acc);           \
    } while (0)
    if (t.a32) {
        if (t.b32) PC_GRAM_CALL(float, float);
        else PC_GRAM_CALL(float, double);
    } else {
        if (t.b32) PC_GRAM_CALL(double, float);
        else PC_GRAM_CALL(double, double);
    }
#undef PC_GRAM_CALL
    // sum the four waves' tiles: red[wave][tile][16 x 16]
    __shared__ double red[kT / 64][kBdMaxM / 16][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    const int nta = (t.ma + 15) / 16;
#pragma unroll
    for (int i = 0; i < kBdMaxM / 16; ++i) {
        if (i >= nta) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave][i][(q + 4 * r) * 16 + n] = acc[i][r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nta * 256; e += kT) {
        const int i = e >> 8, m = (e >> 4) & 15, c = e & 15;
        const int a = 16 * i + m;
        if (a >= t.ma || c >= t.mb) continue;
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kT / 64; ++w) sum += red[w][i][m * 16 + c];
        const long long o = (long long)a + (long long)t.ma * c;
        partial[t.part_off + o * t.nch + blockIdx.x] = sum;
    }
}

// ------------------------------------------------------------------ Gram, narrow B
// Tasks with at most kGNarrowB columns of B (a block of corrections against the
// basis).  On the FP64 matrix cores 3/4 of every 16-wide B tile would be empty,
// and at the sustained FP64 MFMA rate the padded products cost about as much
// time as the memory stream; on the VALU the 8 x 4 products of a row are ~1/6
// of it.  Workgroup = one row chunk x one group of kGNarrowA A columns; each
// thread streams rows in groups of kGNarrowRows (rows g + kT r: coalesced
// columns, all 12 columns of every row of the group loaded before the
// products) and keeps its 32 products in registers; a butterfly
// reduce-scatter over the wave and a sum over the four waves through LDS give
// the chunk's partials.  A columns are loaded nontemporal; the B columns keep
// the default policy, since every tile of the chunk reads them (from the L2).
constexpr int kGNarrowRows = 4;
// NA: A columns loaded per row (8, or 4 for a tile with at most 4 columns)
template <int NA, typename TA, typename TB>
__device__ __forceinline__ void gram_rows_narrow(gptr<const TA> A, long long lda, int na,
                                                 gptr<const TB> B, long long ldb, int nb,
                                                 long long r0, long long r1, long long S,
                                                 double (&acc)[kGNarrowA * kGNarrowB])
{
    constexpr int R = kGNarrowRows;
    // columns past na / nb read column 0 and are weighted 0; rows past r1
    // read row g and get weight 0 on B
    double wa[NA], wb[kGNarrowB];
#pragma unroll
    for (int i = 0; i < NA; ++i) wa[i] = i < na ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < kGNarrowB; ++j) wb[j] = j < nb ? 1.0 : 0.0;
    for (long long g = r0; g < r1; g += S * R) {
        long long gi[R];
        double wr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool in = g + S * r < r1;
            gi[r] = in ? g + S * r : g;
            wr[r] = in ? 1.0 : 0.0;
        }
        TA a[R][NA];
        TB b[R][kGNarrowB];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int i = 0; i < NA; ++i) a[r][i] = ldnt(A + (gi[r] + (i < na ? i : 0) * lda));
#pragma unroll
            for (int j = 0; j < kGNarrowB; ++j) b[r][j] = B[gi[r] + (j < nb ? j : 0) * ldb];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double bv[kGNarrowB];
#pragma unroll
            for (int j = 0; j < kGNarrowB; ++j) bv[j] = (double)b[r][j] * (wb[j] * wr[r]);
#pragma unroll
            for (int i = 0; i < NA; ++i) {
                const double av = (double)a[r][i] * wa[i];
#pragma unroll
                for (int j = 0; j < kGNarrowB; ++j)
                    acc[i * kGNarrowB + j] = fma(av, bv[j], acc[i * kGNarrowB + j]);
            }
        }
    }
}

template <int OFF, int NV>
__device__ __forceinline__ void rs_step_n(double (&v)[NV], int lane)
{
    const bool up = (lane & OFF) != 0;
#pragma unroll
    for (int k = 0; k < OFF; ++k) {
        const double keep = up ? v[k + OFF] : v[k];
        const double give = up ? v[k] : v[k + OFF];
        v[k] = keep + __shfl_xor(give, OFF, 64);
    }
}

__global__ void __launch_bounds__(kT) gram_narrow_kernel(const GramTask *__restrict__ tasks,
                                                         double *__restrict__ partial)
{
    constexpr int NV = kGNarrowA * kGNarrowB;
    static_assert(NV == 32, "reduce-scatter of 32 values");
    const GramTask t = tasks[blockIdx.y];
    if (t.kind != 1) return;
    if (t.skip && *cst(t.skip)) return;
    const int ntile = t.cpb;
    if ((int)blockIdx.x >= t.nch * ntile) return;
    // tile-major numbering: workgroups are dealt to the 8 XCDs round-robin, so
    // with nch a multiple of 8 every tile of a chunk runs on the same XCD,
    // whose L2 then serves the chunk's B columns to all of them
    const int tile = blockIdx.x / t.nch, chunk = blockIdx.x - tile * t.nch;
    const int a0 = tile * kGNarrowA;
    const int na = min(kGNarrowA, t.ma - a0), nb = t.mb;
    const long long r0 = (long long)chunk * kGRowsNarrow;
    const long long r1 = min(r0 + kGRowsNarrow, t.N);
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    const size_t ea = t.a32 ? 4 : 8;
    const void *Aa = static_cast<const unsigned char *>(t.A) + ea * (size_t)t.lda * a0;
    const gptr<const float> Bf = glb(static_cast<const float *>(t.B));
    const gptr<const double> Bd = glb(static_cast<const double *>(t.B));
    const long long g0 = r0 + threadIdx.x;
    if (t.a32) {
        const gptr<const float> A = glb(static_cast<const float *>(Aa));
        if (t.b32) {
            if (na <= 4) gram_rows_narrow<4>(A, t.lda, na, Bf, t.ldb, nb, g0, r1, kT, acc);
            else gram_rows_narrow<8>(A, t.lda, na, Bf, t.ldb, nb, g0, r1, kT, acc);
        } else {
            gram_rows_narrow<8>(A, t.lda, na, Bd, t.ldb, nb, g0, r1, kT, acc);
        }
    } else {
        const gptr<const double> A = glb(static_cast<const double *>(Aa));
        if (t.b32) {
            if (na <= 4) gram_rows_narrow<4>(A, t.lda, na, Bf, t.ldb, nb, g0, r1, kT, acc);
            else gram_rows_narrow<8>(A, t.lda, na, Bf, t.ldb, nb, g0, r1, kT, acc);
        } else {
            gram_rows_narrow<8>(A, t.lda, na, Bd, t.ldb, nb, g0, r1, kT, acc);
        }
    }
    // lane L ends with the sum over its half-wave of value L % 32, then over both halves
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    rs_step_n<16>(acc, lane);
    rs_step_n<8>(acc, lane);
    rs_step_n<4>(acc, lane);
    rs_step_n<2>(acc, lane);
    rs_step_n<1>(acc, lane);
    const double w = acc[0] + __shfl_xor(acc[0], 32, 64);
    __shared__ double red[kT / 64][NV];
    if (lane < NV) red[wv][lane] = w;
    __syncthreads();
    if (threadIdx.x < NV) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < kT / 64; ++q) sum += red[q][threadIdx.x];
        const int i = threadIdx.x / kGNarrowB, j = threadIdx.x - i * kGNarrowB;
        if (i < na && j < nb) {
            const long long o = (long long)(a0 + i) + (long long)t.ma * j;
            partial[t.part_off + o * t.nch + chunk] = sum;
        }
    }
}

// Partial-sum reductions.  Their loads are what the step waits on between a
// streaming pass and the small factor kernels that read its sums, so every
// load of an entry is issued before the first add: a lane sums a strided
// subset of the partials with its loads in flight together, and a fixed
// butterfly over the lanes of the entry gives the total (the same order on
// every run).
// Sum over c < n of p[c] by the L lanes (a power of two, lanes j of an aligned
// group of L) of the calling group: lane j takes c = j, j + L, ... in runs of
// eight loads.
template <int L>
__device__ __forceinline__ double group_sum(const double *__restrict__ p, int n, int j)
{
    double s[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] = 0.0;
    for (int c0 = j; c0 < n; c0 += 8 * L) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int c = c0 + L * u;
            x[u] = p[c < n ? c : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] += (c0 + L * u < n) ? x[u] : 0.0;
    }
    double v = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
#pragma unroll
    for (int off = L / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Gram entries: kGramRedLanes lanes per entry, kT / kGramRedLanes entries per
// workgroup; the entry's chunk partials are summed as above
constexpr int kGramRedLanes = 8;
__global__ void __launch_bounds__(kT) gram_reduce_kernel(const GramTask *__restrict__ tasks,
                                                         const double *__restrict__ partial)
{
    const GramTask t = tasks[blockIdx.y];
    if (t.skip && *cst(t.skip)) return;
    const int o = blockIdx.x * (kT / kGramRedLanes) + threadIdx.x / kGramRedLanes;
    const int j = threadIdx.x & (kGramRedLanes - 1);
    // whole lane groups leave together, so the group's shuffles see only live lanes
    if (o >= t.ma * t.mb) return;
    const double s = group_sum<kGramRedLanes>(partial + t.part_off + (long long)o * t.nch, t.nch, j);
    if (j == 0) t.out[(o % t.ma) + (long long)t.out_ld * (o / t.ma)] = s;
}

// ------------------------------------------------------------------ combine
// Workgroup = kT * RPT rows (thread: rows g0 + kT r, r < RPT, each column load
// coalesced over the wave), MB >= mb columns of Y accumulated in FP64
// registers.  The coefficients are wave-uniform loads from W through the
// constant address space (scalar loads; W is written only by earlier launches):
// no LDS staging and no barrier before the first column load, and each
// coefficient serves RPT rows.  Rows past N load row g0 (in range) and are
// not stored.
// one group of G consecutive columns from column a: G loads per row in flight
// before the products.  Columns of Y past mb get coefficient 0 from a clamped
// (in range) index, with no branch, so a group's loads overlap the products of
// the group before.
template <typename T, int MB, int RPT, int G>
__device__ __forceinline__ void comb_group(gptr<const T> A, long long lda, int a, cptr<double> W,
                                           int ldw, int mb, const long long (&gi)[RPT],
                                           double (&acc)[RPT][MB])
{
    T x[RPT][G];
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int r = 0; r < RPT; ++r) x[r][q] = ldnt(A + ((long long)(a + q) * lda + gi[r]));
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int bc = b < mb ? b : mb - 1;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const double wl = W[a + q + (long long)ldw * bc];
            const double w = b < mb ? wl : 0.0;
#pragma unroll
            for (int r = 0; r < RPT; ++r) acc[r][b] = fma((double)x[r][q], w, acc[r][b]);
        }
    }
}

// columns in groups of 8 (FP32) or 4 (FP64) -- 32 bytes per row in flight --
// then one half group, then single columns
template <typename T, int MB, int RPT>
__device__ __forceinline__ void comb_cols(gptr<const T> A, long long lda, int na, cptr<double> W,
                                          int ldw, int mb, const long long (&gi)[RPT],
                                          double (&acc)[RPT][MB])
{
    constexpr int CW = sizeof(T) == 4 ? 8 : 4;
    int a = 0;
    for (; a + CW <= na; a += CW) comb_group<T, MB, RPT, CW>(A, lda, a, W, ldw, mb, gi, acc);
    if (a + CW / 2 <= na) {
        comb_group<T, MB, RPT, CW / 2>(A, lda, a, W, ldw, mb, gi, acc);
        a += CW / 2;
    }
    for (; a < na; ++a) comb_group<T, MB, RPT, 1>(A, lda, a, W, ldw, mb, gi, acc);
}

template <int MB, int RPT>
__device__ __forceinline__ void comb_source(const void *A, long long lda, int na, int a32,
                                            const double *W, int ldw, int mb,
                                            const long long (&gi)[RPT], double (&acc)[RPT][MB])
{
    if (na <= 0) return;
    if (a32) comb_cols<float, MB, RPT>(glb(static_cast<const float *>(A)), lda, na, cst(W), ldw, mb, gi, acc);
    else comb_cols<double, MB, RPT>(glb(static_cast<const double *>(A)), lda, na, cst(W), ldw, mb, gi, acc);
}

// Launches that read the columns the previous launch read walk the grid in
// reverse order (rev = 1): workgroups are dispatched in order of their linear
// index, so such a launch starts on the rows whose lines the previous one
// left last in the L2 and memory-side caches.
__device__ __forceinline__ int grid_x(int rev) { return rev ? gridDim.x - 1 - blockIdx.x : blockIdx.x; }
__device__ __forceinline__ int grid_y(int rev) { return rev ? gridDim.y - 1 - blockIdx.y : blockIdx.y; }

template <int MB, int RPT>
__global__ void __launch_bounds__(kT) comb_kernel(const CombTask *__restrict__ tasks, int rev)
{
    const CombTask t = tasks[grid_y(rev)];
    if (t.skip && *cst(t.skip)) return;
    const long long g0 = (long long)grid_x(rev) * (kT * RPT) + threadIdx.x;
    if (g0 >= t.N) return;
    long long gi[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) gi[r] = (g0 + kT * r < t.N) ? g0 + kT * r : g0;
    double acc[RPT][MB];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int b = 0; b < MB; ++b) acc[r][b] = 0.0;
    comb_source<MB, RPT>(t.A1, t.lda1, t.ma1, t.a1_32, t.W1, t.ldw1, t.mb, gi, acc);
    comb_source<MB, RPT>(t.A2, t.lda2, t.ma2, t.a2_32, t.W2, t.ldw2, t.mb, gi, acc);
    comb_source<MB, RPT>(t.A3, t.lda3, t.ma3, t.a3_32, t.W3, t.ldw3, t.mb, gi, acc);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const long long g = g0 + kT * r;
        if (g >= t.N) break;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            if (b < t.mb) {
                const long long yi = b * t.ldy + g;
                double v = acc[r][b];
                if (t.mode == 1) v = ld_any(t.Y, yi, t.y32) - v;
                else if (t.mode == 2) v = ld_any(t.Y, yi, t.y32) + v;
                st_any_nt(t.Y, yi, t.y32, v);
                if (t.Y2) st_any_nt(t.Y2, yi, t.y2_32, v);
            }
        }
    }
}

// ------------------------------------------------------------------ eigensolver
// Parallel Jacobi (round-robin pairing, m/2 disjoint rotations per round) on
// the symmetrised projected matrix in LDS; one workgroup per task.
// 0 <= i < mp, 0 <= r < mp - 1: (i - 1 + r) mod (mp - 1) by one subtraction
__device__ inline int rr_index(int i, int r, int mp)
{
    if (i == 0) return 0;
    const int x = i - 1 + r;
    return (x >= mp - 1 ? x - (mp - 1) : x) + 1;
}

constexpr int kEigT = 512;  // measured at m = 28: 10% faster than 256 lanes, as fast as 1024

constexpr int kEigSweeps = 40;

// Cyclic Jacobi with round-robin pairing: every round applies m/2 disjoint
// rotations J as A <- J^T A J in one pass over 2 x 2 blocks (rows of one
// pair, columns of another: each block reads and writes only its own four
// entries) and V <- V J.  A pair is rotated only while its off-diagonal entry
// is above 1e-17 (|a_pp| + |a_qq|); sweeps end after one with no rotation or
// once the off-diagonal Frobenius norm is 1e-14 of the total.
__global__ void __launch_bounds__(kEigT) eig_kernel(const EigTask *__restrict__ tasks)
{
    const EigTask &t = tasks[blockIdx.x];
    constexpr int LD = kBdMaxM + 1;
    __shared__ double A[kBdMaxM * LD];
    __shared__ double V[kBdMaxM * LD];
    __shared__ double cs[kBdMaxM / 2], sn[kBdMaxM / 2];
    __shared__ int pp[kBdMaxM / 2], qq[kBdMaxM / 2];
    __shared__ double red[kEigT / 64][2];
    __shared__ int order[kBdMaxM];
    __shared__ int s_stop, s_rot;
    const int m = t.m, tid = threadIdx.x;
    const int mw = (t.Vfull && t.mwarm > 0 && t.mwarm <= m) ? t.mwarm : 0;
    for (int idx = tid; idx < m * m; idx += kEigT) {
        const int i = idx % m, j = idx / m;
        double g = (i <= j) ? t.G[i + (long long)t.ldg * j] : t.G[j + (long long)t.ldg * i];
        const bool di = t.keep && !t.keep[i];
        const bool dj = t.keep && !t.keep[j];
        if (di || dj) g = (i == j) ? 1e30 : 0.0;
        A[i + LD * j] = g;
        V[i + LD * j] = (i < mw && j < mw) ? t.Vfull[i + (long long)kBdMaxM * j] : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    if (mw > 0) {
        // warm start: V = diag(Vp, I), A <- V^T A V.  Only rows and columns
        // below mw change: Tm = A[:, :mw] Vp, then A[:mw, :mw] = Vp^T Tm[:mw]
        // and A[i, j] = A[j, i] = Tm[i, j] for i >= mw > j.  The leading block
        // comes out diagonal up to rounding (Vp diagonalised it last step), so
        // only the new rows and columns are left to rotate.
        __shared__ double Tm[kBdMaxM * LD];
        for (int idx = tid; idx < m * mw; idx += kEigT) {
            const int i = idx % m, j = idx / m;
            double acc = 0.0;
            for (int l = 0; l < mw; ++l) acc = fma(A[i + LD * l], V[l + LD * j], acc);
            Tm[i + LD * j] = acc;
        }
        __syncthreads();
        for (int idx = tid; idx < m * mw; idx += kEigT) {
            const int i = idx % m, j = idx / m;
            if (i < mw) {
                double acc = 0.0;
                for (int l = 0; l < mw; ++l) acc = fma(V[l + LD * i], Tm[l + LD * j], acc);
                A[i + LD * j] = acc;
            } else {
                A[i + LD * j] = Tm[i + LD * j];
                A[j + LD * i] = Tm[i + LD * j];
            }
        }
        __syncthreads();
    }
    const int mp = m + (m & 1);
    const int npair = mp / 2;
    const int lane = tid & 63, wave = tid >> 6;
    // this thread's 2 x 2 blocks (b = tid + kEigT q) and V items, fixed for
    // every round: their pair indices are divided out once here
    constexpr int kNB = (kBdMaxM / 2) * (kBdMaxM / 2) / kEigT, kNV = kBdMaxM * (kBdMaxM / 2) / kEigT;
    int bP1[kNB], bP2[kNB], vK[kNV], vP[kNV];
#pragma unroll
    for (int q = 0; q < kNB; ++q) {
        const int b = tid + kEigT * q;
        bP1[q] = b % npair;
        bP2[q] = b / npair;
    }
#pragma unroll
    for (int q = 0; q < kNV; ++q) {
        const int item = tid + kEigT * q;
        vK[q] = item % m;
        vP[q] = item / m;
    }
    for (int sweep = 0; sweep < kEigSweeps; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int idx = tid; idx < m * m; idx += kEigT) {
            const int i = idx % m, j = idx / m;
            const double a = A[i + LD * j];
            if (i != j && fabs(A[i + LD * i]) < 1e29 && fabs(A[j + LD * j]) < 1e29) off += a * a;
            else if (i == j && fabs(a) < 1e29) tot += a * a;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            off += __shfl_xor(off, o, 64);
            tot += __shfl_xor(tot, o, 64);
        }
        if (lane == 0) {
            red[wave][0] = off;
            red[wave][1] = tot;
        }
        __syncthreads();
        if (tid == 0) {
            double so = 0, st = 0;
            for (int w = 0; w < kEigT / 64; ++w) {
                so += red[w][0];
                st += red[w][1];
            }
            s_stop = (so <= 1e-28 * (st + so)) ? 1 : 0;
            s_rot = 0;
        }
        __syncthreads();
        if (s_stop) break;
        for (int r = 0; r < mp - 1; ++r) {
            if (tid < npair) {
                int p = rr_index(tid, r, mp), q = rr_index(mp - 1 - tid, r, mp);
                if (p > q) { const int x = p; p = q; q = x; }
                double c = 1.0, s = 0.0;
                if (q < m) {
                    const double apq = A[p + LD * q];
                    const double app = A[p + LD * p], aqq = A[q + LD * q];
                    if (fabs(apq) > 1e-17 * (fabs(app) + fabs(aqq)) + 1e-300) {
                        // t = tan of the annihilating angle to ~1 ulp; any t
                        // gives an exact rotation (c^2 + s^2 = 1 to rounding)
                        const double th = (aqq - app) * rcp64(2.0 * apq);
                        double tt;
                        if (fabs(th) > 1e150) {
                            tt = 0.5 * rcp64(th);
                        } else {
                            const double h = fma(th, th, 1.0);
                            tt = copysign(rcp64(fabs(th) + h * rsq64(h)), th);
                        }
                        c = rsq64(fma(tt, tt, 1.0));
                        s = tt * c;
                        s_rot = 1;  // every writer stores 1
                    }
                }
                pp[tid] = p;
                qq[tid] = q;
                cs[tid] = c;
                sn[tid] = s;
            }
            __syncthreads();
#pragma unroll
            for (int qb = 0; qb < kNB; ++qb) {
                if (tid + kEigT * qb >= npair * npair) break;
                const int P1 = bP1[qb], P2 = bP2[qb];
                const double s1 = sn[P1], s2 = sn[P2];
                if (s1 == 0.0 && s2 == 0.0) continue;
                const double c1 = cs[P1], c2 = cs[P2];
                const int p1 = pp[P1], q1 = qq[P1], p2 = pp[P2], q2 = qq[P2];
                const bool v1 = q1 < m, v2 = q2 < m;
                const double b00 = A[p1 + LD * p2];
                const double b01 = v2 ? A[p1 + LD * q2] : 0.0;
                const double b10 = v1 ? A[q1 + LD * p2] : 0.0;
                const double b11 = (v1 && v2) ? A[q1 + LD * q2] : 0.0;
                const double x00 = b00 * c2 - b01 * s2, x01 = b00 * s2 + b01 * c2;
                const double x10 = b10 * c2 - b11 * s2, x11 = b10 * s2 + b11 * c2;
                A[p1 + LD * p2] = c1 * x00 - s1 * x10;
                if (v2) A[p1 + LD * q2] = c1 * x01 - s1 * x11;
                if (v1) A[q1 + LD * p2] = s1 * x00 + c1 * x10;
                if (v1 && v2) A[q1 + LD * q2] = s1 * x01 + c1 * x11;
            }
#pragma unroll
            for (int qv = 0; qv < kNV; ++qv) {
                if (tid + kEigT * qv >= m * npair) break;
                const int k = vK[qv], pr = vP[qv];
                const double s = sn[pr];
                if (s == 0.0) continue;
                const int p = pp[pr], q = qq[pr];
                const double c = cs[pr];
                const double vkp = V[k + LD * p], vkq = V[k + LD * q];
                V[k + LD * p] = c * vkp - s * vkq;
                V[k + LD * q] = s * vkp + c * vkq;
            }
            __syncthreads();
        }
        if (!s_rot) break;  // uniform: read after the round's final barrier
    }
    // ascending order by rank (ties by index)
    if (tid < m) {
        const double di = A[tid + LD * tid];
        int rank = 0;
        for (int j = 0; j < m; ++j) {
            const double dj = A[j + LD * j];
            rank += (dj < di || (dj == di && j < tid)) ? 1 : 0;
        }
        order[rank] = tid;
    }
    __syncthreads();
    for (int idx = tid; idx < m * t.k; idx += kEigT) {
        const int i = idx % m, c = idx / m;
        t.Y[i + (long long)m * c] = V[i + LD * order[c]];
    }
    if (tid < t.k) t.theta[tid] = A[order[tid] + LD * order[tid]];
    if (t.Vfull)
        for (int idx = tid; idx < m * m; idx += kEigT) {
            const int i = idx % m, j = idx / m;
            t.Vfull[i + (long long)kBdMaxM * j] = V[i + LD * j];
        }
}

// ------------------------------------------------------------------ residual
__global__ void __launch_bounds__(kT) resid_kernel(const ResTask *__restrict__ tasks,
                                                   double *__restrict__ partial)
{
    const ResTask t = tasks[blockIdx.y];
    if ((int)blockIdx.x >= t.nch) return;
    __shared__ double red[kBdMaxK][kT / 64];
    double acc[kBdMaxK];
#pragma unroll
    for (int i = 0; i < kBdMaxK; ++i) acc[i] = 0.0;
    const long long g0 = (long long)blockIdx.x * kBdResChunk;
    for (int e = threadIdx.x; e < kBdResChunk; e += kT) {
        const long long g = g0 + e;
        if (g >= t.N) break;
#pragma unroll
        for (int i = 0; i < kBdMaxK; ++i) {
            if (i >= t.k) break;
            const double r = ldnt(glb(t.HU) + (i * t.ld + g)) -
                             t.theta[i * t.th_ld] * ldnt(glb(t.U) + (i * t.ld + g));
            st_any_nt(t.R, i * t.ld + g, t.r32, r);
            acc[i] += r * r;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kBdMaxK; ++i) {
        double v = acc[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[i][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < t.k) {
        double s = 0;
        for (int w = 0; w < kT / 64; ++w) s += red[threadIdx.x][w];
        partial[t.part_off + (long long)threadIdx.x * t.nch + blockIdx.x] = s;
    }
}

// one wave per column of the block (four per workgroup)
__global__ void __launch_bounds__(kT) resid_reduce_kernel(const ResTask *__restrict__ tasks,
                                                          const double *__restrict__ partial)
{
    const ResTask t = tasks[blockIdx.x];
    const int lane = threadIdx.x & 63;
    for (int b = threadIdx.x >> 6; b < t.k; b += kT / 64) {
        const double s = group_sum<64>(partial + t.part_off + (long long)b * t.nch, t.nch, lane);
        if (lane == 0) {
            t.out[b] = s;
            if (t.gather) {
                t.gather[b] = t.theta[b * t.th_ld];
                t.gather[kBdMaxK + b] = s;
            }
        }
    }
}

// ------------------------------------------------------------------ Ritz vectors + residuals
// G consecutive basis columns from column a (coefficient rows a0 + a ..) for
// RPT rows per thread: the 2 G RPT loads are issued before the products;
// coefficient columns past k are 0 from a clamped index (no branch).
template <typename T, int KB, int RPT, int G>
__device__ __forceinline__ void ritz_group(gptr<const T> A, gptr<const T> HA, long long N, int a,
                                           cptr<double> Y, int m, int a0, int k,
                                           const long long (&gi)[RPT], double (&u)[RPT][KB],
                                           double (&hu)[RPT][KB])
{
    T x[RPT][G], h[RPT][G];
#pragma unroll
    for (int q = 0; q < G; ++q)
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            x[r][q] = ldnt(A + ((long long)(a + q) * N + gi[r]));
            h[r][q] = ldnt(HA + ((long long)(a + q) * N + gi[r]));
        }
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        const int bc = b < k ? b : k - 1;
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const double wl = Y[(a0 + a + q) + (long long)m * bc];
            const double w = b < k ? wl : 0.0;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                u[r][b] = fma((double)x[r][q], w, u[r][b]);
                hu[r][b] = fma((double)h[r][q], w, hu[r][b]);
            }
        }
    }
}

// groups of 4 (FP32) or 2 (FP64) columns (half that for 16-column blocks),
// then single columns
template <typename T, int KB, int RPT>
__device__ __forceinline__ void ritz_cols(gptr<const T> A, gptr<const T> HA, long long N, int na,
                                          cptr<double> Y, int m, int a0, int k,
                                          const long long (&gi)[RPT], double (&u)[RPT][KB],
                                          double (&hu)[RPT][KB])
{
    constexpr int G = (sizeof(T) == 4 ? 4 : 2) / (KB > 8 ? 2 : 1);
    int a = 0;
    for (; a + G <= na; a += G) ritz_group<T, KB, RPT, G>(A, HA, N, a, Y, m, a0, k, gi, u, hu);
    for (; a < na; ++a) ritz_group<T, KB, RPT, 1>(A, HA, N, a, Y, m, a0, k, gi, u, hu);
}

// KB >= k: block size rounded up.  Workgroup = kT * RPT rows (thread: rows
// g0 + kT r); the Ritz coefficients and theta are scalar loads through the
// constant address space.  Rows past N read row 0 and are neither stored nor
// counted.
template <int KB, int RPT>
__global__ void __launch_bounds__(kT) ritz_kernel(const RitzTask *__restrict__ tasks,
                                                  double *__restrict__ partial, int rev)
{
    const RitzTask &t = tasks[grid_y(rev)];
    const int bx = grid_x(rev);
    if (bx >= t.nch) return;
    const int k = t.k, mc = t.mc, m = k + mc;
    __shared__ double red[KB][kT / 64];
    const cptr<double> Y = cst(t.Y);
    const long long g0 = (long long)bx * (kT * RPT) + threadIdx.x;
    long long gi[RPT];
    bool ok[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        ok[r] = g0 + kT * r < t.N;
        gi[r] = ok[r] ? g0 + kT * r : 0;
    }
    double u[RPT][KB], hu[RPT][KB];
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int b = 0; b < KB; ++b) u[r][b] = hu[r][b] = 0.0;
    ritz_cols<double, KB, RPT>(glb(t.X), glb(t.HX), t.N, k, Y, m, 0, k, gi, u, hu);
    if (t.c32)
        ritz_cols<float, KB, RPT>(glb(static_cast<const float *>(t.C)),
                                  glb(static_cast<const float *>(t.HC)), t.N, mc, Y, m, k, k, gi, u, hu);
    else
        ritz_cols<double, KB, RPT>(glb(static_cast<const double *>(t.C)),
                                   glb(static_cast<const double *>(t.HC)), t.N, mc, Y, m, k, k, gi, u,
                                   hu);
    const cptr<double> th = cst(t.theta);
    double nrm[KB];
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        nrm[b] = 0.0;
        if (b < k) {
            const double thb = th[b];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                if (ok[r]) {
                    const long long gr = gi[r];
                    const double res = hu[r][b] - thb * u[r][b];
                    if (t.store & kRitzU) stnt(glb(t.U) + (b * t.N + gr), u[r][b]);
                    if (t.store & kRitzHU) stnt(glb(t.HU) + (b * t.N + gr), hu[r][b]);
                    st_any_nt(t.R, b * t.N + gr, t.r32, res);
                    nrm[b] += res * res;
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < KB; ++b) {
        double v = nrm[b];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[b][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < k) {
        double sum = 0;
        for (int w = 0; w < kT / 64; ++w) sum += red[threadIdx.x][w];
        partial[t.part_off + (long long)threadIdx.x * t.nch + bx] = sum;
    }
}

// one wave per Ritz column (four per workgroup)
__global__ void __launch_bounds__(kT) ritz_reduce_kernel(const RitzTask *__restrict__ tasks,
                                                         const double *__restrict__ partial)
{
    const RitzTask &t = tasks[blockIdx.x];
    const int lane = threadIdx.x & 63;
    for (int b = threadIdx.x >> 6; b < t.k; b += kT / 64) {
        const double s = group_sum<64>(partial + t.part_off + (long long)b * t.nch, t.nch, lane);
        if (lane == 0) {
            t.out[b] = s;
            if (t.gather) {
                t.gather[b] = t.theta[b];
                t.gather[kBdMaxK + b] = s;
            }
        }
    }
}

// ------------------------------------------------------------------ orthogonalisation factor
// One 64-lane workgroup per task: S = G - h^T h (h may be absent), Cholesky
// with dropping, W = L^{-T}, F = -h W.  Shared by the Cholesky-QR factor.
__device__ void orth_factor(const double *__restrict__ hg, int hld, const double *__restrict__ Gin,
                            int gld, int m, int nt, double *__restrict__ Wo,
                            double *__restrict__ Fo, const int *kin, int *kout, int *kcopy,
                            double tol, int *skip = nullptr, double skip_tol = 0.0,
                            const double *part = nullptr, long long part_off = 0, int nch = 0)
{
    constexpr int K = kBdMaxK;
    __shared__ double h[kBdMaxM * kBdMaxK];
    __shared__ double S[K * K], L[K * K], W[K * K], Gs[K * K];
    __shared__ int keep[K];
    const int tid = threadIdx.x;
    if (part) {
        // [V | T]^T T from its chunk partials (fixed order per entry)
        const int ma = m + nt;
        for (int idx = tid; idx < ma * nt; idx += 64) {
            const int a = idx % ma, c = idx / ma;
            const double *p = part + part_off + (long long)(a + ma * c) * nch;
            double sum = 0.0;
            for (int q = 0; q < nch; ++q) sum += p[q];
            if (a < m) h[a + m * c] = sum;
            else Gs[(a - m) + K * c] = sum;
        }
    } else {
        for (int idx = tid; idx < m * nt; idx += 64) {
            const int a = idx % m, c = idx / m;
            h[idx] = hg[a + (long long)hld * c];
        }
        for (int idx = tid; idx < nt * nt; idx += 64) {
            const int i = idx % nt, j = idx / nt;
            Gs[i + K * j] = Gin[i + (long long)gld * j];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, j = idx / nt;
        double v = Gs[i + K * j];
        for (int a = 0; a < m; ++a) v -= h[a + m * i] * h[a + m * j];
        S[i + K * j] = v;
        L[i + K * j] = 0.0;
    }
    __syncthreads();
    for (int j = 0; j < nt; ++j) {
        if (tid == 0) {
            const double d = S[j + K * j];
            const int ki = kin ? kin[j] : 1;
            const int kj = (ki && d > tol * Gs[j + K * j] && d > 0.0) ? 1 : 0;
            keep[j] = kj;
            L[j + K * j] = kj ? sqrt(d) : 0.0;
        }
        __syncthreads();
        if (keep[j]) {
            const double ljj = L[j + K * j];
            for (int i = j + 1 + tid; i < nt; i += 64) L[i + K * j] = S[i + K * j] / ljj;
            __syncthreads();
            // trailing update of the lower triangle
            for (int idx = tid; idx < nt * nt; idx += 64) {
                const int i = idx % nt, l = idx / nt;
                if (l > j && i >= l) S[i + K * l] -= L[i + K * j] * L[l + K * j];
            }
        }
        __syncthreads();
    }
    // W = L^{-T} (upper triangular), one column per thread
    if (tid < nt) {
        const int c = tid;
        for (int i = nt - 1; i >= 0; --i) {
            if (!keep[i] || !keep[c]) {
                W[i + K * c] = 0.0;
                continue;
            }
            double v = (i == c) ? 1.0 : 0.0;
            for (int p = i + 1; p < nt; ++p) v -= L[p + K * i] * W[p + K * c];
            W[i + K * c] = v / L[i + K * i];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, c = idx / nt;
        Wo[i + (long long)nt * c] = W[i + K * c];
    }
    double dev = 0.0;  // max |W - I|, |F| over this thread's entries
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, c = idx / nt;
        dev = fmax(dev, fabs(W[i + K * c] - (i == c ? 1.0 : 0.0)));
    }
    if (Fo)
        for (int idx = tid; idx < m * nt; idx += 64) {
            const int a = idx % m, c = idx / m;
            double f = 0.0;
            for (int i = 0; i < nt; ++i) f -= h[a + m * i] * W[i + K * c];
            Fo[a + (long long)m * c] = f;
            dev = fmax(dev, fabs(f));
        }
    if (tid < nt) {
        kout[tid] = keep[tid];
        if (kcopy) kcopy[tid] = keep[tid];
    }
    if (skip) {
        int dropped = (tid < nt && !keep[tid]) ? 1 : 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            dev = fmax(dev, __shfl_xor(dev, off, 64));
            dropped |= __shfl_xor(dropped, off, 64);
        }
        if (tid == 0) *skip = (!dropped && dev < skip_tol) ? 1 : 0;
    }
}

// Cholesky with dropping of the nt x nt matrix A (LDS, ld kBdMaxK; its lower
// triangle is overwritten), then W = L^{-T} (zero columns for dropped ones).
// Column j is kept if kin[j] (null: all) and its pivot exceeds tol * dref[j].
// All 64 threads take part.
__device__ void chol_winv(double *A, const double *dref, int nt, const int *kin, double tol,
                          double *L, double *W, int *keep)
{
    constexpr int K = kBdMaxK;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < nt * nt; idx += 64) L[(idx % nt) + K * (idx / nt)] = 0.0;
    __syncthreads();
    for (int j = 0; j < nt; ++j) {
        if (tid == 0) {
            const double d = A[j + K * j];
            const int ki = kin ? kin[j] : 1;
            const int kj = (ki && d > tol * dref[j] && d > 0.0) ? 1 : 0;
            keep[j] = kj;
            L[j + K * j] = kj ? sqrt(d) : 0.0;
        }
        __syncthreads();
        if (keep[j]) {
            const double ljj = L[j + K * j];
            for (int i = j + 1 + tid; i < nt; i += 64) L[i + K * j] = A[i + K * j] / ljj;
            __syncthreads();
            for (int idx = tid; idx < nt * nt; idx += 64) {
                const int i = idx % nt, l = idx / nt;
                if (l > j && i >= l) A[i + K * l] -= L[i + K * j] * L[l + K * j];
            }
        }
        __syncthreads();
    }
    if (tid < nt) {
        const int c = tid;
        for (int i = nt - 1; i >= 0; --i) {
            if (!keep[i] || !keep[c]) {
                W[i + K * c] = 0.0;
                continue;
            }
            double v = (i == c) ? 1.0 : 0.0;
            for (int p = i + 1; p < nt; ++p) v -= L[p + K * i] * W[p + K * c];
            W[i + K * c] = v / L[i + K * i];
        }
    }
    __syncthreads();
}

// S is written one column at a time (rows 0 .. of the basis it then had), so
// only its upper triangle is current: S(a, b) = S[min + lds * max]
__device__ __forceinline__ double sym_at(cptr<double> S, int lds, int a, int b)
{
    return a <= b ? S[a + (long long)lds * b] : S[b + (long long)lds * a];
}

__global__ void __launch_bounds__(64) orth2_kernel(const Orth2Task *__restrict__ tasks)
{
    constexpr int K = kBdMaxK, M = kBdMaxM;
    const Orth2Task &t = tasks[blockIdx.x];
    const int m = t.m, nt = t.nt, ma = m + nt, tid = threadIdx.x;
    __shared__ double h0[M * K], F1[M * K], h1[M * K];
    __shared__ double G0[K * K], A[K * K], L[K * K], W1[K * K], W2[K * K], G1[K * K], dref[K];
    __shared__ int keep1[K], keep2[K];
    // h0, G0 from the chunk partials (fixed order per entry)
    for (int idx = tid; idx < ma * nt; idx += 64) {
        const int a = idx % ma, c = idx / ma;
        const double *p = t.part + t.part_off + (long long)(a + ma * c) * t.nch;
        double sum = 0.0;
        for (int q = 0; q < t.nch; ++q) sum += p[q];
        if (a < m) h0[a + M * c] = sum;
        else G0[(a - m) + K * c] = sum;
    }
    __syncthreads();
    // round 1
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, j = idx / nt;
        double v = G0[i + K * j];
        for (int a = 0; a < m; ++a) v -= h0[a + M * i] * h0[a + M * j];
        A[i + K * j] = v;
    }
    if (tid < nt) dref[tid] = G0[tid + K * tid];
    __syncthreads();
    // amplification of round 1: |T_j|^2 / |T_j - V h0_j|^2 over the kept columns
    __shared__ int s_alg;
    if (tid == 0) {
        int ok = 1;
        for (int j = 0; j < nt; ++j) {
            const double d = A[j + K * j];
            if (d > t.tol1 * dref[j] && d > 0.0 && dref[j] > t.amp2_max * d) ok = 0;
        }
        s_alg = ok;
    }
    chol_winv(A, dref, nt, nullptr, t.tol1, L, W1, keep1);
    const bool alg = s_alg != 0;  // written before chol_winv's barriers
    // F1 = -h0 W1; P = h0 W1 (kept in h1 for now)
    for (int idx = tid; idx < m * nt; idx += 64) {
        const int a = idx % m, c = idx / m;
        double pv = 0.0;
        for (int i = 0; i < nt; ++i) pv += h0[a + M * i] * W1[i + K * c];
        h1[a + M * c] = pv;
        F1[a + M * c] = -pv;
    }
    __syncthreads();
    if (!alg) {
        // round 1 alone; the second round runs on the data
        for (int idx = tid; idx < nt * nt; idx += 64) {
            const int i = idx % nt, c = idx / nt;
            t.W[i + (long long)nt * c] = W1[i + K * c];
        }
        for (int idx = tid; idx < m * nt; idx += 64) {
            const int a = idx % m, c = idx / m;
            t.F[a + (long long)m * c] = F1[a + M * c];
        }
        if (tid < nt) t.keep_out[tid] = keep1[tid];
        if (tid == 0) {
            *t.alg = 0;
            *t.skip2 = 0;
            if (t.ndata) *t.ndata += 1;
        }
        return;
    }
    // Q = S F1, then h1 = P + Q and
    // G1 = W1^T G0 W1 + P^T F1 + F1^T P + F1^T Q (P = h0 W1, held in h1)
    const cptr<double> Sv = cst(t.Sv);
    __shared__ double Q[M * K];
    for (int idx = tid; idx < m * nt; idx += 64) {
        const int a = idx % m, c = idx / m;
        double v = 0.0;
        for (int b = 0; b < m; ++b) v += sym_at(Sv, t.lds, a, b) * F1[b + M * c];
        Q[a + M * c] = v;
    }
    __syncthreads();
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, j = idx / nt;
        if (i < j) continue;
        double v = 0.0;
        for (int p = 0; p < nt; ++p) {
            double gw = 0.0;
            for (int q = 0; q < nt; ++q) gw += G0[p + K * q] * W1[q + K * j];
            v += W1[p + K * i] * gw;
        }
        for (int a = 0; a < m; ++a)
            v += h1[a + M * i] * F1[a + M * j] + F1[a + M * i] * h1[a + M * j] + F1[a + M * i] * Q[a + M * j];
        G1[i + K * j] = v;
        G1[j + K * i] = v;
    }
    __syncthreads();
    for (int idx = tid; idx < m * nt; idx += 64) {
        const int a = idx % m, c = idx / m;
        h1[a + M * c] += Q[a + M * c];
    }
    __syncthreads();
    // round 2
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, j = idx / nt;
        double v = G1[i + K * j];
        for (int a = 0; a < m; ++a) v -= h1[a + M * i] * h1[a + M * j];
        A[i + K * j] = v;
    }
    if (tid < nt) dref[tid] = G1[tid + K * tid];
    __syncthreads();
    chol_winv(A, dref, nt, keep1, t.tol2, L, W2, keep2);
    // W = W1 W2 (into W1's place: Wm), F = (F1 - h1) W2 (into F1's place)
    __shared__ double Wm[K * K];
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, c = idx / nt;
        double v = 0.0;
        for (int p = 0; p < nt; ++p) v += W1[i + K * p] * W2[p + K * c];
        Wm[i + K * c] = v;
        t.W[i + (long long)nt * c] = v;
    }
    for (int idx = tid; idx < m * nt; idx += 64) {
        const int a = idx % m, c = idx / m;
        double v = 0.0;
        for (int p = 0; p < nt; ++p) v += (F1[a + M * p] - h1[a + M * p]) * W2[p + K * c];
        Q[a + M * c] = v;  // F
        t.F[a + (long long)m * c] = v;
    }
    __syncthreads();
    // overlaps of the new block T W + V F (basis columns m .. m + nt) with the
    // basis and among themselves: P' = h0 W, Q' = S F,
    // V^T T' = P' + Q',  T'^T T' = W^T G0 W + P'^T F + F^T P' + F^T Q'
    // (the rounding of the stored block, ~1e-10 here, is not captured; every
    // restart rebuilds S from the data)
    for (int idx = tid; idx < m * nt; idx += 64) {
        const int a = idx % m, c = idx / m;
        double pv = 0.0, qv = 0.0;
        for (int p = 0; p < nt; ++p) pv += h0[a + M * p] * Wm[p + K * c];
        for (int b = 0; b < m; ++b) qv += sym_at(Sv, t.lds, a, b) * Q[b + M * c];
        F1[a + M * c] = pv;  // P'
        h1[a + M * c] = qv;  // Q'
    }
    __syncthreads();
    double *So = t.Sout;
    for (int idx = tid; idx < m * nt; idx += 64) {
        const int a = idx % m, c = idx / m;
        So[a + (long long)t.lds * (m + c)] = F1[a + M * c] + h1[a + M * c];
    }
    for (int idx = tid; idx < nt * nt; idx += 64) {
        const int i = idx % nt, j = idx / nt;
        if (i > j) continue;
        double v = 0.0;
        for (int p = 0; p < nt; ++p) {
            double gw = 0.0;
            for (int q = 0; q < nt; ++q) gw += G0[p + K * q] * Wm[q + K * j];
            v += Wm[p + K * i] * gw;
        }
        for (int a = 0; a < m; ++a)
            v += F1[a + M * i] * Q[a + M * j] + Q[a + M * i] * F1[a + M * j] + Q[a + M * i] * h1[a + M * j];
        So[(m + i) + (long long)t.lds * (m + j)] = v;
    }
    if (tid < nt) {
        t.keep_out[tid] = keep2[tid];
        if (t.keep_copy) t.keep_copy[tid] = keep2[tid];
    }
    if (tid == 0) {
        *t.alg = 1;
        *t.skip2 = 1;
    }
}

__global__ void __launch_bounds__(64) orth_kernel(const OrthTask *__restrict__ tasks)
{
    const OrthTask &t = tasks[blockIdx.x];
    if (t.noop && *cst(t.noop)) return;
    if (t.clear && threadIdx.x == 0) *t.clear = 0;
    orth_factor(t.h, t.hld, t.G, t.gld, t.m, t.nt, t.W, t.F, t.keep_in, t.keep_out, t.keep_copy,
                t.tol, t.skip, t.skip_tol, t.part, t.part_off, t.nch);
}

__global__ void __launch_bounds__(64) cholinv_batched_kernel(const CholTask *__restrict__ tasks)
{
    const CholTask &t = tasks[blockIdx.x];
    orth_factor(nullptr, 0, t.G, t.k, 0, t.k, t.W, nullptr, t.keep_in, t.keep_out, t.keep_copy, t.tol);
}

// ------------------------------------------------------------------ correction
__global__ void __launch_bounds__(kT) prec_kernel(const PrecTask *__restrict__ tasks)
{
    const PrecTask &t = tasks[blockIdx.y];
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= t.N) return;
    const double dg = t.D[g];
    for (int j = 0; j < t.nt; ++j) {
        const int r = t.src[j];
        double dv = t.theta[r] - dg;
        if (fabs(dv) < 1e-12) dv = 1e-12;
        st_any(t.T, j * t.N + g, t.t32, t.R[r * t.ld + g] / dv);
    }
}

// ------------------------------------------------------------------ restart directions
// One 64-lane workgroup per task; lane = coefficient row (m <= 64).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(64) coef_kernel(const CoefTask *__restrict__ tasks)
{
    const CoefTask t = tasks[blockIdx.x];
    const int r = threadIdx.x;
    const bool in = r < t.m;
    double y[kBdMaxK];
#pragma unroll
    for (int j = 0; j < kBdMaxK; ++j) y[j] = (in && j < t.k) ? t.Y[r + (long long)t.m * j] : 0.0;
    double q[kBdMaxK];
#pragma unroll
    for (int j = 0; j < kBdMaxK; ++j) q[j] = (in && j < t.k && r >= t.k) ? y[j] : 0.0;
    int keep[kBdMaxK];
#pragma unroll
    for (int c = 0; c < kBdMaxK; ++c) {
        keep[c] = 0;
        if (c >= t.k) continue;
        // two passes of Gram-Schmidt against the Ritz coefficients and the
        // directions kept so far, normalising after each pass
        double nrm = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int j = 0; j < kBdMaxK; ++j) {
                if (j >= t.k) break;
                const double h = wave_sum(y[j] * q[c]);
                q[c] -= h * y[j];
            }
#pragma unroll
            for (int j = 0; j < kBdMaxK; ++j) {
                if (j >= c) break;
                if (!keep[j]) continue;
                const double h = wave_sum(q[j] * q[c]);
                q[c] -= h * q[j];
            }
            nrm = sqrt(wave_sum(q[c] * q[c]));
            if (!(nrm > 1e-14)) break;  // the direction is numerically zero
            q[c] /= nrm;
        }
        keep[c] = (nrm > 0.5 && nrm < 2.0) ? 1 : 0;  // second pass left the unit norm intact
        if (!keep[c]) q[c] = 0.0;
        if (in) t.Q[r + (long long)t.m * c] = q[c];
    }
    if (r < t.k) {
        int kv = 0;
#pragma unroll
        for (int c = 0; c < kBdMaxK; ++c)
            if (c == r) kv = keep[c];
        t.keep_out[r] = kv;
    }
}

inline unsigned nblk(long long N) { return (unsigned)((N + kT - 1) / kT); }

}  // namespace

// Gram configuration: kernel variant (0 VALU register tiles, 1 matrix cores
// with 16-row blocks, 2 matrix cores with 64-row runs) and rows per workgroup.
// Set by the block engine's online tuner (or PISCES_GRAM=<variant>[:rows]).
struct GramCfg {
    int variant, rows;
};
static GramCfg g_gram = {2, kGRows};
void gram_load_config();

int gram_variant() { return g_gram.variant; }

bool gram_narrow(int mb)
{
    static const bool on = [] {
        const char *e = std::getenv("PISCES_GRAM_NARROW");
        return !(e && e[0] == '0');
    }();
    return on && mb <= kGNarrowB;
}

int gram_chunks(long long N) { return (int)((N + g_gram.rows - 1) / g_gram.rows); }

int gram_blocks_per_chunk(int ma, int mb)
{
    if (g_gram.variant != 0) return 1;
    return ((ma + kGTile - 1) / kGTile) * ((mb + kGTile - 1) / kGTile);
}

void bd_gram(const GramTask *d_tasks, int ntasks, int max_blocks, int max_blocks_narrow,
             int max_mamb, double *d_partial, hipStream_t s, bool reduce)
{
    if (ntasks <= 0) return;
    const int v = g_gram.variant, rows = g_gram.rows;
    if (max_blocks_narrow > 0)
        hipLaunchKernelGGL(gram_narrow_kernel, dim3(max_blocks_narrow, ntasks), dim3(kT), 0, s,
                           d_tasks, d_partial);
    if (max_blocks <= 0) {
    } else if (v == 0)
        hipLaunchKernelGGL(gram_valu_kernel, dim3(max_blocks, ntasks), dim3(kT), 0, s, d_tasks, d_partial, rows);
    else if (v == 1)
        hipLaunchKernelGGL(gram_kernel<0>, dim3(max_blocks, ntasks), dim3(kT), 0, s, d_tasks, d_partial, rows);
    else
        hipLaunchKernelGGL(gram_kernel<1>, dim3(max_blocks, ntasks), dim3(kT), 0, s, d_tasks, d_partial, rows);
    if (reduce)
        hipLaunchKernelGGL(gram_reduce_kernel,
                           dim3((max_mamb + kT / kGramRedLanes - 1) / (kT / kGramRedLanes), ntasks), dim3(kT), 0,
                           s, d_tasks, d_partial);
    PC_HIP(hipGetLastError());
}

void gram_set(int variant, int rows)
{
    g_gram.variant = variant;
    g_gram.rows = rows;
}

// The configuration comes from PISCES_GRAM=<variant>[:rows], else from the line
// "gram <variant> <rows>" of tuning.txt in the directory above the library
// (written by the benchmark's tuning mode from whole served runs), else the
// default (64-row matrix-core runs, 4096 rows per workgroup).
void gram_load_config()
{
    static std::once_flag once;
    std::call_once(once, [] {
        int v = 2, rows = kGRows;
        const char *e = std::getenv("PISCES_GRAM");
        if (e && e[0] >= '0' && e[0] <= '2') {
            v = e[0] - '0';
            rows = v == 0 ? kGRowsValu : kGRows;
            if (e[1] == ':') rows = std::atoi(e + 2);
        } else {
            Dl_info info;
            if (dladdr(reinterpret_cast<void *>(&gram_load_config), &info) && info.dli_fname) {
                std::string path(info.dli_fname);
                const size_t slash = path.rfind('/');
                path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) +
                       "/../tuning.txt";
                if (FILE *f = std::fopen(path.c_str(), "r")) {
                    char key[32];
                    int fv, fr;
                    while (std::fscanf(f, "%31s %d %d", key, &fv, &fr) == 3)
                        if (!std::strcmp(key, "gram") && fv >= 0 && fv <= 2) {
                            v = fv;
                            rows = fr;
                        }
                    std::fclose(f);
                }
            }
        }
        rows = std::max(256, std::min(rows, 1 << 20)) & ~255;
        g_gram.variant = v;
        g_gram.rows = rows;
    });
}

int gram_config(int *rows)
{
    if (rows) *rows = g_gram.rows;
    return g_gram.variant;
}

// rows per thread x columns: 4 x 4, 2 x 8, 1 x 16 (the same accumulator count)
void bd_combine(const CombTask *d_tasks, int ntasks, long long max_N, int max_mb, hipStream_t s)
{
    if (ntasks <= 0) return;
    PC_REQUIRE(max_mb >= 1 && max_mb <= kBdMaxK, PISCES_ERR_ARG, "combination width 1..16");
    auto grid = [&](int rpt) { return dim3((unsigned)((max_N + kT * rpt - 1) / (kT * rpt)), ntasks); };
    const int rev = kCombReverse;
    if (max_mb <= 4) hipLaunchKernelGGL((comb_kernel<4, 4>), grid(4), dim3(kT), 0, s, d_tasks, rev);
    else if (max_mb <= 8) hipLaunchKernelGGL((comb_kernel<8, 2>), grid(2), dim3(kT), 0, s, d_tasks, rev);
    else hipLaunchKernelGGL((comb_kernel<16, 1>), grid(1), dim3(kT), 0, s, d_tasks, rev);
    PC_HIP(hipGetLastError());
}

void bd_coef(const CoefTask *d_tasks, int ntasks, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(coef_kernel, dim3(ntasks), dim3(64), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

int ritz_rows(int k) { return kT * (k <= 4 ? 2 : 1); }

void bd_ritz(const RitzTask *d_tasks, int ntasks, int max_nch, int k, double *d_partial,
             hipStream_t s)
{
    if (ntasks <= 0) return;
    PC_REQUIRE(k >= 1 && k <= kBdMaxK, PISCES_ERR_ARG, "block size must be 1..16");
    PC_REQUIRE(max_nch >= 1, PISCES_ERR_ARG, "no rows");
    const dim3 grid(max_nch, ntasks);
    // rows per workgroup = kT * RPT = ritz_rows(k)
    const int rev = kCombReverse;
    if (k <= 2) hipLaunchKernelGGL((ritz_kernel<2, 2>), grid, dim3(kT), 0, s, d_tasks, d_partial, rev);
    else if (k <= 4) hipLaunchKernelGGL((ritz_kernel<4, 2>), grid, dim3(kT), 0, s, d_tasks, d_partial, rev);
    else if (k <= 8) hipLaunchKernelGGL((ritz_kernel<8, 1>), grid, dim3(kT), 0, s, d_tasks, d_partial, rev);
    else hipLaunchKernelGGL((ritz_kernel<16, 1>), grid, dim3(kT), 0, s, d_tasks, d_partial, rev);
    hipLaunchKernelGGL(ritz_reduce_kernel, dim3(ntasks), dim3(kT), 0, s, d_tasks, d_partial);
    PC_HIP(hipGetLastError());
}

void bd_orth2(const Orth2Task *d_tasks, int ntasks, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(orth2_kernel, dim3(ntasks), dim3(64), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

void bd_orth(const OrthTask *d_tasks, int ntasks, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(orth_kernel, dim3(ntasks), dim3(64), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

void bd_eig(const EigTask *d_tasks, int ntasks, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(eig_kernel, dim3(ntasks), dim3(kEigT), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

void bd_resid(const ResTask *d_tasks, int ntasks, int max_nch, double *d_partial, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(resid_kernel, dim3(max_nch, ntasks), dim3(kT), 0, s, d_tasks, d_partial);
    hipLaunchKernelGGL(resid_reduce_kernel, dim3(ntasks), dim3(kT), 0, s, d_tasks, d_partial);
    PC_HIP(hipGetLastError());
}

void bd_precond(const PrecTask *d_tasks, int ntasks, long long max_N, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(prec_kernel, dim3(nblk(max_N), ntasks), dim3(kT), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

void bd_cholinv(const CholTask *d_tasks, int ntasks, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(cholinv_batched_kernel, dim3(ntasks), dim3(64), 0, s, d_tasks);
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
