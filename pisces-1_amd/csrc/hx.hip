// hx.hip -- H*X on a 3-D sine-DVR grid for gfx950.
//
//   Y = V o X + (Tx (x) I (x) I + I (x) Ty (x) I + I (x) I (x) Tz) X
//
// restating DVR::MatrixTimesVector (src/mtx.cpp:22-108) for a block of vectors.
// Vectors are x fastest: g = i + nx*(j + ny*k).
//
// Two passes over each vector, both plane kernels that stage one 2-D plane of
// the vector in LDS and do the contractions of the two axes that lie in it:
//   pass A, plane (i,j) at fixed k:  Y  = V o X + contraction along y (Ty)
//   pass B, plane (i,k) at fixed j:  Y += contraction along x (Tx) + along z (Tz)
// The whole cube of one vector (2 MB at 64^3 FP64) does not fit in LDS, and a
// plane holding full x, y and z lines does not exist, so a single pass would
// need every z-slab in every workgroup.  Two passes re-read X and Y once more;
// at the block sizes of the solver the re-read is served largely from the
// Infinity Cache (a 64^3 x 8 FP64 block is 17 MB against 256 MB).
//
// The contractions are dense N x N by N x N products per plane and run on the
// matrix cores: v_mfma_f64_16x16x4f64 (FP64 form), v_mfma_f32_16x16x4f32 (FP32
// form), and v_mfma_f32_16x16x32_bf16 with a hi/lo split of both operands and
// three products (BF16X2 form, FP32 accumulation).  A plain-VALU variant of the
// same tiling is kept as a fallback and for cross-checking.
#include "common.hpp"

#include <mutex>

namespace pisces {

static int g_variant = 1;
// FP64 matrix-core path enabled (lane layout checked on the device).
static constexpr bool kF64MfmaVerified = true;
void hx_set_variant(int v) { g_variant = v; }
int hx_get_variant() { return g_variant; }

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
};

template <int PASS>
__device__ inline bool plane_geo(const HxJob &J, int p, PlaneGeo &g)
{
    const int nplanes = (PASS == 0) ? J.nz : J.ny;
    if (p >= nplanes * J.nvec) return false;
    const int v = p / nplanes, pl = p - v * nplanes;
    g.ni = J.nx;
    if (PASS == 0) {
        g.nc = J.ny;
        g.cs = J.nx;
        g.vbase = (long long)J.nx * J.ny * pl;
        g.Tc = J.T[1];
        g.Ti = nullptr;
    } else {
        g.nc = J.nz;
        g.cs = (long long)J.nx * J.ny;
        g.vbase = (long long)J.nx * pl;
        g.Tc = J.T[2];
        g.Ti = J.T[0];
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
    const HxJob J = jobs[blockIdx.y];
    PlaneGeo g;
    if (!plane_geo<PASS>(J, blockIdx.x, g)) return;
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
            for (int r = 0; r < 4; ++r) tv[r] = (c0 + r < g.nc) ? Tc[(long long)cp * g.nc + c0 + r] : T(0);
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
                for (int a = 0; a < 4; ++a) tv[a] = (i0 + a < g.ni) ? Ti[(long long)ip * g.ni + i0 + a] : T(0);
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

// ---------------------------------------------------------------- MFMA form
// Output tiles are 16(c) x 16(i) so that the 16 lanes of a row write 16
// consecutive i (128 B for FP64).  With the matrix-core operand layout
//   A (16 x K): lane l holds A[l%16][k0 + l/16]      (K = 4 per instruction)
//   B (K x 16): lane l holds B[k0 + l/16][l%16]
//   D (16x16): lane l, element r holds D[row(l/16, r)][l%16] (see MfmaTraits)
// the two contractions of a plane are
//   along c: D(c,i) += sum_k Tc[c][k] P[k][i]   A = Tc (symmetric), B = P
//   along i: D(c,i) += sum_k P[c][k] Ti[k][i]   A = P,              B = Ti
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct MfmaTraits;
// Result rows held by lane group lq = lane/16: FP32 rows 4*lq + r; FP64 rows
// lq + 4*r (the FP64 instruction interleaves the four row groups).
template <> struct MfmaTraits<double> {
    typedef f64x4 acc_t;
    __device__ static inline acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lq, int r) { return lq + 4 * r; }
};
template <> struct MfmaTraits<float> {
    typedef f32x4 acc_t;
    __device__ static inline acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lq, int r) { return 4 * lq + r; }
};

// LDS row pitch (elements) of a plane whose rows have n valid entries: padded to
// a multiple of 16 plus one element so that 16 lanes reading one column of 16
// rows hit different banks.
__host__ __device__ inline int pitch_for(int n) { return ((n + 15) & ~15) + 1; }

template <typename T, int PASS>
__global__ void __launch_bounds__(kThreads) hx_plane_mfma(const HxJob *__restrict__ jobs)
{
    typedef MfmaTraits<T> M;
    typedef typename M::acc_t acc_t;
    const HxJob J = jobs[blockIdx.y];
    PlaneGeo g;
    if (!plane_geo<PASS>(J, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    T *P = reinterpret_cast<T *>(smem);
    const int nip = (g.ni + 15) & ~15, ncp = (g.nc + 15) & ~15;
    const int ld = pitch_for(g.ni);
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
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int nti = nip / 16, ntc = ncp / 16;
    const int kc = (g.nc + 3) & ~3, ki = (g.ni + 3) & ~3;
    for (int t = wave; t < nti * ntc; t += kThreads / 64) {
        const int ti = t % nti, tc = t / nti;
        const int i0 = 16 * ti, c0 = 16 * tc;
        acc_t acc = {};
        // along c: A[m][k] = Tc[c0+m][k] = Tc[k][c0+m], B[k][n] = P[k][i0+n]
        const int cA = c0 + l16;
        for (int k0 = 0; k0 < kc; k0 += 4) {
            const int k = k0 + lq;
            const T a = (k < g.nc && cA < g.nc) ? Tc[(long long)k * g.nc + cA] : T(0);
            const T b = P[k * ld + i0 + l16];
            acc = M::mma(a, b, acc);
        }
        if (PASS == 1) {
            // along i: A[m][k] = P[c0+m][k], B[k][n] = Ti[k][i0+n]
            const int iB = i0 + l16;
            for (int k0 = 0; k0 < ki; k0 += 4) {
                const int k = k0 + lq;
                const T a = P[(c0 + l16) * ld + k];
                const T b = (k < g.ni && iB < g.ni) ? Ti[(long long)k * g.ni + iB] : T(0);
                acc = M::mma(a, b, acc);
            }
        }
        const int i = i0 + l16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = c0 + M::row(lq, r);
            if (i < g.ni && c < g.nc) {
                const long long gi = g.base + i + c * g.cs;
                if (PASS == 0) Y[gi] = V[g.vbase + i + c * g.cs] * P[c * ld + i] + acc[r];
                else Y[gi] = Y[gi] + acc[r];
            }
        }
    }
}

// ------------------------------------------------------------- BF16X2 form
// X and T are FP32; each operand is split x = hi + lo with hi = bf16(x),
// lo = bf16(x - hi), and A*B ~ Ahi*Bhi + Ahi*Blo + Alo*Bhi (FP32 accumulate).
// v_mfma_f32_16x16x32_bf16: A lane l holds A[l%16][8*(l/16) + e], e < 8;
// B lane l holds B[8*(l/16) + e][l%16]; D as in the FP32 form.
__device__ inline unsigned short bf16_bits_rne(float x)
{
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    const unsigned short h = bf16_bits_rne(x);
    const float hf = __uint_as_float((unsigned int)h << 16);
    const unsigned short l = bf16_bits_rne(x - hf);
    hi = __builtin_bit_cast(__bf16, h);
    lo = __builtin_bit_cast(__bf16, l);
}

template <int PASS>
__global__ void __launch_bounds__(kThreads) hx_plane_bf16x2(const HxJob *__restrict__ jobs)
{
    const HxJob J = jobs[blockIdx.y];
    PlaneGeo g;
    if (!plane_geo<PASS>(J, blockIdx.x, g)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    // plane stored split: hi plane then lo plane, row pitch of 32-multiples + 8
    const int nip = (g.ni + 31) & ~31, ncp = (g.nc + 31) & ~31;
    const int ld = nip + 8;
    __bf16 *Ph = reinterpret_cast<__bf16 *>(smem);
    __bf16 *Pl = Ph + ncp * ld;
    float *Pf = reinterpret_cast<float *>(Pl + ncp * ld);  // original values for V o X
    const float *X = static_cast<const float *>(J.X);
    float *Y = static_cast<float *>(J.Y);
    for (int idx = threadIdx.x; idx < ncp * nip; idx += kThreads) {
        const int c = idx / nip, i = idx - c * nip;
        const float x = (i < g.ni && c < g.nc) ? X[g.base + i + c * g.cs] : 0.0f;
        __bf16 h, l;
        split_bf16(x, h, l);
        Ph[c * ld + i] = h;
        Pl[c * ld + i] = l;
        if (PASS == 0) Pf[c * nip + i] = x;
    }
    __syncthreads();
    const float *Tc = static_cast<const float *>(g.Tc);
    const float *Ti = static_cast<const float *>(g.Ti);
    const float *V = static_cast<const float *>(J.V);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int nti = (g.ni + 15) / 16, ntc = (g.nc + 15) / 16;
    const int kc = (g.nc + 31) & ~31, ki = (g.ni + 31) & ~31;
    for (int t = wave; t < nti * ntc; t += kThreads / 64) {
        const int ti = t % nti, tc = t / nti;
        const int i0 = 16 * ti, c0 = 16 * tc;
        f32x4 acc = {};
        const int cA = c0 + l16;
        for (int k0 = 0; k0 < kc; k0 += 32) {
            bf16x8 ah, al, bh, bl;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + 8 * lq + e;
                const float a = (k < g.nc && cA < g.nc) ? Tc[(long long)k * g.nc + cA] : 0.0f;
                __bf16 h, l;
                split_bf16(a, h, l);
                ah[e] = h;
                al[e] = l;
                bh[e] = Ph[k * ld + i0 + l16];
                bl[e] = Pl[k * ld + i0 + l16];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
        }
        if (PASS == 1) {
            const int iB = i0 + l16;
            for (int k0 = 0; k0 < ki; k0 += 32) {
                bf16x8 ah, al, bh, bl;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = k0 + 8 * lq + e;
                    ah[e] = Ph[(c0 + l16) * ld + k];
                    al[e] = Pl[(c0 + l16) * ld + k];
                    const float b = (k < g.ni && iB < g.ni) ? Ti[(long long)k * g.ni + iB] : 0.0f;
                    __bf16 h, l;
                    split_bf16(b, h, l);
                    bh[e] = h;
                    bl[e] = l;
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
            }
        }
        const int i = i0 + l16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = c0 + 4 * lq + r;
            if (i < g.ni && c < g.nc) {
                const long long gi = g.base + i + c * g.cs;
                if (PASS == 0) Y[gi] = V[g.vbase + i + c * g.cs] * Pf[c * nip + i] + acc[r];
                else Y[gi] = Y[gi] + acc[r];
            }
        }
    }
}

template <typename K>
void set_lds_limit(K kernel)
{
    static_cast<void>(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          160 * 1024));
}

void init_lds_limits()
{
    static std::once_flag once;
    std::call_once(once, [] {
    });
}

}  // namespace

void hx_launch(const HxJob *d_jobs, int njobs, int max_planes_a, int max_planes_b,
               int max_n, int precision, hipStream_t s)
{
    if (njobs <= 0) return;
    PC_REQUIRE(max_n > 0 && max_n <= 128, PISCES_ERR_UNSUPPORTED, "grid dimension above 128");
    init_lds_limits();
    dim3 blk(kThreads);
    dim3 ga(max_planes_a, njobs), gb(max_planes_b, njobs);
    if (precision == PISCES_BF16X2) {
        const int np = (max_n + 31) & ~31;
        const size_t lds = (size_t)np * (np + 8) * 2 * sizeof(__bf16) + (size_t)np * np * sizeof(float);
        hipLaunchKernelGGL(hx_plane_bf16x2<0>, ga, blk, lds, s, d_jobs);
        hipLaunchKernelGGL(hx_plane_bf16x2<1>, gb, blk, lds, s, d_jobs);
    } else if (precision == PISCES_FP32) {
        if (g_variant == 0) {
            const size_t lds = (size_t)((max_n + 3) & ~3) * ((max_n + 3) & ~3) * sizeof(float);
            hipLaunchKernelGGL((hx_plane_valu<float, 0>), ga, blk, lds, s, d_jobs);
            hipLaunchKernelGGL((hx_plane_valu<float, 1>), gb, blk, lds, s, d_jobs);
        } else {
            const size_t lds = (size_t)((max_n + 15) & ~15) * pitch_for(max_n) * sizeof(float);
            hipLaunchKernelGGL((hx_plane_mfma<float, 0>), ga, blk, lds, s, d_jobs);
            hipLaunchKernelGGL((hx_plane_mfma<float, 1>), gb, blk, lds, s, d_jobs);
        }
    } else {
        if (g_variant == 0 || !kF64MfmaVerified) {
            const size_t lds = (size_t)((max_n + 3) & ~3) * ((max_n + 3) & ~3) * sizeof(double);
            hipLaunchKernelGGL((hx_plane_valu<double, 0>), ga, blk, lds, s, d_jobs);
            hipLaunchKernelGGL((hx_plane_valu<double, 1>), gb, blk, lds, s, d_jobs);
        } else {
            const size_t lds = (size_t)((max_n + 15) & ~15) * pitch_for(max_n) * sizeof(double);
            hipLaunchKernelGGL((hx_plane_mfma<double, 0>), ga, blk, lds, s, d_jobs);
            hipLaunchKernelGGL((hx_plane_mfma<double, 1>), gb, blk, lds, s, d_jobs);
        }
    }
    PC_HIP(hipGetLastError());
}

}  // namespace pisces
