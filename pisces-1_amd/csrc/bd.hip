// bd.hip -- batched kernels of the block-Davidson engine (gfx950).
//
// One launch serves every live instance: blockIdx.y (or blockIdx.x for the
// one-workgroup-per-task kernels) picks the task.  Loads of FP32 or FP64 columns
// go through ld_any(); the flag is uniform per task, so the branch is uniform.
#include "bd.hpp"

#include <mutex>

namespace pisces {

namespace {

constexpr int kT = 256;

__device__ inline double ld_any(const void *p, long long i, int is32)
{
    return is32 ? (double)static_cast<const float *>(p)[i] : static_cast<const double *>(p)[i];
}
__device__ inline void st_any(void *p, long long i, int is32, double v)
{
    if (is32) static_cast<float *>(p)[i] = (float)v;
    else static_cast<double *>(p)[i] = v;
}

// ------------------------------------------------------------------ Gram
__global__ void __launch_bounds__(kT) gram_kernel(const GramTask *__restrict__ tasks,
                                                  double *__restrict__ partial)
{
    const GramTask t = tasks[blockIdx.y];
    if ((int)blockIdx.x >= t.nch) return;
    extern __shared__ __align__(16) unsigned char smem[];
    double *sa = reinterpret_cast<double *>(smem);
    constexpr int P = kBdChunk + 1;
    double *sb = sa + t.ma * P;
    // each thread owns up to 4 (a, b) pairs and accumulates over cpb chunks
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int cc = 0; cc < t.cpb; ++cc) {
        const long long g0 = ((long long)blockIdx.x * t.cpb + cc) * kBdChunk;
        if (g0 >= t.N) break;
        const int len = (int)((t.N - g0) < kBdChunk ? (t.N - g0) : kBdChunk);
        __syncthreads();  // the previous chunk is consumed
        for (int idx = threadIdx.x; idx < t.ma * kBdChunk; idx += kT) {
            const int a = idx / kBdChunk, e = idx - a * kBdChunk;
            sa[a * P + e] = (e < len) ? ld_any(t.A, a * t.lda + g0 + e, t.a32) : 0.0;
        }
        for (int idx = threadIdx.x; idx < t.mb * kBdChunk; idx += kT) {
            const int b = idx / kBdChunk, e = idx - b * kBdChunk;
            sb[b * P + e] = (e < len) ? ld_any(t.B, b * t.ldb + g0 + e, t.b32) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int o = threadIdx.x + u * kT;
            if (o >= t.ma * t.mb) break;
            const int a = o % t.ma, b = o / t.ma;
            const double *pa = sa + a * P, *pb = sb + b * P;
            double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            for (int e = 0; e < kBdChunk; e += 4) {
                s0 += pa[e] * pb[e];
                s1 += pa[e + 1] * pb[e + 1];
                s2 += pa[e + 2] * pb[e + 2];
                s3 += pa[e + 3] * pb[e + 3];
            }
            acc[u] += (s0 + s1) + (s2 + s3);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int o = threadIdx.x + u * kT;
        if (o < t.ma * t.mb) partial[t.part_off + (long long)o * t.nch + blockIdx.x] = acc[u];
    }
}

__global__ void __launch_bounds__(kT) gram_reduce_kernel(const GramTask *__restrict__ tasks,
                                                         const double *__restrict__ partial)
{
    const GramTask t = tasks[blockIdx.y];
    const int o = blockIdx.x;
    if (o >= t.ma * t.mb) return;
    __shared__ double red[kT];
    double s = 0;
    for (int c = threadIdx.x; c < t.nch; c += kT) s += partial[t.part_off + (long long)o * t.nch + c];
    red[threadIdx.x] = s;
    for (int w = kT / 2; w > 0; w >>= 1) {
        __syncthreads();
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    }
    if (threadIdx.x == 0) t.out[(o % t.ma) + (long long)t.out_ld * (o / t.ma)] = red[0];
}

// ------------------------------------------------------------------ combine
__global__ void __launch_bounds__(kT) comb_kernel(const CombTask *__restrict__ tasks)
{
    const CombTask t = tasks[blockIdx.y];
    __shared__ double w[kBdMaxM * kBdMaxK];
    const int ma = t.ma1 + t.ma2;
    for (int idx = threadIdx.x; idx < ma * t.mb; idx += kT) {
        const int a = idx % ma, b = idx / ma;
        w[idx] = (a < t.ma1) ? t.W1[a + (long long)t.ldw1 * b] : t.W2[(a - t.ma1) + (long long)t.ldw2 * b];
    }
    __syncthreads();
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= t.N) return;
    double acc[kBdMaxK];
#pragma unroll
    for (int b = 0; b < kBdMaxK; ++b) acc[b] = 0.0;
    for (int a = 0; a < t.ma1; ++a) {
        const double x = ld_any(t.A1, a * t.lda1 + g, t.a1_32);
#pragma unroll
        for (int b = 0; b < kBdMaxK; ++b)
            if (b < t.mb) acc[b] += x * w[a + ma * b];
    }
    for (int a = 0; a < t.ma2; ++a) {
        const double x = ld_any(t.A2, a * t.lda2 + g, t.a2_32);
#pragma unroll
        for (int b = 0; b < kBdMaxK; ++b)
            if (b < t.mb) acc[b] += x * w[t.ma1 + a + ma * b];
    }
#pragma unroll
    for (int b = 0; b < kBdMaxK; ++b) {
        if (b >= t.mb) break;
        const long long yi = b * t.ldy + g;
        double v = acc[b];
        if (t.mode == 1) v = ld_any(t.Y, yi, t.y32) - v;
        else if (t.mode == 2) v = ld_any(t.Y, yi, t.y32) + v;
        st_any(t.Y, yi, t.y32, v);
    }
}

// ------------------------------------------------------------------ eigensolver
// Parallel Jacobi (round-robin pairing, m/2 disjoint rotations per round) on
// the symmetrised projected matrix in LDS; one workgroup per task.
__device__ inline int rr_index(int i, int r, int mp)
{
    return (i == 0) ? 0 : ((i - 1 + r) % (mp - 1)) + 1;
}

constexpr int kEigT = 64;

__global__ void __launch_bounds__(kEigT) eig_kernel(const EigTask *__restrict__ tasks)
{
    const EigTask t = tasks[blockIdx.x];
    constexpr int LD = kBdMaxM + 1;
    __shared__ double A[kBdMaxM * LD];
    __shared__ double V[kBdMaxM * LD];
    __shared__ double cs[kBdMaxM / 2], sn[kBdMaxM / 2];
    __shared__ int pp[kBdMaxM / 2], qq[kBdMaxM / 2];
    __shared__ double red[kEigT];
    __shared__ int order[kBdMaxM];
    __shared__ int s_stop;
    const int m = t.m;
    for (int idx = threadIdx.x; idx < m * m; idx += kEigT) {
        const int i = idx % m, j = idx / m;
        double g = (i <= j) ? t.G[i + (long long)t.ldg * j] : t.G[j + (long long)t.ldg * i];
        const bool di = t.keep && !t.keep[i];
        const bool dj = t.keep && !t.keep[j];
        if (di || dj) g = (i == j) ? 1e30 : 0.0;
        A[i + LD * j] = g;
        V[i + LD * j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int mp = m + (m & 1);
    const int npair = mp / 2;
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int idx = threadIdx.x; idx < m * m; idx += kEigT) {
            const int i = idx % m, j = idx / m;
            const double a = A[i + LD * j];
            if (i != j && fabs(A[i + LD * i]) < 1e29 && fabs(A[j + LD * j]) < 1e29) off += a * a;
            else if (i == j && fabs(a) < 1e29) tot += a * a;
        }
        red[threadIdx.x] = off;
        __syncthreads();
        for (int w = kEigT / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        const double offs = red[0];
        __syncthreads();
        red[threadIdx.x] = tot;
        __syncthreads();
        for (int w = kEigT / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) s_stop = (offs <= 1e-30 * (red[0] + offs)) ? 1 : 0;
        __syncthreads();
        if (s_stop) break;
        for (int r = 0; r < mp - 1; ++r) {
            if (threadIdx.x < npair) {
                int p = rr_index(threadIdx.x, r, mp), q = rr_index(mp - 1 - threadIdx.x, r, mp);
                if (p > q) { const int x = p; p = q; q = x; }
                double c = 1.0, s = 0.0;
                if (q < m) {
                    const double apq = A[p + LD * q];
                    if (apq != 0.0) {
                        const double th = (A[q + LD * q] - A[p + LD * p]) / (2.0 * apq);
                        double tt;
                        if (fabs(th) > 1e150) tt = 0.5 / th;
                        else tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        s = tt * c;
                    }
                }
                pp[threadIdx.x] = p;
                qq[threadIdx.x] = q;
                cs[threadIdx.x] = c;
                sn[threadIdx.x] = s;
            }
            __syncthreads();
            for (int item = threadIdx.x; item < m * npair; item += kEigT) {
                const int k = item % m, pr = item / m;
                const int p = pp[pr], q = qq[pr];
                if (q >= m || sn[pr] == 0.0) continue;
                const double c = cs[pr], s = sn[pr];
                const double akp = A[k + LD * p], akq = A[k + LD * q];
                A[k + LD * p] = c * akp - s * akq;
                A[k + LD * q] = s * akp + c * akq;
                const double vkp = V[k + LD * p], vkq = V[k + LD * q];
                V[k + LD * p] = c * vkp - s * vkq;
                V[k + LD * q] = s * vkp + c * vkq;
            }
            __syncthreads();
            for (int item = threadIdx.x; item < m * npair; item += kEigT) {
                const int k = item % m, pr = item / m;
                const int p = pp[pr], q = qq[pr];
                if (q >= m || sn[pr] == 0.0) continue;
                const double c = cs[pr], s = sn[pr];
                const double apk = A[p + LD * k], aqk = A[q + LD * k];
                A[p + LD * k] = c * apk - s * aqk;
                A[q + LD * k] = s * apk + c * aqk;
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        for (int i = 0; i < m; ++i) order[i] = i;
        for (int i = 0; i < m - 1; ++i) {
            int b = i;
            for (int j = i + 1; j < m; ++j)
                if (A[order[j] + LD * order[j]] < A[order[b] + LD * order[b]]) b = j;
            const int x = order[i]; order[i] = order[b]; order[b] = x;
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < m * t.k; idx += kEigT) {
        const int i = idx % m, c = idx / m;
        t.Y[i + (long long)m * c] = V[i + LD * order[c]];
    }
    if (threadIdx.x < t.k) t.theta[threadIdx.x] = A[order[threadIdx.x] + LD * order[threadIdx.x]];
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
            const double r = t.HU[i * t.ld + g] - t.theta[i] * t.U[i * t.ld + g];
            t.R[i * t.ld + g] = r;
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

__global__ void __launch_bounds__(64) resid_reduce_kernel(const ResTask *__restrict__ tasks,
                                                          const double *__restrict__ partial)
{
    const ResTask t = tasks[blockIdx.x];
    if ((int)threadIdx.x >= t.k) return;
    double s = 0;
    for (int c = 0; c < t.nch; ++c) s += partial[t.part_off + (long long)threadIdx.x * t.nch + c];
    t.out[threadIdx.x] = s;
    if (t.gather) {
        t.gather[threadIdx.x] = t.theta[threadIdx.x];
        t.gather[kBdMaxK + threadIdx.x] = s;
    }
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
        t.T[j * t.N + g] = t.R[r * t.ld + g] / dv;
    }
}

__global__ void cholinv_batched_kernel(const CholTask *__restrict__ tasks)
{
    const CholTask t = tasks[blockIdx.x];
    if (threadIdx.x != 0) return;
    const int k = t.k;
    double L[kBdMaxK * kBdMaxK];
    int keep[kBdMaxK];
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) L[i + kBdMaxK * j] = 0.0;
    for (int j = 0; j < k; ++j) {
        double d = t.G[j + k * j];
        for (int p = 0; p < j; ++p) d -= L[j + kBdMaxK * p] * L[j + kBdMaxK * p];
        const int kin = t.keep_in ? t.keep_in[j] : 1;
        keep[j] = (kin && d > t.tol * t.G[j + k * j] && d > 0.0) ? 1 : 0;
        if (!keep[j]) continue;
        const double ljj = sqrt(d);
        L[j + kBdMaxK * j] = ljj;
        for (int i = j + 1; i < k; ++i) {
            double s = t.G[i + k * j];
            for (int p = 0; p < j; ++p) s -= L[i + kBdMaxK * p] * L[j + kBdMaxK * p];
            L[i + kBdMaxK * j] = s / ljj;
        }
    }
    for (int c = 0; c < k; ++c)
        for (int i = k - 1; i >= 0; --i) {
            if (!keep[i] || !keep[c]) { t.W[i + k * c] = 0.0; continue; }
            double s = (i == c) ? 1.0 : 0.0;
            for (int p = i + 1; p < k; ++p) s -= L[p + kBdMaxK * i] * t.W[p + k * c];
            t.W[i + k * c] = s / L[i + kBdMaxK * i];
        }
    for (int c = 0; c < k; ++c) {
        t.keep_out[c] = keep[c];
        if (t.keep_copy) t.keep_copy[c] = keep[c];
    }
}

inline unsigned nblk(long long N) { return (unsigned)((N + kT - 1) / kT); }

void init_attrs()
{
    static std::once_flag once;
    std::call_once(once, [] {
        static_cast<void>(hipFuncSetAttribute(reinterpret_cast<const void *>(gram_kernel),
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              160 * 1024));
    });
}

}  // namespace

void bd_gram(const GramTask *d_tasks, int ntasks, int max_nch, int max_mamb, int max_cols,
             double *d_partial, hipStream_t s)
{
    if (ntasks <= 0) return;
    init_attrs();
    const size_t lds = (size_t)max_cols * (kBdChunk + 1) * sizeof(double);
    PC_REQUIRE(lds <= 160 * 1024, PISCES_ERR_ARG, "gram block too wide");
    hipLaunchKernelGGL(gram_kernel, dim3(max_nch, ntasks), dim3(kT), lds, s, d_tasks, d_partial);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(max_mamb, ntasks), dim3(kT), 0, s, d_tasks,
                       d_partial);
    PC_HIP(hipGetLastError());
}

void bd_combine(const CombTask *d_tasks, int ntasks, long long max_N, hipStream_t s)
{
    if (ntasks <= 0) return;
    hipLaunchKernelGGL(comb_kernel, dim3(nblk(max_N), ntasks), dim3(kT), 0, s, d_tasks);
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
    hipLaunchKernelGGL(resid_reduce_kernel, dim3(ntasks), dim3(64), 0, s, d_tasks, d_partial);
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
