// blockops.hpp -- tall-skinny block operations used by the Davidson solvers.
#pragma once

#include "common.hpp"

namespace pisces {

// G(a, b) = sum_g A[a][g] B[b][g]  for a < ma, b < mb, accumulated in FP64;
// written to d_out[a + out_ld * b] on the device.  d_partial needs
// ceil(N / gram_chunk(ma, mb)) * ma * mb doubles.
int gram_chunk(int ma, int mb);
void tsgram(const double *A, long long lda, int ma, const double *B, long long ldb, int mb,
            long long N, double *d_partial, size_t partial_cap, double *d_out, int out_ld,
            hipStream_t s);
void tsgram(const float *A, long long lda, int ma, const double *B, long long ldb, int mb,
            long long N, double *d_partial, size_t partial_cap, double *d_out, int out_ld,
            hipStream_t s);
void tsgram(const float *A, long long lda, int ma, const float *B, long long ldb, int mb,
            long long N, double *d_partial, size_t partial_cap, double *d_out, int out_ld,
            hipStream_t s);
void tsgram(const double *A, long long lda, int ma, const float *B, long long ldb, int mb,
            long long N, double *d_partial, size_t partial_cap, double *d_out, int out_ld,
            hipStream_t s);

// Y[b] (=, -=, +=) sum_a A[a] W[a + ldw b]  (mode 0, 1, 2); W on the device,
// ma <= 64, mb <= 16.  FP64 accumulation.
void combine(const double *A, long long lda, int ma, const double *d_W, int ldw, int mb,
             double *Y, long long ldy, long long N, int mode, hipStream_t s);
void combine(const float *A, long long lda, int ma, const double *d_W, int ldw, int mb,
             double *Y, long long ldy, long long N, int mode, hipStream_t s);
void combine(const float *A, long long lda, int ma, const double *d_W, int ldw, int mb,
             float *Y, long long ldy, long long N, int mode, hipStream_t s);
void combine(const double *A, long long lda, int ma, const double *d_W, int ldw, int mb,
             float *Y, long long ldy, long long N, int mode, hipStream_t s);

void residual(double *U, const double *ZV, double *R, long long ld, long long N, int nr,
              const double *d_scale, const double *d_lam, hipStream_t s);
void precond(const double *R, long long ldr, const double *diag, double *C, long long ldc,
             long long N, int k, const int *d_src, const double *d_lam, hipStream_t s);
void jd_precond(const double *R, long long ldr, const double *U, const double *diag, double *C,
                double *J, long long ldc, long long N, int k, const int *d_src,
                const int *d_usrc, const double *d_lam, hipStream_t s);
void jd_finish(double *C, const double *J, long long ldc, long long N, int k,
               const double *d_dots, hipStream_t s);
void cholinv(const double *d_G, int k, double *d_W, int *d_info, double rel_tol, hipStream_t s);
void diag_h(const double *V, const double *Tx, const double *Ty, const double *Tz, int nx,
            int ny, int nz, double *D, hipStream_t s);
void to_f32(const double *x, float *y, long long n, hipStream_t s);
void to_f64(const float *x, double *y, long long n, hipStream_t s);
// x[i] = uniform in [-1, 1) from a counter-based hash of (seed, i)
void fill_uniform(void *x, long long n, bool f64, unsigned seed, hipStream_t s);
// dst = src, bytes a multiple of 16 (copy_probe_kernel)
void copy_probe(const void *src, void *dst, long long bytes, hipStream_t s);
// read-only stream over `bytes` (a multiple of 16), nontemporal loads if
// asked (stream_probe_kernel); sink: one unsigned of device memory
void stream_probe(const void *src, long long bytes, bool nontemporal, unsigned *sink, hipStream_t s);

// Symmetric eigensolver on the host for the projected matrix (m <= 64):
// cyclic Jacobi, ascending eigenvalues; reads the upper triangle of A
// (column-major, lda).  Replaces LAPACK dsyev at src/Davidson.cpp:224.
void syev_host(int n, const double *A, int lda, double *W, double *V, int ldv);

}  // namespace pisces
