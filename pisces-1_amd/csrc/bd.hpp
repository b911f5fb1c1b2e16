// bd.hpp -- batched kernels of the block-Davidson engine.
//
// Every launch covers a list of independent tasks (one or more per problem
// instance); blockIdx.y selects the task.  Vectors are FP64 or FP32 columns
// (a32/b32/y32 flags, uniform per task), accumulation is always FP64.
#pragma once

#include "common.hpp"

namespace pisces {

constexpr int kBdMaxM = 64;   // max subspace dimension (block + corrections)
constexpr int kBdMaxK = 16;   // max block size
constexpr int kGramMaxB = 16; // max B columns of one Gram task (A: kBdMaxM)

// partial Gram sums of A^T B over one chunk of rows; out[a + out_ld * b]
struct GramTask {
    const void *A; long long lda; int ma; int a32;
    const void *B; long long ldb; int mb; int b32;
    long long N;
    double *out; int out_ld;
    long long part_off;  // this task's region of the partial buffer (doubles)
    int nch;             // row chunks (gram_chunks(N))
    int cpb;             // workgroups per chunk (gram_blocks_per_chunk(ma, mb))
    int kind;            // 0: the configured variant; 1: narrow-B VALU kernel (gram_narrow)
    const int *skip;     // optional: the task does nothing when *skip != 0 (set by an earlier launch)
};

// Y (set|-=|+=) A1 W1 + A2 W2 + A3 W3, W in device memory (column-major);
// the third source is optional (ma3 = 0).  Y2 (optional, mode 0 only, leading
// dimension ldy) receives a second copy of Y, FP32 if y2_32.  Every row is
// read in full before it is written, so Y may be one of the sources (the
// combination is then done in place).  `skip` (optional) is a flag written by
// an earlier launch (orth_kernel) that turns the task into a no-op.
struct CombTask {
    const void *A1; long long lda1; int ma1; int a1_32; const double *W1; int ldw1;
    const void *A2; long long lda2; int ma2; int a2_32; const double *W2; int ldw2;
    void *Y; long long ldy; int y32; int mb; int mode;
    long long N;
    const void *A3; long long lda3; int ma3; int a3_32; const double *W3; int ldw3;
    void *Y2; int y2_32; int pad;
    const int *skip;  // optional: the task does nothing when *skip != 0 (set by an earlier launch)
};

// Ritz vectors and residuals of one instance in one pass over the basis:
//   U = X Y_X + C Y_C,  HU = HX Y_X + HC Y_C,  R = HU - U diag(theta),
// plus partial sums of ||R_i||^2 (reduced into out, and into gather as in
// ResTask).  X, HX, U, HU FP64; C, HC FP32 if c32; R FP32 if r32.  U and HU
// are stored only as `store` asks (kRitzU | kRitzHU); R always.
constexpr int kRitzU = 1, kRitzHU = 2;
struct RitzTask {
    const double *X, *HX; const void *C, *HC;
    int k, mc, c32, r32;
    const double *Y; const double *theta;
    double *U, *HU; void *R;
    long long N;
    long long part_off; int nch; int store;
    double *out; double *gather;
};

// Factor of one orthogonalisation round of a new block T (nt columns) against
// an orthonormal basis V (m columns), from h = V^T T (m x nt, leading
// dimension hld) and G = T^T T (nt x nt, leading dimension gld): S = G - h^T h
// = L L^T (columns whose pivot falls below tol * G_jj, or that keep_in marks
// dead, are dropped), W = L^{-T}, F = -h W (m x nt, ld m); then (T - V h) W =
// T W + V F has orthonormal columns orthogonal to V.
// skip (optional): set to 1 when every column is kept and |W - I|, |F| stay
// below skip_tol everywhere (T W + V F then equals T to well below the
// rounding of FP32 columns, so the combination may be left out), else 0.
struct OrthTask {
    const double *h; const double *G; int m; int nt;
    double *W; double *F;
    const int *keep_in; int *keep_out; int *keep_copy;
    double tol;
    int hld, gld;
    int *skip; double skip_tol;
    const int *noop;  // optional: the task does nothing when *noop != 0 (set by an earlier launch)
    int *clear;       // optional: set to 0 (a flag of orth2 that this task's path replaces)
    // optional: h and G are the row-chunk partials of the Gram [V | T]^T T
    // (GramTask layout, (m + nt) x nt entries) and are summed here; the Gram
    // launch then needs no reduction of its own
    const double *part; long long part_off; int nch; int pad;
};

// Both orthogonalisation rounds of a new block T (nt columns) against the
// stored basis V (m columns) from one Gram: h0 = V^T T and G0 = T^T T come as
// the row-chunk partials of [V | T]^T T (GramTask layout), S = V^T V (the
// exact overlaps of the stored columns, leading dimension lds; only its upper
// triangle is read) gives the
// second round without a pass over the data:
//   round 1: G0 - h0^T h0 = L1 L1^T (dropping, tol1), W1 = L1^{-T}, F1 = -h0 W1;
//   T1 = T W1 + V F1 has V^T T1 = h1 = h0 W1 + S F1 and
//   T1^T T1 = G1 = W1^T G0 W1 + P^T F1 + F1^T P + F1^T S F1 (P = h0 W1);
//   round 2: G1 - h1^T h1 = L2 L2^T (dropping, tol2, after round 1's), W2 = L2^{-T};
// out: W = W1 W2 (nt x nt, ld nt) and F = (F1 - h1) W2 (m x nt, ld m), so that
// one combination T W + V F gives the block of both rounds.
// Where a kept column's |T|^2 / |T - V h0|^2 exceeds amp2_max (round 1 then
// cancels digits that only a second pass over the data restores), the task
// writes round 1 alone (W1, F1, its mask) and *alg = 0; otherwise the merged
// factors, the final mask and *alg = 1, *skip2 = 1 (the caller's second-round
// launches read them and do nothing).
// With *alg = 1 the task also writes S's columns m .. m + nt - 1 (rows
// 0 .. m + nt - 1, upper triangle) for the new block from the same algebra, so
// the caller need not form them from the data.
struct Orth2Task {
    const double *part; long long part_off; int nch; int m;
    const double *Sv; int lds; int nt;
    double *W; double *F;
    int *keep_out; int *keep_copy;
    double tol1, tol2, amp2_max;
    int *alg; int *skip2;
    double *Sout;  // S itself (its new columns)
    int *ndata;    // optional: incremented where the second round is left to the data
};
void bd_orth2(const Orth2Task *d_tasks, int ntasks, hipStream_t s);

// Symmetric eigenproblem of the projected matrix: G (m x m, upper triangle,
// leading dimension ldg) -> lowest k eigenvalues theta and eigenvectors Y (m x k,
// leading dimension m).  Columns flagged dead (keep[c] == 0 for correction
// column c) are excluded.
struct EigTask {
    const double *G; int ldg; int m; int k;
    double *theta; double *Y;
    const int *keep;  // keep mask of all m columns (may be null)
    int mwarm, pad;
    // full eigenvector matrix (m x m, ld kBdMaxM, columns unsorted), written
    // at the end; if mwarm > 0 its leading mwarm x mwarm block (the last solve
    // on the first mwarm columns of this basis) starts the rotations
    double *Vfull;
};

// R = HU - theta U for k columns (theta[i * th_ld]; R FP32 if r32), and
// partial sums of ||R_i||^2
struct ResTask {
    const double *U; const double *HU; void *R; long long ld; int k; int r32;
    const double *theta; long long N; long long part_off; int nch; int th_ld;
    double *out;     // ||R_i||^2 after reduction
    double *gather;  // host-bound copy: theta[0..16) then ||R_i||^2 [16..32)
};

// T[t] = R[src[t]] / (theta[src[t]] - D), |.| >= 1e-12 (Davidson correction)
struct PrecTask {
    const double *R; long long ld; const double *theta; const double *D;
    void *T; long long N; int nt; int src[kBdMaxK]; int t32;  // T FP32 if t32
};

// W = L^{-T} of the k x k Gram G (CholQR).  keep_out[c] = 0 marks a column
// whose remaining norm is below tol (relative): its column of W is zero.
// keep_in (may be null = all kept) carries the mask of an earlier pass.
struct CholTask {
    const double *G; int k; int pad; double *W; const int *keep_in; int *keep_out; double tol;
    int *keep_copy;  // optional second destination of keep_out
};

// max_blocks = max over tasks of nch; max_mamb = max over tasks of ma * mb;
// requires ma <= kBdMaxM and mb <= 16
// Search directions of a restart in the coefficients of the old basis [X|C]:
// Q = orth([0; Y_C] against Y), Y the m x k Ritz coefficients (ld m) whose
// first k rows belong to X.  Q (m x k, ld m) has orthonormal columns
// orthogonal to those of Y; a column whose direction is numerically zero is
// zeroed with keep_out[c] = 0.  Working on the coefficients avoids forming
// the difference of two nearly equal Ritz blocks in the full space.
struct CoefTask {
    const double *Y; int m; int k; double *Q; int *keep_out;
};

// max_blocks / max_blocks_narrow: max over the tasks of kind 0 / 1 of nch * cpb;
// reduce = false leaves the partials to the consumer (OrthTask::part)
void bd_gram(const GramTask *d_tasks, int ntasks, int max_blocks, int max_blocks_narrow,
             int max_mamb, double *d_partial, hipStream_t s, bool reduce = true);
// max_mb: max over the tasks of mb
void bd_combine(const CombTask *d_tasks, int ntasks, long long max_N, int max_mb, hipStream_t s);
void bd_eig(const EigTask *d_tasks, int ntasks, hipStream_t s);
void bd_resid(const ResTask *d_tasks, int ntasks, int max_nch, double *d_partial, hipStream_t s);
void bd_precond(const PrecTask *d_tasks, int ntasks, long long max_N, hipStream_t s);
void bd_cholinv(const CholTask *d_tasks, int ntasks, hipStream_t s);
// k: the (common) block size of the tasks
void bd_ritz(const RitzTask *d_tasks, int ntasks, int max_nch, int k, double *d_partial,
             hipStream_t s);
void bd_orth(const OrthTask *d_tasks, int ntasks, hipStream_t s);
void bd_coef(const CoefTask *d_tasks, int ntasks, hipStream_t s);

constexpr int kGRows = 4096;        // rows per Gram workgroup, matrix-core variants
constexpr int kGTile = 8;           // VALU variant: 8 x 8 tiles of (A, B) columns
constexpr int kGRowsValu = 256 * 32; // VALU variant: rows per chunk (32 per thread)
// Narrow-B tasks (mb <= kGNarrowB) run on the VALU with kGNarrowA x kGNarrowB
// register tiles, kGRowsNarrow rows per workgroup (PISCES_GRAM_NARROW=0 sends
// them to the configured variant instead).  4096 rows: a launch's last round
// of workgroups is less empty than with 8192 (served (H2O)24 64^3 batch on
// three lanes: 544 against 535 solves/s; 2048 rows: 538).
constexpr int kGNarrowA = 8, kGNarrowB = 4, kGRowsNarrow = 256 * 16;
bool gram_narrow(int mb);
// Gram kernel variant (0 VALU register tiles, 1 matrix cores with 16-row
// blocks, 2 matrix cores with 64-row runs) and rows per workgroup, and the
// task sizing they imply.
void gram_set(int variant, int rows);
// loads the configuration once per process (PISCES_GRAM, tuning.txt or default)
void gram_load_config();
int gram_config(int *rows);
int gram_variant();
int gram_chunks(long long N);
int gram_blocks_per_chunk(int ma, int mb);
constexpr int kBdResChunk = 1024; // rows per residual chunk
// rows per Ritz workgroup for block size k (RitzTask::nch = ceil(N / ritz_rows(k)))
int ritz_rows(int k);

}  // namespace pisces
