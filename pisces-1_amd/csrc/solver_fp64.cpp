// solver_fp64.cpp -- the plain FP64 Davidson on the GPU.
//
// Follows the reference's reverse-communication Davidson step by step
// (src/Davidson.cpp:61-376; driver src/davdriver.cpp:28-96): the same start
// orthonormalisation, projected matrix S = Z^T B, Ritz vectors renormalised,
// residuals, convergence by ||r|| < 10^-ptol, locking of converged roots by
// their weight on the converged space, restart when the subspace is full or a
// root converges, one correction vector per unconverged root per iteration.
// All vectors stay on the device; the host only sees the m x m projected matrix
// and the per-root norms.  Differences from the reference, all at rounding
// level: S is extended by its new columns instead of recomputed, the
// orthogonalisation of the corrections is block classical Gram-Schmidt done
// twice plus a Cholesky QR of the new block (the reference: modified
// Gram-Schmidt twice, vector by vector), and the small eigenproblem is solved
// by Jacobi rotations instead of LAPACK dsyev.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"

namespace pisces {

namespace {

struct Dev {
    pisces_ctx *c;
    hipStream_t s;
    long long N;
    double *B, *Z, *U, *ZV, *R, *C, *J, *diag, *part, *small;
    int *info;
    size_t part_cap;
    double *hs;  // pinned host scratch
};

// device small-buffer layout (doubles)
constexpr int kS = 0;          // S columns: 64*64
constexpr int kW = 4096;       // W / V block: 64*64
constexpr int kH = 8192;       // CGS coefficients: 64*16
constexpr int kG = 9216;       // Gram of the new block: 16*16
constexpr int kWc = 9472;      // CholQR factor 16*16
constexpr int kScale = 9728;   // per-root scale (64)
constexpr int kLam = 9792;     // per-root lambda (64)
constexpr int kNrm = 9856;     // per-root norms^2 (64*64 Gram)
constexpr int kDots = 13952;   // JD dots (2*16)
constexpr int kSrc = 14000;    // int indices (as ints, 64)
constexpr int kSmallDoubles = 16384;

void d2h(Dev &d, double *h, const double *dv, size_t n)
{
    PC_HIP(hipMemcpyAsync(h, dv, n * sizeof(double), hipMemcpyDeviceToHost, d.s));
}
void h2d(Dev &d, double *dv, const double *h, size_t n)
{
    PC_HIP(hipMemcpyAsync(dv, h, n * sizeof(double), hipMemcpyHostToDevice, d.s));
}
void dcopy(Dev &d, double *dst, const double *src, long long n)
{
    if (dst != src)
        PC_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, d.s));
}

// Orthonormalise the k columns of C against B[0..nb) (two CGS passes) and
// among themselves (Cholesky QR, twice); the result goes to dst.  Returns a
// keep mask; dependent columns are dropped by the caller.
int ortho_block(Dev &d, int nb, int k, double *dst, std::vector<int> &keep)
{
    const long long N = d.N;
    for (int pass = 0; pass < 2 && nb > 0; ++pass) {
        tsgram(d.B, N, nb, d.C, N, k, N, d.part, d.part_cap, d.small + kH, nb, d.s);
        combine(d.B, N, nb, d.small + kH, nb, k, d.C, N, N, 1, d.s);
    }
    double *src = d.C, *out = dst;
    int bad = 0;
    keep.assign(k, 1);
    for (int pass = 0; pass < 2; ++pass) {
        tsgram(src, N, k, src, N, k, N, d.part, d.part_cap, d.small + kG, k, d.s);
        cholinv(d.small + kG, k, d.small + kWc, d.info, 1e-12, d.s);
        // second pass writes back to C and is then copied to dst
        double *o = (pass == 0) ? out : d.J;
        combine(src, N, k, d.small + kWc, k, k, o, N, N, 0, d.s);
        PC_HIP(hipMemcpyAsync(d.hs, d.info, sizeof(int) * (k + 1), hipMemcpyDeviceToHost, d.s));
        PC_HIP(hipStreamSynchronize(d.s));
        const int *hi = reinterpret_cast<const int *>(d.hs);
        bad += hi[0];
        for (int t = 0; t < k; ++t)
            if (!hi[1 + t]) keep[t] = 0;
        src = o;
    }
    dcopy(d, dst, d.J, N * k);
    return bad;
}

}  // namespace

int solve_fp64(pisces_ctx *c, int nroots, int maxsub, int maxmacro, int ptol, int corrflag,
               double *vecs, double *evals, pisces_stats *st, const Fp64Hooks *hooks)
{
    const bool ext = hooks && hooks->matvec;
    PC_REQUIRE(ext || c->has_v, PISCES_ERR_STATE, "no potential set");
    PC_REQUIRE(!ext || hooks->host_diag, PISCES_ERR_ARG, "matrix-free solve needs diag(H)");
    PC_REQUIRE(nroots >= 1 && nroots <= 16 && maxsub >= nroots + 1 && maxsub <= 64,
               PISCES_ERR_ARG, "need 1 <= nroots <= 16 and nroots < maxsub <= 64");
    PC_REQUIRE(corrflag >= 0 && corrflag <= 2, PISCES_ERR_ARG, "corrflag must be 0, 1 or 2");
    auto t0 = std::chrono::steady_clock::now();
    const long long N = c->ngp;
    c->wB.ensure(sizeof(double) * N * maxsub);
    c->wZ.ensure(sizeof(double) * N * maxsub);
    c->wU.ensure(sizeof(double) * N * nroots);
    c->wZV.ensure(sizeof(double) * N * nroots);
    c->wR.ensure(sizeof(double) * N * nroots);
    c->wC.ensure(sizeof(double) * N * nroots);
    c->wJ.ensure(sizeof(double) * N * nroots);
    c->wdiag.ensure(sizeof(double) * N);
    const size_t part_cap = (size_t)((N + 31) / 32 + 1) * 64 * 16;
    c->wpart.ensure(sizeof(double) * part_cap);
    c->wsmall.ensure(sizeof(double) * kSmallDoubles);
    c->winfo.ensure(sizeof(int) * 64);
    Dev d{c, c->stream, N, c->wB.as<double>(), c->wZ.as<double>(), c->wU.as<double>(),
          c->wZV.as<double>(), c->wR.as<double>(), c->wC.as<double>(), c->wJ.as<double>(),
          c->wdiag.as<double>(), c->wpart.as<double>(), c->wsmall.as<double>(),
          c->winfo.as<int>(), part_cap, c->hsmall};

    PC_HIP(hipMemcpyAsync(d.B, vecs, sizeof(double) * N * nroots, hipMemcpyHostToDevice, d.s));
    if (ext)
        PC_HIP(hipMemcpyAsync(d.diag, hooks->host_diag, sizeof(double) * N, hipMemcpyHostToDevice,
                              d.s));
    else
        diag_h(c->dV64.as<double>(), c->dT64[0].as<double>(), c->dT64[1].as<double>(),
               c->dT64[2].as<double>(), c->n[0], c->n[1], c->n[2], d.diag, d.s);

    // GramSchmidt of the start vectors (src/Davidson.cpp:482-514)
    std::vector<int> keep;
    dcopy(d, d.C, d.B, N * nroots);
    ortho_block(d, 0, nroots, d.B, keep);
    for (int t = 0; t < nroots; ++t)
        PC_REQUIRE(keep[t], PISCES_ERR_NUMERIC, "start vectors are linearly dependent");

    const double thresh = std::pow(10.0, -ptol);
    std::vector<double> S((size_t)maxsub * maxsub, 0.0), Vv((size_t)maxsub * maxsub), sse(maxsub);
    std::vector<double> resid(nroots, 0.0), cw(nroots, 0.0);
    std::vector<int> ccode(nroots, 2), bidx(nroots, 0);
    int nmtx = 0, nconv = 0, jmacro = 1, nsub = 0, nnewB = nroots, nnewZ = 0, jvalid = 0;
    int iters = 0;
    bool converged = false, force_restart = false;
    double *hs = c->hsmall;

    for (;;) {
        nnewZ = 0;
        if (nnewB > 0) {  // the reverse interface's "return 1": Z = H B (new columns)
            if (ext) hooks->matvec(d.B + N * nsub, d.Z + N * nsub, nsub, nnewB);
            else c->hx(d.B + N * nsub, d.Z + N * nsub, N, nnewB, PISCES_FP64);
            nmtx += nnewB;
            nnewZ = nnewB;
            nnewB = 0;
        }
        ++iters;
        const int m = nsub + nnewZ;
        // new columns of S = Z^T B: S(i, j) = Z_i . B_j for j in [jvalid, m)
        tsgram(d.Z, N, m, d.B + N * jvalid, N, m - jvalid, N, d.part, d.part_cap, d.small + kS,
               m, d.s);
        d2h(d, hs, d.small + kS, (size_t)m * (m - jvalid));
        PC_HIP(hipStreamSynchronize(d.s));
        for (int j = jvalid; j < m; ++j)
            for (int i = 0; i < m; ++i) S[i + (size_t)maxsub * j] = hs[i + (size_t)m * (j - jvalid)];
        nsub = m;
        jvalid = m;
        syev_host(nsub, S.data(), maxsub, sse.data(), Vv.data(), maxsub);
        for (int r = 0; r < nroots; ++r) {
            double s2 = 0;
            for (int k = 0; k < nconv; ++k) s2 += Vv[k + (size_t)maxsub * r] * Vv[k + (size_t)maxsub * r];
            cw[r] = nconv > 0 ? std::sqrt(s2) : 0.0;
        }
        // Ritz vectors and Z v (src/Davidson.cpp:240-259)
        for (int r = 0; r < nroots; ++r)
            for (int k = 0; k < nsub; ++k) hs[k + (size_t)nsub * r] = Vv[k + (size_t)maxsub * r];
        h2d(d, d.small + kW, hs, (size_t)nsub * nroots);
        combine(d.B, N, nsub, d.small + kW, nsub, nroots, d.U, N, N, 0, d.s);
        combine(d.Z, N, nsub, d.small + kW, nsub, nroots, d.ZV, N, N, 0, d.s);
        tsgram(d.U, N, nroots, d.U, N, nroots, N, d.part, d.part_cap, d.small + kNrm, nroots, d.s);
        d2h(d, hs, d.small + kNrm, (size_t)nroots * nroots);
        PC_HIP(hipStreamSynchronize(d.s));
        for (int r = 0; r < nroots; ++r) {
            const double nu = std::sqrt(hs[r + (size_t)nroots * r]);
            hs[4096 + r] = 1.0 / nu;
            hs[4096 + 64 + r] = sse[r];
        }
        h2d(d, d.small + kScale, hs + 4096, 64 + nroots);
        residual(d.U, d.ZV, d.R, N, N, nroots, d.small + kScale, d.small + kLam, d.s);
        tsgram(d.R, N, nroots, d.R, N, nroots, N, d.part, d.part_cap, d.small + kNrm, nroots, d.s);
        d2h(d, hs, d.small + kNrm, (size_t)nroots * nroots);
        PC_HIP(hipStreamSynchronize(d.s));
        for (int r = 0; r < nroots; ++r) resid[r] = std::sqrt(hs[r + (size_t)nroots * r]);

        // convergence bookkeeping (src/Davidson.cpp:272-297)
        int nnewconv = 0;
        for (int r = 0; r < nroots; ++r) {
            if (cw[r] > 0.95) {
                ccode[r] = 0;
            } else if (resid[r] < thresh) {
                ccode[r] = 1;
                bidx[r] = nconv;
                dcopy(d, d.B + N * nconv, d.U + N * r, N);
                ++nconv;
                ++nnewconv;
            } else {
                ccode[r] = 2;
            }
        }
        if (nconv == nroots) {
            for (int r = 0; r < nroots; ++r) evals[r] = sse[r];
            PC_HIP(hipMemcpyAsync(vecs, d.U, sizeof(double) * N * nroots, hipMemcpyDeviceToHost, d.s));
            PC_HIP(hipStreamSynchronize(d.s));
            converged = true;
            break;
        }
        // restart (src/Davidson.cpp:303-346)
        if (nsub == maxsub || nnewconv > 0 || force_restart) {
            force_restart = false;
            nnewB = 0;
            for (int r = 0; r < nroots; ++r)
                if (ccode[r] == 2) {
                    dcopy(d, d.B + N * (nnewB + nconv), d.U + N * r, N);
                    bidx[r] = nconv + nnewB;
                    ++nnewB;
                }
            for (int r = 0; r < nroots; ++r)
                if (ccode[r] > 0) dcopy(d, d.Z + N * bidx[r], d.ZV + N * r, N);
            // the reference builds these Z columns inside its Ritz-vector array,
            // which the Jacobi-Davidson correction then reads (src/Davidson.cpp:319-363)
            if (corrflag == 2)
                for (int r = 0; r < nroots; ++r)
                    if (ccode[r] > 0) dcopy(d, d.U + N * bidx[r], d.ZV + N * r, N);
            nnewB = 0;
            nsub = nroots;
            jvalid = 0;
            jmacro += 1;
            if (jmacro >= maxmacro) break;
        }
        // correction vectors, one per unconverged root (src/Davidson.cpp:353-371)
        std::vector<int> roots;
        for (int r = 0; r < nroots; ++r) {
            if (nsub + (int)roots.size() == maxsub) break;
            if (ccode[r] == 2) roots.push_back(r);
        }
        const int k = (int)roots.size();
        if (k == 0) continue;
        int *hsrc = reinterpret_cast<int *>(hs + 8192);
        for (int t = 0; t < k; ++t) {
            hsrc[t] = roots[t];
            hs[4096 + t] = sse[roots[t]];
        }
        PC_HIP(hipMemcpyAsync(d.small + kSrc, hsrc, sizeof(int) * k, hipMemcpyHostToDevice, d.s));
        h2d(d, d.small + kLam, hs + 4096, k);
        const int *dsrc = reinterpret_cast<const int *>(d.small + kSrc);
        if (corrflag == 1) {
            precond(d.R, N, d.diag, d.C, N, N, k, dsrc, d.small + kLam, d.s);
        } else if (corrflag == 2) {
            jd_precond(d.R, N, d.U, d.diag, d.C, d.J, N, N, k, dsrc, dsrc, d.small + kLam, d.s);
            // f = (u . c) / (u . j) per vector
            for (int t = 0; t < k; ++t) {
                tsgram(d.U + N * roots[t], N, 1, d.C + N * t, N, 1, N, d.part, d.part_cap,
                       d.small + kDots + 2 * t, 1, d.s);
                tsgram(d.U + N * roots[t], N, 1, d.J + N * t, N, 1, N, d.part, d.part_cap,
                       d.small + kDots + 2 * t + 1, 1, d.s);
            }
            jd_finish(d.C, d.J, N, N, k, d.small + kDots, d.s);
        } else {
            for (int t = 0; t < k; ++t) dcopy(d, d.C + N * t, d.R + N * roots[t], N);
        }
        std::vector<int> kp;
        ortho_block(d, nsub, k, d.B + N * nsub, kp);
        int kept = 0;
        for (int t = 0; t < k; ++t)
            if (kp[t]) {
                if (kept != t) dcopy(d, d.B + N * (nsub + kept), d.B + N * (nsub + t), N);
                ++kept;
            }
        nnewB = kept;
        // every correction fell into the current subspace: restart from the
        // Ritz vectors next time so the iteration cannot stall
        if (kept == 0) force_restart = true;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->nconv = nconv;
        st->n_matvec = nmtx;
        st->n_iter = iters;
        st->n_macro = jmacro;
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        for (int r = 0; r < nroots && r < 64; ++r) st->resid[r] = resid[r];
    }
    if (!converged) {
        // maxmacro exhausted: like the reference, leave evals untouched and
        // return B's first columns (converged vectors first, then the current
        // Ritz vectors of the unconverged roots; src/Davidson.cpp:303-341)
        PC_HIP(hipMemcpyAsync(vecs, d.B, sizeof(double) * N * nroots, hipMemcpyDeviceToHost, d.s));
        PC_HIP(hipStreamSynchronize(d.s));
    }
    return nconv;
}

}  // namespace pisces
