/*
 * pisces_oracle.c -- CPU restatement of the PISCES excess-electron eigensolve
 * path (sine-DVR H*x, DPP electron-water potential, rigid DPP water model with
 * self-consistent induced dipoles, reverse-communication Davidson).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library's parity tests and
 * the CPU baseline of the benchmark.  Nothing in the product links this file.
 *
 * Each function names the reference lines it restates.  The arithmetic keeps
 * the reference's operation order where that is cheap (so FP64 results agree to
 * a few ulps); it is written from the algorithm, not copied.
 */
#include "pisces_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants as the reference defines them.  PI is the 15-digit value of
 * src/constants.h:4 and src/sine_dvr.cpp:24 (not M_PI): keeping it keeps the
 * T matrices and the Euler rotations identical to the reference. */
#define OR_PI 3.14159265358979
#define OR_BOHR2ANGS 0.52917720859                 /* src/constants.h:17 */
#define OR_ANGS2BOHR (1.0 / OR_BOHR2ANGS)
/* Rigid TTM/DPP monomer, src/ttm_data.h:8-58 */
#define OR_TTM_QH 0.5742185854046506
#define OR_TTM_OZ 0.0650980307
#define OR_TTM_HZ (-0.5207842453)
#define OR_TTM_HY 0.756950327
#define OR_DPP_DELTAM 0.25
/* DPP damping constants of the water model (chargeselector 2),
 * src/Water.cpp:241-243: Damping[1] (charge-dipole), Damping[2] (dipole-dipole) */
#define OR_DAMP_CD 0.21
#define OR_DAMP_DD 0.3

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
static int nthreads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ======================================================================= a1 */
static void sine_dvr_impl(int n, double mass, double x0, double xnp1, double *x,
                          double *T, int packed)
{
    /* src/sine_dvr.cpp:47-111 */
    double dx = (xnp1 - x0) / (double)(n + 1);
    for (int j = 0; j < n; ++j) x[j] = x0 + dx * (double)(j + 1);
    double f1 = 0.5 * OR_PI * OR_PI / (mass * dx * dx);
    double f2 = OR_PI / (double)(n + 1);
    double f3 = (double)(n + 1) * (double)(n + 1);
    for (int a = 0; a < n; ++a) {
        double s = sin((a + 1) * f2);
        double taa = f1 * (1.0 / 3.0 + 1.0 / (6.0 * f3) - 0.5 / (f3 * (s * s)));
        if (packed) T[a * (a + 1) / 2 + a] = taa;
        else T[a + n * a] = taa;
    }
    for (int a = 0; a < n; ++a) {
        for (int b = 0; b < a; ++b) {
            double dc = cos(f2 * (a + 1)) - cos(f2 * (b + 1));
            double tab = f1 * 2.0 / f3 * sin(f2 * (a + 1)) * sin(f2 * (b + 1)) / (dc * dc);
            if ((a - b) % 2 == 1) tab = -tab;
            if (packed) {
                T[a * (a + 1) / 2 + b] = tab;
            } else {
                T[a + n * b] = tab;
                T[b + n * a] = tab;
            }
        }
    }
}

void or_sine_dvr(int n, double mass, double x0, double xnp1, double *x, double *T)
{
    sine_dvr_impl(n, mass, x0, xnp1, x, T, 0);
}

void or_sine_dvr_packed(int n, double mass, double x0, double xnp1, double *x, double *Tp)
{
    sine_dvr_impl(n, mass, x0, xnp1, x, Tp, 1);
}

/* ======================================================================= a3 */
void or_apply_h(int nx, int ny, int nz, const double *Tx, const double *Ty,
                const double *Tz, const double *V, const double *X, double *Y,
                int nvec)
{
    const long ng = (long)nx * ny * nz;
    const long sy = nx, sz = (long)nx * ny;
    for (int v = 0; v < nvec; ++v) {
        const double *x = X + (long)v * ng;
        double *y = Y + (long)v * ng;
        /* y = V o x  (src/mtx.cpp:34-38) */
#pragma omp parallel for num_threads(nthreads()) schedule(static)
        for (long g = 0; g < ng; ++g) y[g] = V[g] * x[g];
        /* x sweep (src/mtx.cpp:58-73): y(:,j,k) += Tx x(:,j,k) */
#pragma omp parallel for collapse(2) num_threads(nthreads()) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) {
                const double *xl = x + j * sy + k * sz;
                double *yl = y + j * sy + k * sz;
                for (int a = 0; a < nx; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < nx; ++b) s += Tx[a + nx * b] * xl[b];
                    yl[a] += s;
                }
            }
        /* y sweep (src/mtx.cpp:77-88): stride nx */
#pragma omp parallel for collapse(2) num_threads(nthreads()) schedule(static)
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                const double *xl = x + i + k * sz;
                double *yl = y + i + k * sz;
                for (int a = 0; a < ny; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < ny; ++b) s += Ty[a + ny * b] * xl[b * sy];
                    yl[a * sy] += s;
                }
            }
        /* z sweep (src/mtx.cpp:92-103): stride nx*ny */
#pragma omp parallel for collapse(2) num_threads(nthreads()) schedule(static)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const double *xl = x + i + j * sy;
                double *yl = y + i + j * sy;
                for (int a = 0; a < nz; ++a) {
                    double s = 0.0;
                    for (int b = 0; b < nz; ++b) s += Tz[a + nz * b] * xl[b * sz];
                    yl[a * sz] += s;
                }
            }
    }
}

/* ======================================================================= a4 */
void or_diag(int nx, int ny, int nz, const double *Tx, const double *Ty,
             const double *Tz, const double *V, double *D)
{
    /* src/DVR.cpp:1808-1881: diag = V, then += Tx_ii, Ty_jj, Tz_kk in turn */
#pragma omp parallel for collapse(2) num_threads(nthreads()) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long g = i + (long)nx * (j + (long)ny * k);
                double d = V[g];
                d += Tx[i + nx * i];
                d += Ty[j + ny * j];
                d += Tz[k + nz * k];
                D[g] = d;
            }
}

/* ====================================================================== a11 */
void or_piab_start(int nx, int ny, int nz, const double *xg, const double *yg,
                   const double *zg, double *wf)
{
    /* src/DVR.cpp:1725-1805: x0 = first grid point, L = last - first; the
     * function is the lowest state of the central half box, zero outside. */
    const double *g[3] = {xg, yg, zg};
    int n[3] = {nx, ny, nz};
    double x0[3], L[3];
    for (int d = 0; d < 3; ++d) {
        x0[d] = g[d][0];
        L[d] = g[d][n[d] - 1] - x0[d];
    }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int ii[3] = {i, j, k};
                double w = 1.0;
                for (int d = 0; d < 3; ++d) {
                    double q = g[d][ii[d]];
                    if (q < x0[d] / 2 || q > -x0[d] / 2) w = 0.0;
                    else w *= sin(OR_PI / L[d] * 2.0 * (q - x0[d] / 2.0));
                }
                wf[i + (long)nx * (j + (long)ny * k)] = w;
            }
}

/* ======================================================================= a8 */
/* Reference (standard-orientation) positions of one monomer relative to its
 * centre of mass, bohr: Water::SetupCharge case 2, src/Water.cpp:216-236. */
static void monomer_ref(double ref[4][3])
{
    const double a2b = OR_ANGS2BOHR;
    ref[0][0] = 0; ref[0][1] = 0;                 ref[0][2] = OR_TTM_OZ * a2b;
    ref[1][0] = 0; ref[1][1] = OR_TTM_HY * a2b;   ref[1][2] = OR_TTM_HZ * a2b;
    ref[2][0] = 0; ref[2][1] = -OR_TTM_HY * a2b;  ref[2][2] = OR_TTM_HZ * a2b;
    ref[3][0] = 0; ref[3][1] = 0;                 ref[3][2] = (OR_TTM_OZ - OR_DPP_DELTAM) * a2b;
}

/* Water::UpdatePositions -> ResetPositions + ApplyEuler, src/Water.cpp:837-943 */
static void monomer_place(const double conf[6], double out[4][3])
{
    double ref[4][3];
    monomer_ref(ref);
    double c1 = cos(conf[3]), c2 = cos(conf[4]), c3 = cos(conf[5]);
    double s1 = sin(conf[3]), s2 = sin(conf[4]), s3 = sin(conf[5]);
    double R[9];
    R[0] = c1 * c2 * c3 - s1 * s3; R[1] = -c1 * c2 * s3 - s1 * c3; R[2] = c1 * s2;
    R[3] = s1 * c2 * c3 + c1 * s3; R[4] = -s1 * c2 * s3 + c1 * c3; R[5] = s1 * s2;
    R[6] = -s2 * c3;               R[7] = s2 * s3;                 R[8] = c2;
    for (int a = 0; a < 4; ++a) {
        /* positions are first stored absolute (origin + ref) and the rotation
         * acts on (position - origin), as the reference does */
        double p0 = conf[0] + ref[a][0], p1 = conf[1] + ref[a][1], p2 = conf[2] + ref[a][2];
        double o0 = p0 - conf[0], o1 = p1 - conf[1], o2 = p2 - conf[2];
        double n0 = R[0] * o0 + R[1] * o1 + R[2] * o2;
        double n1 = R[3] * o0 + R[4] * o1 + R[5] * o2;
        double n2 = R[6] * o0 + R[7] * o1 + R[8] * o2;
        out[a][0] = n0 + conf[0];
        out[a][1] = n1 + conf[1];
        out[a][2] = n2 + conf[2];
    }
}

/* Water::Reposition, src/Water.cpp:695-825 (input in bohr) */
static void monomer_project(const double *w, double conf[6])
{
    double o[3] = {w[0], w[1], w[2]}, h1[3] = {w[3], w[4], w[5]}, h2[3] = {w[6], w[7], w[8]};
    double r1[3], r3[3], z[3], y[3], x[3];
    for (int a = 0; a < 3; ++a) { r1[a] = h1[a] - o[a]; r3[a] = h2[a] - o[a]; }
    for (int a = 0; a < 3; ++a) z[a] = r1[a] + r3[a];
    double zm = sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (int a = 0; a < 3; ++a) { z[a] /= zm; z[a] = -z[a]; }
    double p = r1[0] * z[0] + r1[1] * z[1] + r1[2] * z[2];
    for (int a = 0; a < 3; ++a) y[a] = r1[a] - p * z[a];
    double ym = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    for (int a = 0; a < 3; ++a) y[a] /= ym;
    x[0] = y[1] * z[2] - y[2] * z[1];
    x[1] = y[2] * z[0] - y[0] * z[2];
    x[2] = y[0] * z[1] - y[1] * z[0];
    double phi, theta, psi;
    theta = acos(z[2]);
    double st = sin(theta);
    if (fabs(st) > 1e-5) {
        double tc = -x[2] / st;
        double ac = acos(tc);
        if (tc <= -1) ac = OR_PI;
        psi = (y[2] >= 0) ? ac : 2 * OR_PI - ac;
        tc = z[0] / st;
        if (tc > 1) tc = 1;
        if (tc < -1) tc = -1;
        phi = (z[1] >= 0) ? acos(tc) : 2 * OR_PI - acos(tc);
    } else {
        phi = 0.0;
        psi = (x[1] >= 0) ? acos(y[1]) : 2 * OR_PI - acos(y[1]);
    }
    conf[0] = (h1[0] + h2[0] + 16.0 * o[0]) / 18.0;
    conf[1] = (h1[1] + h2[1] + 16.0 * o[1]) / 18.0;
    conf[2] = (h1[2] + h2[2] + 16.0 * o[2]) / 18.0;
    conf[3] = phi;
    conf[4] = theta;
    conf[5] = psi;
}

int or_water_sites(int nw, const double *xyz_ang, double *sites, double *conf)
{
    for (int w = 0; w < nw; ++w) {
        double b[9], c[6], pos[4][3];
        /* Water::Reposition_A converts to bohr, src/Water.cpp:679-687 */
        for (int a = 0; a < 9; ++a) b[a] = xyz_ang[9 * w + a] * OR_ANGS2BOHR;
        monomer_project(b, c);
        monomer_place(c, pos);
        for (int s = 0; s < 4; ++s)
            for (int a = 0; a < 3; ++a) {
                sites[12 * w + 3 * s + a] = pos[s][a];
                if (!isfinite(pos[s][a])) return -1;
            }
        if (conf)
            for (int a = 0; a < 6; ++a) conf[6 * w + a] = c[a];
    }
    return 0;
}

/* per-site charge, damping polarizability ("sigma") and polarizability,
 * src/Water.cpp:233-236 */
static void site_params(int s, double *q, double *sig, double *alpha)
{
    const double a3 = OR_ANGS2BOHR * OR_ANGS2BOHR * OR_ANGS2BOHR;
    const double polO = 1.22 * a3, polH = 0.28 * a3;
    switch (s) {
    case 0: *q = 0.0; *sig = 0.0; *alpha = polO; break;
    case 1: case 2: *q = OR_TTM_QH; *sig = polH; *alpha = polH; break;
    default: *q = -2.0 * OR_TTM_QH; *sig = polO; *alpha = 0.0; break;
    }
}

/* ElectrostaticDCDamp / ElectrostaticDDDamp, src/DPP.cpp:1798-1812 */
static double damp_cd(double r, double ai, double aj, double a)
{
    double A = pow(ai * aj, 1.0 / 6.0);
    double x = a * pow(r / A, 3);
    return 1.0 - exp(-x);
}
static double damp_dd(double r, double ai, double aj, double a)
{
    double A = pow(ai * aj, 1.0 / 6.0);
    double x = a * pow(r / A, 3);
    return 1.0 - (1 + x) * exp(-x);
}

void or_damp(int n, const double *r, const double *ai, const double *aj, const double *a,
             double *dc, double *dd)
{
    for (int i = 0; i < n; ++i) {
        dc[i] = damp_cd(r[i], ai[i], aj[i], a[i]);
        dd[i] = damp_dd(r[i], ai[i], aj[i], a[i]);
    }
}

int or_water_induced(int nw, const double *sites, double *mu, double *epc,
                     double *e_pol, int *iters)
{
    const int ns = 3 * nw; /* polarizable sites O,H1,H2 per water */
    double *al = malloc(sizeof(double) * ns), *xs = malloc(sizeof(double) * 3 * ns);
    double *f1 = calloc(3 * ns, sizeof(double)), *f2 = malloc(sizeof(double) * 3 * ns);
    double *dp = malloc(sizeof(double) * 3 * ns), *od = malloc(sizeof(double) * 3 * ns);
    for (int w = 0; w < nw; ++w)
        for (int s = 0; s < 3; ++s) {
            double q, sg, a;
            site_params(s, &q, &sg, &a);
            int i = 3 * w + s;
            al[i] = a;
            for (int c = 0; c < 3; ++c) xs[3 * i + c] = sites[12 * w + 3 * s + c];
        }
    /* field of the permanent charges (src/DPP.cpp:830-865), skipping charges of
     * the site's own water */
    for (int w = 0; w < nw; ++w)
        for (int s = 0; s < 4; ++s) {
            double q, sg, a;
            site_params(s, &q, &sg, &a);
            if (q == 0) continue;
            const double *pc = sites + 12 * w + 3 * s;
            for (int i = 0; i < ns; ++i) {
                if (i / 3 == w) continue;
                double rx = xs[3 * i] - pc[0], ry = xs[3 * i + 1] - pc[1], rz = xs[3 * i + 2] - pc[2];
                double t = q * pow(rx * rx + ry * ry + rz * rz, -1.5);
                double dist = sqrt(rx * rx + ry * ry + rz * rz);
                double dmp = damp_cd(dist, al[i], sg, OR_DAMP_CD);
                f1[3 * i] += t * rx * dmp;
                f1[3 * i + 1] += t * ry * dmp;
                f1[3 * i + 2] += t * rz * dmp;
            }
        }
    const double newfac = 0.5, oldfac = 1 - newfac;
    for (int i = 0; i < ns; ++i)
        for (int c = 0; c < 3; ++c) dp[3 * i + c] = newfac * al[i] * f1[3 * i + c];
    /* 0.5-mixed fixed-point iteration (src/DPP.cpp:953-1006) */
    int maxit = 5 * (ns + 10), conv = 0, it;
    double tester = 0;
    for (it = 0; it < maxit; ++it) {
        for (int k = 0; k < 3 * ns; ++k) f2[k] = 0;
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j) {
                if (j == i) continue;
                /* FieldOfDipole(dipole j at site j, evaluated at site i),
                 * src/DPP.cpp:2377-2410 */
                double rx = xs[3 * i] - xs[3 * j], ry = xs[3 * i + 1] - xs[3 * j + 1],
                       rz = xs[3 * i + 2] - xs[3 * j + 2];
                double r2 = rx * rx + ry * ry + rz * rz;
                double r5 = pow(r2, -2.5);
                double d1 = damp_dd(sqrt(r2), al[i], al[j], OR_DAMP_DD);
                double d2 = damp_cd(sqrt(r2), al[i], al[j], OR_DAMP_DD);
                double txx = (d1 * 3. * rx * rx - d2 * r2) * r5;
                double tyy = (d1 * 3. * ry * ry - d2 * r2) * r5;
                double tzz = (d1 * 3. * rz * rz - d2 * r2) * r5;
                double txy = d1 * 3. * rx * ry * r5;
                double txz = d1 * 3. * rx * rz * r5;
                double tyz = d1 * 3. * ry * rz * r5;
                double mx = dp[3 * j], my = dp[3 * j + 1], mz = dp[3 * j + 2];
                f2[3 * i] += txx * mx + txy * my + txz * mz;
                f2[3 * i + 1] += txy * mx + tyy * my + tyz * mz;
                f2[3 * i + 2] += txz * mx + tyz * my + tzz * mz;
            }
        for (int i = 0; i < ns; ++i)
            for (int c = 0; c < 3; ++c) {
                od[3 * i + c] = dp[3 * i + c];
                dp[3 * i + c] = oldfac * od[3 * i + c] + newfac * al[i] * (f1[3 * i + c] + f2[3 * i + c]);
            }
        tester = 0;
        for (int i = 0; i < ns; ++i) {
            double tx = od[3 * i] - dp[3 * i], ty = od[3 * i + 1] - dp[3 * i + 1],
                   tz = od[3 * i + 2] - dp[3 * i + 2];
            tester += tx * tx + ty * ty + tz * tz;
        }
        tester = sqrt(tester) / 3 / ns;
        if (tester < 1e-12) { conv = 1; break; }
    }
    if (iters) *iters = conv ? it + 1 : it;
    if (mu) memcpy(mu, dp, sizeof(double) * 3 * ns);
    if (epc) memcpy(epc, f1, sizeof(double) * 3 * ns);
    if (e_pol) {
        /* sum over pairs of WaterWaterPolarizationEnergy, src/DPP.cpp:1935-2024 */
        double e = 0;
        for (int wi = 0; wi < nw - 1; ++wi)
            for (int wj = wi + 1; wj < nw; ++wj)
                for (int k = 0; k < 3; ++k) { /* polarizable centres O,H1,H2 */
                    double q, sg, a;
                    site_params(k, &q, &sg, &a);
                    const double *x1 = sites + 12 * wi + 3 * k, *x2 = sites + 12 * wj + 3 * k;
                    const double *m1 = dp + 3 * (3 * wi + k), *m2 = dp + 3 * (3 * wj + k);
                    for (int j = 0; j < 4; ++j) { /* W2 polarizing W1 */
                        double q2, s2, a2;
                        site_params(j, &q2, &s2, &a2);
                        if (q2 == 0) continue;
                        const double *p2 = sites + 12 * wj + 3 * j;
                        double rx = x1[0] - p2[0], ry = x1[1] - p2[1], rz = x1[2] - p2[2];
                        double dist = sqrt(rx * rx + ry * ry + rz * rz);
                        double dmp = damp_cd(dist, a, s2, OR_DAMP_CD);
                        e -= q2 * (m1[0] * rx + m1[1] * ry + m1[2] * rz) * dmp *
                             pow(rx * rx + ry * ry + rz * rz, -1.5) / 2.0;
                    }
                    for (int j = 0; j < 4; ++j) { /* W1 polarizing W2 */
                        double q1, s1, a1;
                        site_params(j, &q1, &s1, &a1);
                        if (q1 == 0) continue;
                        const double *p1 = sites + 12 * wi + 3 * j;
                        double rx = x2[0] - p1[0], ry = x2[1] - p1[1], rz = x2[2] - p1[2];
                        double dist = sqrt(rx * rx + ry * ry + rz * rz);
                        double dmp = damp_cd(dist, a, s1, OR_DAMP_CD);
                        e -= q1 * (m2[0] * rx + m2[1] * ry + m2[2] * rz) * dmp *
                             pow(rx * rx + ry * ry + rz * rz, -1.5) / 2.0;
                    }
                }
        *e_pol = e;
    }
    free(al); free(xs); free(f1); free(f2); free(dp); free(od);
    return conv ? 0 : 1;
}

/* Pair terms of the neutral water cluster, src/DPP.cpp:1820-2024, 2256-2303.
 * WaterWaterElectrostaticEnergy: charge-charge, undamped. */
static double ww_elec(const double *s1, const double *s2)
{
    double e = 0;
    for (int i = 0; i < 4; ++i) {
        double q1, g1, a1;
        site_params(i, &q1, &g1, &a1);
        if (q1 == 0) continue;
        for (int j = 0; j < 4; ++j) {
            double q2, g2, a2;
            site_params(j, &q2, &g2, &a2);
            if (q2 == 0) continue;
            double dx = s1[3 * i] - s2[3 * j], dy = s1[3 * i + 1] - s2[3 * j + 1],
                   dz = s1[3 * i + 2] - s2[3 * j + 2];
            e += q1 * q2 / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    return e;
}

/* WaterWaterPolarizationEnergy: induced dipole (O,H1,H2) in the damped field
 * of the other water's charges, halved; m1, m2 the dipoles of the two waters */
static double ww_pol(const double *s1, const double *s2, const double *m1, const double *m2)
{
    double e = 0;
    for (int k = 0; k < 3; ++k) {
        double q, g, a;
        site_params(k, &q, &g, &a);
        const double *x1 = s1 + 3 * k, *x2 = s2 + 3 * k, *d1 = m1 + 3 * k, *d2 = m2 + 3 * k;
        for (int j = 0; j < 4; ++j) { /* W2 polarizing W1 */
            double q2, g2, a2;
            site_params(j, &q2, &g2, &a2);
            if (q2 == 0) continue;
            double rx = x1[0] - s2[3 * j], ry = x1[1] - s2[3 * j + 1], rz = x1[2] - s2[3 * j + 2];
            double dist = sqrt(rx * rx + ry * ry + rz * rz);
            double dmp = damp_cd(dist, a, g2, OR_DAMP_CD);
            e -= q2 * (d1[0] * rx + d1[1] * ry + d1[2] * rz) * dmp *
                 pow(rx * rx + ry * ry + rz * rz, -1.5) / 2.0;
        }
        for (int j = 0; j < 4; ++j) { /* W1 polarizing W2 */
            double q1, g1, a1;
            site_params(j, &q1, &g1, &a1);
            if (q1 == 0) continue;
            double rx = x2[0] - s1[3 * j], ry = x2[1] - s1[3 * j + 1], rz = x2[2] - s1[3 * j + 2];
            double dist = sqrt(rx * rx + ry * ry + rz * rz);
            double dmp = damp_cd(dist, a, g1, OR_DAMP_CD);
            e -= q1 * (d2[0] * rx + d2[1] * ry + d2[2] * rz) * dmp *
                 pow(rx * rx + ry * ry + rz * rz, -1.5) / 2.0;
        }
    }
    return e;
}

/* WaterWaterDispersionEnergy: O-O exponential repulsion plus Tang-Toennies
 * damped C6, H-H and O-H exponential repulsion (DPP parameters of
 * src/ttm_data.h:76-119, Water::SetupCharge case 2, src/Water.cpp:225-235) */
#define OR_AU2KCAL (6.02214179e23 * 4.35974394e-18 / 4.184 / 1000)
static double ww_disp(const double *s1, const double *s2)
{
    const double a2b = OR_ANGS2BOHR;
    const double delta = 2.23 / a2b, c6 = -1300.0000000 / OR_AU2KCAL * pow(a2b, 6);
    const double A1 = 115271.77066806905 / OR_AU2KCAL, b1 = 4.167868773196562 / a2b;
    const double A2 = 119085.35806729361 / OR_AU2KCAL, b2 = 5.581698710292979 / a2b;
    const double A3 = 1803.662375255084 / OR_AU2KCAL, b3 = 3.5354602029164868 / a2b;
    double e = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double dx = s1[3 * i] - s2[3 * j], dy = s1[3 * i + 1] - s2[3 * j + 1],
                   dz = s1[3 * i + 2] - s2[3 * j + 2];
            double R2 = dx * dx + dy * dy + dz * dz;
            double R = sqrt(R2);
            if (i == 0 && j == 0) {
                double R6 = R2 * R2 * R2, dR = delta * R;
                double damp = 1 - (1.0 + dR + pow(dR, 2) / 2.0 + pow(dR, 3) / 6.0 +
                                   pow(dR, 4) / 24.0 + pow(dR, 5) / 120.0 + pow(dR, 6) / 720.0) *
                                      exp(-dR);
                e = e + A1 * exp(-b1 * R) + damp * c6 / R6;
            } else if (i != 0 && j != 0) {
                e = e + A2 * exp(-b2 * R);
            } else {
                e = e + A3 * exp(-b3 * R);
            }
        }
    return e;
}

int or_water_energy(int nw, const double *sites, const double *mu, double *out)
{
    /* WaterCluster::CalcIntermolecularPotential, src/DPP.cpp:1648-1700: the
     * pair loop i < j, E0 = sum(elec + pol) + sum(disp) */
    double nre = 0, lj = 0, el = 0, pol = 0, dis = 0;
    for (int i = 0; i < nw - 1; ++i)
        for (int j = i + 1; j < nw; ++j) {
            const double *s1 = sites + 12 * i, *s2 = sites + 12 * j;
            double e = ww_elec(s1, s2);
            double p = ww_pol(s1, s2, mu + 9 * i, mu + 9 * j);
            double d = ww_disp(s1, s2);
            el += e;
            nre += e + p;
            lj += d;
            pol += p;
            dis += d;
        }
    out[0] = nre + lj;
    out[1] = el;
    out[2] = pol;
    out[3] = dis;
    return 0;
}

int or_water_lists(int nw, const double *sites, const double *mu, double *q,
                   int *iq, double *d, int *id)
{
    (void)sites;
    int nq = 0;
    for (int w = 0; w < nw; ++w)
        for (int s = 0; s < 4; ++s) {
            double qq, sg, a;
            site_params(s, &qq, &sg, &a);
            if (qq != 0) { q[nq] = qq; iq[nq] = 4 * w + s; ++nq; }
        }
    int nd = 0;
    for (int w = 0; w < nw; ++w)
        for (int s = 0; s < 4; ++s) {
            double qq, sg, a;
            site_params(s, &qq, &sg, &a);
            if (a != 0) {
                for (int c = 0; c < 3; ++c) d[3 * nd + c] = mu[3 * nd + c];
                id[nd] = 4 * w + s;
                ++nd;
            }
        }
    return nq;
}

void or_fft_kinetic(int n, double mass, double x0, double xnp1, double *x, double *T,
                    double *ek)
{
    /* Tdiag, src/KE_diag.cpp */
    const double pi = 3.14159265358979323846;
    double dx = (xnp1 - x0) / (double)(n + 1);
    for (int j = 0; j < n; ++j) x[j] = x0 + dx * (double)(j + 1);
    double dk = 2.0 * pi / ((double)n * dx);
    double *e = malloc(sizeof(double) * n);
    int *m = malloc(sizeof(int) * n);
    for (int a = 0; a < n; ++a) {
        m[a] = (a >= n / 2) ? a - n : a;
        e[a] = (dk * m[a]) * (dk * m[a]) / (2.0 * mass);
        if (ek) ek[a] = e[a];
    }
    for (int j = 0; j < n; ++j)
        for (int l = 0; l < n; ++l) {
            double t = 0;
            for (int a = 0; a < n; ++a) t += e[a] * cos(2.0 * pi * (double)m[a] * (double)(j - l) / (double)n);
            T[(long)j * n + l] = t / (double)n;
        }
    free(e);
    free(m);
}

/* ==================================================================== a5 a6 */
typedef struct {
    int damp_type, rep_type, pol_type, adiabatic;
    double qdamp, ddamp, pdamp, alpha, eps;
    int ns, nq, nd, npps, nrep, nion;
    const double *sites, *q, *d, *ion_xyz, *ion_q;
    const int *iq, *id;
    double ion_r;
    int *ipps;
    double *rep;  /* exponent, coefficient */
    int *irep;
    /* PolType 3: np polarizable sites ip[], their charge field epc[3np],
     * InvA (3np x 3np, both triangles from the lower one) and VpolWaterWater */
    int np;
    const int *ip;
    const double *epc;
    double *inva;
    double vpolww;
} or_pot;

static double reff(double r, double r0)
{
    /* effective distance, src/Potential.cpp:1191-1196 */
    if (r < r0) {
        double t = r / r0;
        return r0 * (0.5 + t * t * t * (1.0 - 0.5 * t));
    }
    return r;
}

static int pot_setup(or_pot *P, const int *flags, const double *pp, int nsites)
{
    /* Potential::SetupDPPGTOP, src/Potential.cpp:234-593 */
    int pf = flags[0];
    if (pf < 1 || pf > 4) return -2;
    P->damp_type = flags[1];
    P->pol_type = flags[3];
    P->adiabatic = flags[4];
    if (P->pol_type != 0 && P->pol_type != 1 && P->pol_type != 3) return -2;
    P->alpha = pp[0];
    double vrep = pp[2], vrepO = (flags[2] == 1) ? pp[5] : pp[2];
    if (P->damp_type == 1) {
        P->qdamp = 1.0 / (pp[3] * pp[3]);
        P->ddamp = 1.0 / (pp[4] * pp[4]);
        P->pdamp = 1.0 / (pp[1] * pp[1]);
    } else if (P->damp_type == 2) {
        P->qdamp = fabs(pp[3]);
        P->ddamp = fabs(pp[4]);
        P->pdamp = fabs(pp[1]);
    } else return -2;
    P->eps = pp[6];
    if (nsites % 4) return -3;
    int nw = nsites / 4;
    P->npps = (P->pol_type == 1) ? nw : 0;
    P->ipps = malloc(sizeof(int) * (nw + 1));
    for (int i = 0; i < nw; ++i) P->ipps[i] = 4 * i + 3; /* M sites */
    /* repulsive cores: src/Potential.cpp:598-728 */
    double sH = 1.0 / vrep, sO = 1.0 / vrepO;
    int per = (pf == 1) ? 6 : (pf == 2) ? 12 : (pf == 3) ? 4 : 0;
    P->nrep = per * nw;
    P->rep = malloc(sizeof(double) * 2 * (P->nrep + 1));
    P->irep = malloc(sizeof(int) * (P->nrep + 1));
    P->rep_type = (pf == 3) ? 2 : (pf == 4) ? 0 : 1;
    for (int w = 0; w < nw; ++w) {
        double *g = P->rep + 2 * per * w;
        int *s = P->irep + per * w;
        if (pf == 1) {
            static const double e6[3] = {33.87, 5.095, 1.159};
            static const double c6[3] = {2.84230876640242, 4.55401933663180, 6.91243747981459};
            for (int h = 0; h < 2; ++h)
                for (int k = 0; k < 3; ++k) {
                    s[3 * h + k] = 4 * w + 1 + h;
                    g[2 * (3 * h + k)] = e6[k];
                    g[2 * (3 * h + k) + 1] = c6[k] * (1.0 / vrep);
                }
        } else if (pf == 2) {
            static const double eo[6] = {322.037, 48.4308, 10.4206, 7.40294, 1.5762, 0.373684};
            static const double co[6] = {41.69216064, 58.35946922, 35.19878892, -3.27944113, 4.44843633, 2.36213092};
            static const double eh[3] = {5.447178, 0.824547, 0.183192};
            static const double ch[3] = {0.69233452, 0.66686241, 0.07193908};
            for (int k = 0; k < 6; ++k) { s[k] = 4 * w; g[2 * k] = eo[k]; g[2 * k + 1] = co[k] * sO; }
            for (int h = 0; h < 2; ++h)
                for (int k = 0; k < 3; ++k) {
                    int m = 6 + 3 * h + k;
                    s[m] = 4 * w + 1 + h;
                    g[2 * m] = eh[k];
                    g[2 * m + 1] = ch[k] * sH;
                }
        } else if (pf == 3) {
            s[0] = 4 * w; s[1] = 4 * w; s[2] = 4 * w + 1; s[3] = 4 * w + 2;
            g[0] = 9.98526; g[1] = 192.608 * sO;
            g[2] = 1.67075; g[3] = 13.8937 * sO;
            g[4] = 1.64121; g[5] = 1.7643 * sH;
            g[6] = 1.64121; g[7] = 1.7643 * sH;
        }
    }
    return 0;
}

static void pot_free(or_pot *P)
{
    free(P->ipps); free(P->rep); free(P->irep); free(P->inva);
}

/* EvaluateDPPGTOP at one point with the per-site distance table of
 * Potential::Evaluate (src/Potential.cpp:136-146). */
static double pot_eval(const or_pot *P, const double *r, double *R, double *R2,
                       double *Rx, double *Ry, double *Rz, double *Rm3, double *parts,
                       double *work)
{
    for (int i = 0; i < P->ns; ++i) {
        Rx[i] = r[0] - P->sites[3 * i];
        Ry[i] = r[1] - P->sites[3 * i + 1];
        Rz[i] = r[2] - P->sites[3 * i + 2];
        R2[i] = Rx[i] * Rx[i] + Ry[i] * Ry[i] + Rz[i] * Rz[i];
        R[i] = sqrt(R2[i]);
        Rm3[i] = 1.0 / (R2[i] * R[i]);
    }
    /* charges, src/Potential.cpp:1168-1210 */
    double vpc = 0;
    for (int i = 0; i < P->nq; ++i) {
        int s = P->iq[i];
        if (P->damp_type == 1) vpc += -P->q[i] / R[s] * (1.0 - exp(-P->qdamp * R2[s]));
        else vpc += -P->q[i] / reff(R[s], P->qdamp);
    }
    /* dipoles, src/Potential.cpp:1222-1270 */
    double vind = 0;
    for (int i = 0; i < P->nd; ++i) {
        int s = P->id[i];
        const double *m = P->d + 3 * i;
        double dot = Rx[s] * m[0] + Ry[s] * m[1] + Rz[s] * m[2];
        if (P->damp_type == 1) {
            double dm = 1.0 - exp(-P->ddamp * R2[s]);
            vind += -dot * Rm3[s] * dm * dm;
        } else {
            double rm3 = Rm3[s];
            if (R[s] < P->ddamp) {
                double re = reff(R[s], P->ddamp);
                rm3 = 1.0 / (re * re * re);
            }
            vind += -dot * rm3;
        }
    }
    /* repulsive cores, src/Potential.cpp:1282-1315 */
    double vrep = 0;
    for (int i = 0; i < P->nrep; ++i) {
        int s = P->irep[i];
        if (P->rep_type == 1) vrep += P->rep[2 * i + 1] * exp(-P->rep[2 * i] * R2[s]);
        else if (P->rep_type == 2) vrep += P->rep[2 * i + 1] * exp(-P->rep[2 * i] * R[s]);
    }
    /* polarization, src/Potential.cpp:1331-1374 and :875-887 */
    double vpol = 0;
    if (P->pol_type == 1) {
        double sp = 0;
        for (int i = 0; i < P->npps; ++i) {
            int s = P->ipps[i];
            double rsq = R2[s];
            if (P->damp_type == 1) {
                double dm = 1.0 - exp(-P->pdamp * rsq);
                sp += P->alpha * dm * dm / (rsq * rsq);
            } else {
                if (R[s] < P->pdamp) {
                    double re = reff(R[s], P->pdamp);
                    double re2 = re * re;
                    sp += P->alpha * rsq / (re2 * re2 * re2);
                } else sp += P->alpha / (rsq * rsq);
            }
        }
        if (P->adiabatic) vpol = P->eps - sqrt(P->eps * (P->eps + sp));
        else vpol = -0.5 * sp;
    } else if (P->pol_type == 3) {
        /* SelfConsistentPolarizability, src/Potential.cpp:1495-1592: the
         * electron's field at the polarizable sites (damped R^-3), plus the
         * permanent-charge field, mu = InvA E (dsymv on the lower triangle),
         * Vpol = -E.mu / 2 - VpolWaterWater */
        const int n3 = 3 * P->np;
        double *E = work, *mu = work + n3;
        for (int i = 0; i < P->np; ++i) {
            int s = P->ip[i];
            double g = Rm3[s];
            if (P->damp_type == 1) {
                g *= 1.0 - exp(-P->pdamp * R[s] * R2[s]);
            } else if (R[s] < P->pdamp) {
                double re = reff(R[s], P->pdamp);
                g = 1.0 / (re * re * re);
            }
            E[3 * i] = g * Rx[s] + P->epc[3 * i];
            E[3 * i + 1] = g * Ry[s] + P->epc[3 * i + 1];
            E[3 * i + 2] = g * Rz[s] + P->epc[3 * i + 2];
        }
        for (int a = 0; a < n3; ++a) {
            const double *row = P->inva + (long)n3 * a;
            double acc = 0;
            for (int b = 0; b < n3; ++b) acc += row[b] * E[b];
            mu[a] = acc;
        }
        double em = 0;
        for (int a = 0; a < n3; ++a) em += E[a] * mu[a];
        vpol = -0.5 * em - P->vpolww;
    }
    /* ions (Bloomfield erf-damped charge, src/NaCl.cpp:300-307) */
    double vion = 0;
    for (int i = 0; i < P->nion; ++i) {
        double dx = r[0] - P->ion_xyz[3 * i], dy = r[1] - P->ion_xyz[3 * i + 1],
               dz = r[2] - P->ion_xyz[3 * i + 2];
        double rr = sqrt(dx * dx + dy * dy + dz * dz);
        vion += -P->ion_q[i] * erf(rr / P->ion_r) / rr;
    }
    double vt = vpc + vind + vrep + vpol + vion;
    if (parts) {
        parts[0] = vpc + vion; parts[1] = vind; parts[2] = vrep; parts[3] = vpol; parts[4] = vt;
    }
    return vt;
}

static int pot_init(or_pot *P, const int *flags, const double *pp, int nsites,
                    const double *sites, int nq, const double *q, const int *iq,
                    int nd, const double *d, const int *id, int nion,
                    const double *ion_xyz, const double *ion_q, double ion_r)
{
    memset(P, 0, sizeof(*P));
    int rc = pot_setup(P, flags, pp, nsites);
    if (rc) return rc;
    P->ns = nsites; P->sites = sites;
    P->nq = nq; P->q = q; P->iq = iq;
    P->nd = nd; P->d = d; P->id = id;
    P->nion = nion; P->ion_xyz = ion_xyz; P->ion_q = ion_q; P->ion_r = ion_r;
    return 0;
}

static void pot_points(const or_pot *P, int npts, const double *pts, double *v, double *parts)
{
    const int ns = P->ns, n3 = 3 * P->np;
#pragma omp parallel num_threads(nthreads())
    {
        double *buf = malloc(sizeof(double) * (6 * (ns + 1) + 2 * n3 + 2));
        double *R = buf, *R2 = buf + ns, *Rx = buf + 2 * ns, *Ry = buf + 3 * ns,
               *Rz = buf + 4 * ns, *Rm3 = buf + 5 * ns, *work = buf + 6 * (ns + 1);
#pragma omp for schedule(static)
        for (int p = 0; p < npts; ++p)
            v[p] = pot_eval(P, pts + 3 * p, R, R2, Rx, Ry, Rz, Rm3, parts ? parts + 5 * p : NULL,
                            work);
        free(buf);
    }
}

static void pot_grid(const or_pot *P, int nx, int ny, int nz, const double *xg,
                     const double *yg, const double *zg, double *v)
{
    const int ns = P->ns, n3 = 3 * P->np;
    const long ng = (long)nx * ny * nz;
#pragma omp parallel num_threads(nthreads())
    {
        double *buf = malloc(sizeof(double) * (6 * (ns + 1) + 2 * n3 + 2));
        double *R = buf, *R2 = buf + ns, *Rx = buf + 2 * ns, *Ry = buf + 3 * ns,
               *Rz = buf + 4 * ns, *Rm3 = buf + 5 * ns, *work = buf + 6 * (ns + 1);
#pragma omp for schedule(static)
        for (long g = 0; g < ng; ++g) {
            int i = (int)(g % nx), j = (int)((g / nx) % ny), k = (int)(g / ((long)nx * ny));
            double r[3] = {xg[i], yg[j], zg[k]};
            v[g] = pot_eval(P, r, R, R2, Rx, Ry, Rz, Rm3, NULL, work);
        }
        free(buf);
    }
}

int or_potential_points(const int *flags, const double *potpara, int nsites,
                        const double *sites, int nq, const double *q, const int *iq,
                        int nd, const double *d, const int *id, int nion,
                        const double *ion_xyz, const double *ion_q, double ion_r,
                        int npts, const double *pts, double *v, double *parts)
{
    or_pot P;
    int rc = pot_init(&P, flags, potpara, nsites, sites, nq, q, iq, nd, d, id, nion,
                      ion_xyz, ion_q, ion_r);
    if (!rc && P.pol_type == 3) rc = -2;  /* needs the polarizable-site lists */
    if (rc) { pot_free(&P); return rc; }
    pot_points(&P, npts, pts, v, parts);
    pot_free(&P);
    return 0;
}

int or_potential_grid(const int *flags, const double *potpara, int nsites,
                      const double *sites, int nq, const double *q, const int *iq,
                      int nd, const double *d, const int *id, int nion,
                      const double *ion_xyz, const double *ion_q, double ion_r,
                      int nx, int ny, int nz, const double *xg, const double *yg,
                      const double *zg, double *v)
{
    or_pot P;
    int rc = pot_init(&P, flags, potpara, nsites, sites, nq, q, iq, nd, d, id, nion,
                      ion_xyz, ion_q, ion_r);
    if (!rc && P.pol_type == 3) rc = -2;
    if (rc) { pot_free(&P); return rc; }
    pot_grid(&P, nx, ny, nz, xg, yg, zg, v);
    pot_free(&P);
    return 0;
}

/* ------------------------------------------------------------ PolType 3 */
/* Thole-damped dipole-dipole tensor of a site pair, CalcDDTensor
 * (src/MolPolAux.cpp): T = 3 f5 R R^T / r^5 - f3 I / r^3 with
 * x = a r^3 / sqrt(alpha_i alpha_j), f3 = 1 - e^-x, f5 = 1 - (1 + x) e^-x. */
static void dd_tensor(const double *rij, double ai, double aj, double a, double T[9])
{
    double r2 = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
    double r = sqrt(r2), r3 = r * r2, r5 = r2 * r3;
    double x = a * r3 / sqrt(ai * aj), e = exp(-x);
    double f3 = 1 - e, f5 = 1 - (1 + x) * e;
    for (int u = 0; u < 3; ++u)
        for (int w = 0; w < 3; ++w)
            T[3 * u + w] = 3.0 * f5 * rij[u] * rij[w] / r5 - (u == w ? f3 / r3 : 0.0);
}

int or_pol3_inva(int np, const double *rpol, const double *alpha, double *inva)
{
    /* BuildPolarizationMatrix (src/MolPolAux.cpp:100-136): alpha^-1 on the
     * diagonal 3x3 blocks, -T_ij off the diagonal for every pair (Thole
     * a = 0.3, src/Potential.cpp:425); then the inverse (InvertMatrix uses
     * dsytrf/dsytri; here Gauss-Jordan elimination with partial pivoting). */
    const int n = 3 * np;
    double *A = malloc(sizeof(double) * (size_t)n * n);
    memset(A, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) A[(size_t)n * (3 * i + c) + 3 * i + c] = 1.0 / alpha[i];
    for (int i = 0; i < np; ++i)
        for (int j = i + 1; j < np; ++j) {
            double rij[3] = {rpol[3 * i] - rpol[3 * j], rpol[3 * i + 1] - rpol[3 * j + 1],
                             rpol[3 * i + 2] - rpol[3 * j + 2]};
            double T[9];
            dd_tensor(rij, alpha[i], alpha[j], 0.3, T);
            for (int u = 0; u < 3; ++u)
                for (int w = 0; w < 3; ++w) {
                    A[(size_t)n * (3 * i + u) + 3 * j + w] = -T[3 * u + w];
                    A[(size_t)n * (3 * j + u) + 3 * i + w] = -T[3 * w + u];
                }
        }
    for (long k = 0; k < (long)n * n; ++k) inva[k] = (k % (n + 1) == 0) ? 1.0 : 0.0;
    int rc = 0;
    for (int c = 0; c < n && !rc; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabs(A[(size_t)n * r + c]) > fabs(A[(size_t)n * piv + c])) piv = r;
        if (A[(size_t)n * piv + c] == 0.0) { rc = -1; break; }
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                double t = A[(size_t)n * c + k]; A[(size_t)n * c + k] = A[(size_t)n * piv + k]; A[(size_t)n * piv + k] = t;
                t = inva[(size_t)n * c + k]; inva[(size_t)n * c + k] = inva[(size_t)n * piv + k]; inva[(size_t)n * piv + k] = t;
            }
        double f = 1.0 / A[(size_t)n * c + c];
        for (int k = 0; k < n; ++k) { A[(size_t)n * c + k] *= f; inva[(size_t)n * c + k] *= f; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double m = A[(size_t)n * r + c];
            if (m == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                A[(size_t)n * r + k] -= m * A[(size_t)n * c + k];
                inva[(size_t)n * r + k] -= m * inva[(size_t)n * c + k];
            }
        }
    }
    free(A);
    return rc;
}

static int pol3_init(or_pot *P, int np, const double *alpha, const int *ip, const double *epc,
                     double vpolww)
{
    /* Potential::SetupDPPGTOP case 3 (src/Potential.cpp:387-432): the
     * polarizable sites, their charge field and InvA; the point dipoles of the
     * water model are ignored (:327-335) */
    if (P->pol_type != 3 || np <= 0) return -2;
    for (int i = 0; i < np; ++i)
        if (ip[i] < 0 || ip[i] >= P->ns) return -3;
    const int n = 3 * np;
    double *rp = malloc(sizeof(double) * n);
    for (int i = 0; i < np; ++i)
        for (int c = 0; c < 3; ++c) rp[3 * i + c] = P->sites[3 * ip[i] + c];
    P->inva = malloc(sizeof(double) * (size_t)n * n);
    int rc = or_pol3_inva(np, rp, alpha, P->inva);
    free(rp);
    if (rc) return rc;
    /* dsymv("L") reads the lower triangle only */
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b) P->inva[(size_t)n * a + b] = P->inva[(size_t)n * b + a];
    P->np = np; P->ip = ip; P->epc = epc; P->vpolww = vpolww;
    P->nd = 0;
    return 0;
}

int or_potential_points_pol3(const int *flags, const double *potpara, int nsites,
                             const double *sites, int nq, const double *q, const int *iq,
                             int np, const double *alpha, const int *ip, const double *epc,
                             double vpolww, int npts, const double *pts, double *v,
                             double *parts)
{
    or_pot P;
    int rc = pot_init(&P, flags, potpara, nsites, sites, nq, q, iq, 0, NULL, NULL, 0, NULL,
                      NULL, 1.0);
    if (!rc) rc = pol3_init(&P, np, alpha, ip, epc, vpolww);
    if (rc) { pot_free(&P); return rc; }
    pot_points(&P, npts, pts, v, parts);
    pot_free(&P);
    return 0;
}

int or_potential_grid_pol3(const int *flags, const double *potpara, int nsites,
                           const double *sites, int nq, const double *q, const int *iq,
                           int np, const double *alpha, const int *ip, const double *epc,
                           double vpolww, int nx, int ny, int nz, const double *xg,
                           const double *yg, const double *zg, double *v)
{
    or_pot P;
    int rc = pot_init(&P, flags, potpara, nsites, sites, nq, q, iq, 0, NULL, NULL, 0, NULL,
                      NULL, 1.0);
    if (!rc) rc = pol3_init(&P, np, alpha, ip, epc, vpolww);
    if (rc) { pot_free(&P); return rc; }
    pot_grid(&P, nx, ny, nz, xg, yg, zg, v);
    pot_free(&P);
    return 0;
}

/* ================================================================ a9, a10 */
static double dot(long n, const double *a, const double *b)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s) num_threads(nthreads()) schedule(static)
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double nrm2(long n, const double *a) { return sqrt(dot(n, a, a)); }
static void scal(long n, double f, double *a)
{
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (long i = 0; i < n; ++i) a[i] *= f;
}
static void axpy(long n, double f, const double *x, double *y)
{
#pragma omp parallel for num_threads(nthreads()) schedule(static)
    for (long i = 0; i < n; ++i) y[i] += f * x[i];
}

int or_syev(int n, const double *A, int lda, double *W, double *Vec, int ldv)
{
    /* cyclic Jacobi on a copy symmetrised from the upper triangle */
    double *a = malloc(sizeof(double) * n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            double x = A[i + (long)lda * j];
            a[i + n * j] = x;
            a[j + n * i] = x;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) Vec[i + (long)ldv * j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                tot += a[i + n * j] * a[i + n * j];
                if (i != j) off += a[i + n * j] * a[i + n * j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p + n * q];
                if (apq == 0.0) continue;
                double app = a[p + n * p], aqq = a[q + n * q];
                double th = (aqq - app) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) { /* columns p,q */
                    double akp = a[k + n * p], akq = a[k + n * q];
                    a[k + n * p] = c * akp - s * akq;
                    a[k + n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* rows p,q */
                    double apk = a[p + n * k], aqk = a[q + n * k];
                    a[p + n * k] = c * apk - s * aqk;
                    a[q + n * k] = s * apk + c * aqk;
                }
                a[p + n * q] = a[q + n * p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    double vp = Vec[k + (long)ldv * p], vq = Vec[k + (long)ldv * q];
                    Vec[k + (long)ldv * p] = c * vp - s * vq;
                    Vec[k + (long)ldv * q] = s * vp + c * vq;
                }
            }
    }
    for (int i = 0; i < n; ++i) W[i] = a[i + n * i];
    /* ascending order, eigenvectors follow */
    for (int i = 0; i < n - 1; ++i) {
        int m = i;
        for (int j = i + 1; j < n; ++j)
            if (W[j] < W[m]) m = j;
        if (m != i) {
            double t = W[i]; W[i] = W[m]; W[m] = t;
            for (int k = 0; k < n; ++k) {
                double x = Vec[k + (long)ldv * i];
                Vec[k + (long)ldv * i] = Vec[k + (long)ldv * m];
                Vec[k + (long)ldv * m] = x;
            }
        }
    }
    free(a);
    return 0;
}

int or_davidson(int ndim, int maxsub, int nroots, int maxmacro, int ptol,
                int corrflag, or_matvec_fn mv, void *ctx, const double *diag,
                double *B, double *evals, int *n_mtx, int *n_macro, double *resid)
{
    const long N = ndim;
    const double thresh = pow(10.0, -ptol);
    double *Z = calloc(N * maxsub, sizeof(double));
    double *S = calloc((long)maxsub * maxsub, sizeof(double));
    double *Vv = calloc((long)maxsub * maxsub, sizeof(double));
    double *sse = calloc(maxsub, sizeof(double));
    /* Ritz vectors, residual vectors and the Jacobi-Davidson vector are one
     * contiguous block, as in the reference's davwork (src/Davidson.cpp:181-187):
     * the reference's bookkeeping can index past the Ritz block into the
     * residual block, and this layout reproduces what it then computes. */
    double *ritz = calloc(N * (2 * nroots + 1), sizeof(double));
    double *res = ritz + N * nroots;
    double *jd = res + N * nroots;
    double *residuals = calloc(nroots, sizeof(double));
    double *cw = calloc(nroots, sizeof(double));
    int *ccode = calloc(nroots, sizeof(int)), *bidx = calloc(nroots, sizeof(int));
    int nmtx = 0, nconv = 0, jmacro = 1, nsub = 0, nnewB = nroots, nnewZ = 0, done = 0;

    /* GramSchmidt on the start vectors, src/Davidson.cpp:482-514 */
    for (int i = 0; i < nroots; ++i) {
        double *vi = B + N * i;
        scal(N, 1.0 / nrm2(N, vi), vi);
        for (int j = 0; j < i; ++j)
            for (int r = 0; r < 2; ++r) {
                axpy(N, -dot(N, vi, B + N * j), B + N * j, vi);
                scal(N, 1.0 / nrm2(N, vi), vi);
            }
    }
    while (!done) {
        if (nnewB > 0) { /* return 1 of the reverse interface: Z = H B for new columns */
            mv(B + N * nsub, Z + N * nsub, nnewB, ctx);
            nmtx += nnewB;
            nnewZ = nnewB;
            nnewB = 0;
        }
        /* ComputeS: S = Z^T B over the whole subspace (src/Davidson.cpp:385-410) */
        int m = nsub + nnewZ;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i <= j; ++i) S[i + (long)maxsub * j] = dot(N, Z + N * i, B + N * j);
        nsub = m;
        or_syev(nsub, S, maxsub, sse, Vv, maxsub);
        for (int r = 0; r < nroots; ++r) {
            if (nconv > 0) {
                double s = 0;
                for (int k = 0; k < nconv; ++k) s += Vv[k + (long)maxsub * r] * Vv[k + (long)maxsub * r];
                cw[r] = sqrt(s);
            } else cw[r] = 0;
        }
        /* Ritz vectors u = B v, renormalised (src/Davidson.cpp:240-246) */
        for (int r = 0; r < nroots; ++r) {
            double *u = ritz + N * r;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
            for (long g = 0; g < N; ++g) {
                double s = 0;
                for (int k = 0; k < nsub; ++k) s += B[g + N * k] * Vv[k + (long)maxsub * r];
                u[g] = s;
            }
            scal(N, 1.0 / nrm2(N, u), u);
        }
        /* residuals r = Z v - lambda u (src/Davidson.cpp:251-269) */
        for (int r = 0; r < nroots; ++r) {
            double *rc = res + N * r, lam = sse[r];
            const double *u = ritz + N * r;
#pragma omp parallel for num_threads(nthreads()) schedule(static)
            for (long g = 0; g < N; ++g) {
                double s = 0;
                for (int k = 0; k < nsub; ++k) s += Z[g + N * k] * Vv[k + (long)maxsub * r];
                rc[g] = s - lam * u[g];
            }
            residuals[r] = nrm2(N, rc);
        }
        /* convergence bookkeeping (src/Davidson.cpp:272-297) */
        int nnewconv = 0;
        for (int r = 0; r < nroots; ++r) {
            if (cw[r] > 0.95) ccode[r] = 0;
            else if (residuals[r] < thresh) {
                ccode[r] = 1;
                bidx[r] = nconv;
                memcpy(B + N * nconv, ritz + N * r, sizeof(double) * N);
                ++nconv;
                ++nnewconv;
            } else ccode[r] = 2;
        }
        if (nconv == nroots) {
            for (int r = 0; r < nroots; ++r) evals[r] = sse[r];
            memcpy(B, ritz, sizeof(double) * N * nroots);
            done = 1;
            break;
        }
        /* restart (src/Davidson.cpp:303-346) */
        if (nsub == maxsub || nnewconv > 0) {
            nnewB = 0;
            for (int r = 0; r < nroots; ++r)
                if (ccode[r] == 2) {
                    memcpy(B + N * (nnewB + nconv), ritz + N * r, sizeof(double) * N);
                    bidx[r] = nconv + nnewB;
                    ++nnewB;
                }
            for (int r = 0; r < nroots; ++r)
                if (ccode[r] > 0) {
                    double *dst = ritz + N * bidx[r];
#pragma omp parallel for num_threads(nthreads()) schedule(static)
                    for (long g = 0; g < N; ++g) {
                        double s = 0;
                        for (int k = 0; k < nsub; ++k) s += Z[g + N * k] * Vv[k + (long)maxsub * r];
                        dst[g] = s;
                    }
                }
            long off = (long)(nconv - nnewconv) * N;
            memcpy(Z + off, ritz + off, sizeof(double) * N * (nnewB + nnewconv));
            nsub = nroots;
            jmacro += 1;
            if (jmacro >= maxmacro) {
                /* maxmacro exhausted: return without touching evals, as the
                 * reference does (src/Davidson.cpp:332-341); B holds the
                 * current Ritz vectors */
                break;
            }
        }
        /* correction vectors (src/Davidson.cpp:353-371) */
        nnewB = 0;
        for (int r = 0; r < nroots; ++r) {
            if (nsub + nnewB == maxsub) break;
            if (ccode[r] < 2) continue;
            double *rc = res + N * r, lam = sse[r];
            if (corrflag == 1) {
                for (long g = 0; g < N; ++g) {
                    double dv = lam - diag[g];
                    if (fabs(dv) < 1e-12) dv = 1e-12;
                    rc[g] /= dv;
                }
            } else if (corrflag == 2) {
                const double *u = ritz + N * r;
                for (long g = 0; g < N; ++g) {
                    double mk = diag[g] - lam;
                    if (fabs(mk) < 1e-12) mk = 1e-12;
                    mk = 1.0 / mk;
                    rc[g] *= -mk;
                    jd[g] = mk * u[g];
                }
                double f = dot(N, u, rc) / dot(N, u, jd);
                axpy(N, f, jd, rc);
            }
            /* OrthoVecOnB (src/Davidson.cpp:447-474) */
            int nb = nsub + nnewB;
            if (nb > 0) {
                scal(N, 1.0 / nrm2(N, rc), rc);
                for (int i = 0; i < nb; ++i)
                    for (int t = 0; t < 2; ++t) {
                        axpy(N, -dot(N, rc, B + N * i), B + N * i, rc);
                        scal(N, 1.0 / nrm2(N, rc), rc);
                    }
            }
            memcpy(B + N * (nsub + nnewB), rc, sizeof(double) * N);
            ++nnewB;
        }
    }
    if (n_mtx) *n_mtx = nmtx;
    if (n_macro) *n_macro = jmacro;
    if (resid) memcpy(resid, residuals, sizeof(double) * nroots);
    free(Z); free(S); free(Vv); free(sse); free(ritz);
    free(residuals); free(cw); free(ccode); free(bidx);
    return nconv;
}

typedef struct {
    int nx, ny, nz;
    const double *Tx, *Ty, *Tz, *V;
} grid_ctx;

static void grid_mv(const double *x, double *y, int nvec, void *c)
{
    grid_ctx *g = (grid_ctx *)c;
    or_apply_h(g->nx, g->ny, g->nz, g->Tx, g->Ty, g->Tz, g->V, x, y, nvec);
}

int or_solve_grid(int nx, int ny, int nz, const double *Tx, const double *Ty,
                  const double *Tz, const double *V, int nroots, int maxsub,
                  int maxmacro, int ptol, int corrflag, double *B, double *evals,
                  int *n_mtx, int *n_macro, double *resid)
{
    long N = (long)nx * ny * nz;
    double *D = malloc(sizeof(double) * N);
    or_diag(nx, ny, nz, Tx, Ty, Tz, V, D);
    grid_ctx g = {nx, ny, nz, Tx, Ty, Tz, V};
    int rc = or_davidson((int)N, maxsub, nroots, maxmacro, ptol, corrflag, grid_mv, &g,
                         D, B, evals, n_mtx, n_macro, resid);
    free(D);
    return rc;
}

/* ======================================================================= f2 */
/* Potential::EvaluateDPPGTOPGradient (src/Gradients.cpp:79-176) at one point,
 * the terms weighted by w = psi^2 and summed as DVR::ComputeGradient does
 * (src/DVR.cpp:1099-1176): the point-charge term with the effective distance,
 * the Slater repulsive cores, and the electron's field derivative at the
 * polarizable sites dotted with the induced dipoles mu = InvA E
 * (DerivElecField, :179-225); MuCrossMu (:227-238) accumulates psi^2 mu and
 * psi^2 mu mu^T.  R = r - site throughout. */
static void grad_point(const or_pot *P, const double *r, double w, double psi, double *R,
                       double *Rx, double *Ry, double *Rz, double *Rm3, double *E,
                       double *mu, double *gel, double *tmu, double *mcm)
{
    for (int i = 0; i < P->ns; ++i) {
        Rx[i] = r[0] - P->sites[3 * i];
        Ry[i] = r[1] - P->sites[3 * i + 1];
        Rz[i] = r[2] - P->sites[3 * i + 2];
        double r2 = Rx[i] * Rx[i] + Ry[i] * Ry[i] + Rz[i] * Rz[i];
        R[i] = sqrt(r2);
        Rm3[i] = 1.0 / (r2 * R[i]);
    }
    for (int i = 0; i < P->nq; ++i) {
        int s = P->iq[i];
        double re = R[s], g;
        if (re < P->qdamp) {
            double t = re / P->qdamp;
            re = P->qdamp * (0.5 + t * t * t * (1.0 - 0.5 * t));
            g = P->q[i] * (3.0 * t * t - 2.0 * t * t * t) / (re * re);
        } else {
            g = P->q[i] / (R[s] * R[s]);
        }
        gel[3 * s] += w * (-g * Rx[s] / R[s]);
        gel[3 * s + 1] += w * (-g * Ry[s] / R[s]);
        gel[3 * s + 2] += w * (-g * Rz[s] / R[s]);
    }
    for (int i = 0; i < P->nrep; ++i) {
        int s = P->irep[i];
        double de = P->rep[2 * i + 1] * (-P->rep[2 * i]) * exp(-P->rep[2 * i] * R[s]);
        gel[3 * s] += w * (-de * Rx[s] / R[s]);
        gel[3 * s + 1] += w * (-de * Ry[s] / R[s]);
        gel[3 * s + 2] += w * (-de * Rz[s] / R[s]);
    }
    const int n = 3 * P->np;
    for (int i = 0; i < P->np; ++i) {
        int s = P->ip[i];
        double g = Rm3[s], re = R[s];
        if (re < P->pdamp) {
            double t = re / P->pdamp;
            re = P->pdamp * (0.5 + t * t * t * (1.0 - 0.5 * t));
            g = 1.0 / (re * re * re);
        }
        E[3 * i] = g * Rx[s] + P->epc[3 * i];
        E[3 * i + 1] = g * Ry[s] + P->epc[3 * i + 1];
        E[3 * i + 2] = g * Rz[s] + P->epc[3 * i + 2];
    }
    for (int a = 0; a < n; ++a) {
        double v = 0;
        for (int b = 0; b < n; ++b) v += P->inva[(size_t)n * b + a] * E[b];
        mu[a] = v;
    }
    for (int i = 0; i < P->np; ++i) {
        int s = P->ip[i];
        double g = Rm3[s], re = R[s], dg = -3.0 / (re * re * re * re);
        if (re < P->pdamp) {
            double t = re / P->pdamp;
            re = P->pdamp * (0.5 + t * t * t * (1.0 - 0.5 * t));
            dg = (3.0 * t * t - 0.5 * 4.0 * t * t * t) * (-3.0 / (re * re * re * re));
            g = 1.0 / (re * re * re);
        }
        const double Rv[3] = {Rx[s], Ry[s], Rz[s]};
        const double *m = mu + 3 * i;
        for (int a = 0; a < 3; ++a) {
            /* dE/dR_site along a, dotted with mu (src/Gradients.cpp:206-220) */
            double dE0 = -dg * Rv[a] / R[s] * Rv[0] - (a == 0 ? g : 0.0);
            double dE1 = -dg * Rv[a] / R[s] * Rv[1] - (a == 1 ? g : 0.0);
            double dE2 = -dg * Rv[a] / R[s] * Rv[2] - (a == 2 ? g : 0.0);
            gel[3 * s + a] += w * (-(dE0 * m[0] + dE1 * m[1] + dE2 * m[2]));
        }
    }
    for (int a = 0; a < n; ++a) {
        double t0 = psi * psi * mu[a];
        tmu[a] += t0;
        for (int b = 0; b < n; ++b) mcm[(size_t)n * a + b] += t0 * mu[b];
    }
}

int or_gradient_grid(const int *flags, const double *potpara, int nsites, const double *sites,
                     int nq, const double *q, const int *iq, int np, const double *alpha,
                     const int *ip, const double *epc, int nx, int ny, int nz,
                     const double *xg, const double *yg, const double *zg, const double *psi,
                     double *gel, double *tmu, double *mcm)
{
    or_pot P;
    int rc = pot_init(&P, flags, potpara, nsites, sites, nq, q, iq, 0, NULL, NULL, 0, NULL,
                      NULL, 1.0);
    if (!rc) rc = pol3_init(&P, np, alpha, ip, epc, 0.0);
    if (!rc && (flags[0] != 3 || P.damp_type != 2)) rc = -2;
    if (rc) { pot_free(&P); return rc; }
    const int n = 3 * np;
    memset(gel, 0, sizeof(double) * 3 * nsites);
    memset(tmu, 0, sizeof(double) * n);
    memset(mcm, 0, sizeof(double) * (size_t)n * n);
    double *R = malloc(sizeof(double) * 5 * nsites), *E = malloc(sizeof(double) * 2 * n);
    double *Rx = R + nsites, *Ry = R + 2 * nsites, *Rz = R + 3 * nsites, *Rm3 = R + 4 * nsites;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long g = i + (long)nx * (j + (long)ny * k);
                const double r[3] = {xg[i], yg[j], zg[k]};
                grad_point(&P, r, psi[g] * psi[g], psi[g], R, Rx, Ry, Rz, Rm3, E, E + n, gel,
                           tmu, mcm);
            }
    free(R); free(E);
    pot_free(&P);
    return 0;
}

void or_gradient_final(int nsites, int np, const double *gel, const double *tmu,
                       const double *mcm, const double *dTx, const double *dTy,
                       const double *dTz, const double *dEfield, const double *gww,
                       const double *polgrad, double *gfinal)
{
    /* Potential::FinalGradient (src/Gradients.cpp:240-268), the sum into the
     * water-water gradient and SubtractWWGradient (:270-284), in the order of
     * DVR::ComputeGradient (src/DVR.cpp:1160-1176) */
    const int n3 = 3 * np, n9 = 9 * np;
    double *loc = malloc(sizeof(double) * 3 * nsites);
    memcpy(loc, gel, sizeof(double) * 3 * nsites);
    for (int j = 0; j < nsites; ++j) {
        for (int d = 0; d < 3; ++d) {
            const double *de = dEfield + (size_t)n3 * (3 * j + d);
            double s = 0;
            for (int k = 0; k < n3; ++k) s += de[k] * tmu[k];
            loc[3 * j + d] += -0.5 * 2.0 * s;
        }
        int aj = 3 * (j / 4) + j % 4;
        if ((j + 1) % 4 != 0) {
            const double *T[3] = {dTx, dTy, dTz};
            for (int d = 0; d < 3; ++d) {
                double s = 0;
                for (int k = 0; k < n9; ++k)
                    s += T[d][(size_t)aj * n9 + k] * mcm[(size_t)aj * n9 + k];
                loc[3 * j + d] += -0.5 * 2.0 * s;
            }
        }
    }
    for (int c = 0; c < 3 * nsites; ++c) gfinal[c] = gww[c] + loc[c] - polgrad[c];
    free(loc);
}

/* ======================================================================= f3 */
/* Dual-grid coarse -> fine transfer, cubic grids (ClusterAnion's orchestration,
 * src/ClusterAnion.cpp:207-299).  Coarse grid: odd n -> (n + 1) / 2 points at
 * twice the spacing; even n -> (n / 2 + 1) / 3 * 2 points at three times it
 * (:209-232).  Values on the fine grid: a coarse point where the fine point
 * coincides with one, else a weighted sum of the 2, 4 or 8 coarse corners
 * around it: midpoint averages for double spacing (InterpolVdouble,
 * ExtendWfDouble, src/DVR.cpp:1380-1421, 1646-1722), inverse-distance weights
 * with the reference's six-digit constants for triple spacing (InterpolVtriple
 * for potentials, ExtendWfThree for wavefunctions, src/DVR.cpp:1240-1378,
 * 1446-1644), which differ in three weights (below). */
void or_dual_coarse(int n, double L, int *nc, double *Lc)
{
    double h = L / (n + 1);
    if (n % 2 == 1) {
        *nc = (n + 1) / 2;
        *Lc = *nc * 2 * h + 2 * h;
    } else {
        *nc = (n / 2 + 1) / 3 * 2;
        *Lc = *nc * 3 * h + 3 * h;
    }
}

/* weights of the corners (a, b, c) in {0,1}^3, index a + 2 b + 4 c, for a fine
 * point with residues (rx, ry, rz) in {0,1,2}^3 of triple spacing; kind 0 the
 * wavefunction (ExtendWfThree), 1 the potential (InterpolVtriple) */
static void tri_weights(int kind, int rx, int ry, int rz, double *w)
{
    const double w1 = 0.707107 / 1.955087, w2 = 0.447214 / 1.955087, w3 = 0.353553 / 1.955087;
    const double g1 = 0.57735 / 3.09077, g2 = 0.408248 / 3.09077, g3 = 0.33333 / 3.09077,
                 g4 = 0.288675 / 3.09077;
    for (int k = 0; k < 8; ++k) w[k] = 0.0;
    int r[3] = {rx, ry, rz}, nz = (rx != 0) + (ry != 0) + (rz != 0);
    if (nz == 0) { w[0] = 1.0; return; }
    if (nz == 1) {
        int ax = rx ? 0 : ry ? 1 : 2, s = 1 << ax;
        w[0] = r[ax] == 1 ? 2.0 / 3.0 : 1.0 / 3.0;
        w[s] = r[ax] == 1 ? 1.0 / 3.0 : 2.0 / 3.0;
        return;
    }
    if (nz == 2) {
        /* the two moving axes u < v; corners (0,0), (u), (v), (u,v) */
        int u = rx ? 0 : 1, v = rz ? 2 : 1;
        int su = 1 << u, sv = 1 << v, ru = r[u], rv = r[v];
        double c00, cu, cv, cuv;
        if (ru == 1 && rv == 1) { c00 = w1; cu = w2; cv = w2; cuv = w3; }
        else if (ru == 1 && rv == 2) { c00 = w2; cu = w3; cv = w1; cuv = w2; }
        else if (ru == 2 && rv == 1) {
            c00 = w2; cu = w1; cv = w3; cuv = w2;
            /* ExtendWfThree, plane (x, z): w1 on (0, c+1) (src/DVR.cpp:1570) */
            if (kind == 0 && u == 0 && v == 2) cv = w1;
        } else { c00 = w3; cu = w2; cv = w2; cuv = w1; }
        w[0] = c00; w[su] = cu; w[sv] = cv; w[su + sv] = cuv;
        return;
    }
    /* nz == 3: corners 0, x, y, z, xy, yz, xz, xyz */
    static const double *tab = 0;
    (void)tab;
    double t[8];
    int key = 100 * rx + 10 * ry + rz;
    switch (key) {
    case 111: { double v[8] = {g1, g2, g2, g2, g3, g3, g3, g4}; memcpy(t, v, sizeof(t)); } break;
    case 211: { double v[8] = {g2, g1, g3, g3, g2, g4, g2, g3}; memcpy(t, v, sizeof(t)); } break;
    case 121: { double v[8] = {g2, g3, g1, g3, g2, g2, g4, g3}; memcpy(t, v, sizeof(t)); } break;
    case 112: { double v[8] = {g2, g3, g3, g1, g4, g2, g2, g3}; memcpy(t, v, sizeof(t)); } break;
    case 221: { double v[8] = {g3, g2, g2, g4, g1, g3, g3, g2}; memcpy(t, v, sizeof(t)); } break;
    case 212:
        if (kind == 0) { /* ExtendWfThree (src/DVR.cpp:1615-1620): x and xy swapped */
            double v[8] = {g3, g3, g4, g2, g2, g3, g1, g2}; memcpy(t, v, sizeof(t));
        } else {
            double v[8] = {g3, g2, g4, g2, g3, g3, g1, g2}; memcpy(t, v, sizeof(t));
        }
        break;
    case 122: { double v[8] = {g3, g4, g2, g2, g3, g1, g3, g2}; memcpy(t, v, sizeof(t)); } break;
    default: { double v[8] = {g4, g3, g3, g3, g2, g2, g2, g1}; memcpy(t, v, sizeof(t)); } break;
    }
    /* t in the reference's order 0, x, y, z, xy, yz, xz, xyz -> corner index */
    static const int idx[8] = {0, 1, 2, 4, 3, 6, 5, 7};
    for (int k = 0; k < 8; ++k) w[idx[k]] = t[k];
}

void or_dual_weights(int kind, int rx, int ry, int rz, double *w) { tri_weights(kind, rx, ry, rz, w); }

/* fine[n^3] from coarse[nc^3] (both x fastest); kind 0 wavefunction, 1 potential */
int or_dual_extend(int n, int nc, int kind, const double *coarse, double *fine)
{
    if (n % 2 == 1) {
        if (nc != (n + 1) / 2) return -1;
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    int p[3] = {i / 2, j / 2, k / 2}, odd[3] = {i & 1, j & 1, k & 1};
                    int m = odd[0] + odd[1] + odd[2];
                    double s = 0.0;
                    for (int c = 0; c < 8; ++c) {
                        int a = c & 1, b = (c >> 1) & 1, e = (c >> 2) & 1;
                        if ((a && !odd[0]) || (b && !odd[1]) || (e && !odd[2])) continue;
                        s += coarse[(p[0] + a) + (long)nc * ((p[1] + b) + (long)nc * (p[2] + e))];
                    }
                    fine[i + (long)n * (j + (long)n * k)] = s / (double)(1 << m);
                }
        return 0;
    }
    if (nc != (n / 2 + 1) / 3 * 2) return -1;
    int check = n - (nc * 3 / 2 - 1) * 2, sh = check / 2;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int q[3] = {i - sh, j - sh, k - sh};
                double v = 0.0;
                int out = 0;
                for (int d = 0; d < 3; ++d)
                    if (q[d] < 0 || q[d] >= n - check) out = 1;
                if (!out) {
                    double w[8];
                    tri_weights(kind, q[0] % 3, q[1] % 3, q[2] % 3, w);
                    int p[3] = {q[0] / 3, q[1] / 3, q[2] / 3};
                    for (int c = 0; c < 8; ++c) {
                        if (w[c] == 0.0) continue;
                        int a = c & 1, b = (c >> 1) & 1, e = (c >> 2) & 1;
                        v += coarse[(p[0] + a) + (long)nc * ((p[1] + b) + (long)nc * (p[2] + e))] * w[c];
                    }
                }
                fine[i + (long)n * (j + (long)n * k)] = v;
            }
    return 0;
}
