// host.cpp -- host-side FP64 pieces of the PISCES MI355X library: sine-DVR
// kinetic matrices, the rigid-water projection, the potential parameter tables
// and the small symmetric eigensolver of the projected Davidson matrix.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "blockops.hpp"
#include "common.hpp"

namespace pisces {

namespace {
// 15-digit pi of the reference (src/sine_dvr.cpp:24, src/constants.h:4); using
// it keeps the kinetic matrices and Euler rotations identical to the reference.
constexpr double kPi = 3.14159265358979;
constexpr double kBohr2Angs = 0.52917720859;  // src/constants.h:17
thread_local std::string g_last_error;
}  // namespace

void set_last_error(const std::string &s) { g_last_error = s; }
const char *last_error() { return g_last_error.c_str(); }

// sine-DVR grid and full kinetic matrix: src/sine_dvr.cpp:45-131, eq. (B.62),
// (B.65) of Phys. Rep. 324, 1 (2000).
void sine_dvr_full(int n, double mass, double x0, double xnp1, double *x, double *T)
{
    const double dx = (xnp1 - x0) / (double)(n + 1);
    for (int j = 0; j < n; ++j) x[j] = x0 + dx * (double)(j + 1);
    const double f1 = 0.5 * kPi * kPi / (mass * dx * dx);
    const double f2 = kPi / (double)(n + 1);
    const double f3 = (double)(n + 1) * (double)(n + 1);
    for (int a = 0; a < n; ++a) {
        const double s = std::sin((a + 1) * f2);
        T[a + n * a] = f1 * (1.0 / 3.0 + 1.0 / (6.0 * f3) - 0.5 / (f3 * (s * s)));
    }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < a; ++b) {
            const double dc = std::cos(f2 * (a + 1)) - std::cos(f2 * (b + 1));
            double t = f1 * 2.0 / f3 * std::sin(f2 * (a + 1)) * std::sin(f2 * (b + 1)) / (dc * dc);
            if ((a - b) % 2 == 1) t = -t;
            T[a + n * b] = t;
            T[b + n * a] = t;
        }
}

// DVRType 3 (src/KE_diag.cpp Tdiag, applied by VectorFFT::apply through FFTW,
// src/VectorFFTW.cpp): grid points as for the sine DVR, x_j = x0 + (j + 1) dx,
// dx = L / (n + 1); the kinetic energy is diagonal in the discrete Fourier basis
// of the n points, e_a = k_a^2 / (2 mass) with k_a = 2 pi m_a / (n dx),
// m_a = a for a < n/2 and a - n above.  In the grid basis that is the real
// symmetric circulant T_jl = (1/n) sum_a e_a cos(2 pi m_a (j - l) / n), so the
// FFT kinetic of a 3-D grid is the same axis sum T_x + T_y + T_z as the sine DVR
// with these matrices (the reference's 3-D KE_diag is e_x + e_y + e_z).
void fft_kinetic_full(int n, double mass, double x0, double xnp1, double *x, double *T)
{
    const double dx = (xnp1 - x0) / (double)(n + 1);
    for (int j = 0; j < n; ++j) x[j] = x0 + dx * (double)(j + 1);
    const double dk = 2.0 * kPi / ((double)n * dx);
    std::vector<double> e(n);
    std::vector<int> m(n);
    for (int a = 0; a < n; ++a) {
        m[a] = a >= n / 2 ? a - n : a;
        e[a] = (dk * m[a]) * (dk * m[a]) / (2.0 * mass);
    }
    for (int d = 0; d < n; ++d) {
        // entry for j - l = d (mod n); the phase index m d is reduced mod n first
        double t = 0.0;
        for (int a = 0; a < n; ++a) {
            const long long r = (((long long)m[a] * d) % n + n) % n;
            t += e[a] * std::cos(2.0 * kPi * (double)r / (double)n);
        }
        t /= (double)n;
        for (int j = 0; j < n; ++j) T[(size_t)j * n + (j + n - d) % n] = t;  // T[j][j - d]
    }
}

// DVRType 1 (ho_dvr, src/ho_dvr.cpp): the harmonic-oscillator DVR of frequency w
// (mass 1).  The grid points are the eigenvalues of the position operator in
// the oscillator basis (tridiagonal, off-diagonal sqrt(i + 1)) scaled by
// 1/sqrt(2 w); with Z its eigenvectors (each signed so that its first component
// is positive), T = Z^T T_FBR Z, T_FBR the pentadiagonal kinetic matrix of the
// basis: diagonal w (i + 1/2) / 2, (i, i + 2) entries -w sqrt((i + 1)(i + 2)) / 4.
void ho_dvr_full(int n, double w, double *x, double *T)
{
    std::vector<double> X((size_t)n * n, 0.0), Z((size_t)n * n), ev(n);
    for (int i = 0; i + 1 < n; ++i)
        X[(size_t)i * n + i + 1] = X[(size_t)(i + 1) * n + i] = std::sqrt((double)(i + 1));
    syev_host(n, X.data(), n, ev.data(), Z.data(), n);  // Z(r, c) = Z[r + n c]
    const double s = 1.0 / std::sqrt(2.0 * w);
    for (int i = 0; i < n; ++i) x[i] = ev[i] * s;
    for (int c = 0; c < n; ++c)
        if (Z[(size_t)n * c] < 0)
            for (int r = 0; r < n; ++r) Z[r + (size_t)n * c] = -Z[r + (size_t)n * c];
    std::vector<double> F((size_t)n * n, 0.0), FZ((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        F[(size_t)i * n + i] = 0.5 * w * (i + 0.5);
        if (i + 2 < n)
            F[(size_t)i * n + i + 2] = F[(size_t)(i + 2) * n + i] =
                -0.25 * w * std::sqrt((double)(i + 1) * (double)(i + 2));
    }
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < n; ++c) {
            double a = 0.0;
            for (int k = 0; k < n; ++k) a += F[(size_t)i * n + k] * Z[k + (size_t)n * c];
            FZ[(size_t)i * n + c] = a;
        }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b <= a; ++b) {
            double t = 0.0;
            for (int k = 0; k < n; ++k) t += Z[k + (size_t)n * a] * FZ[(size_t)k * n + b];
            T[(size_t)a * n + b] = T[(size_t)b * n + a] = t;
        }
}

// DVRType 20 (cm_ham / cm_dvr, src/cm_dvr.cpp): the Colbert-Miller kinetic
// matrix of n points over the length ru - rl, with grid points rl + i (ru - rl)/n
// as the reference places them.
void cm_dvr_full(int n, double mass, double rl, double ru, double *x, double *T)
{
    const double len = ru - rl, dr = len / (double)n;
    for (int i = 0; i < n; ++i) x[i] = rl + dr * (double)i;
    const int m = n + 1;
    const double f1 = kPi * kPi / (4.0 * mass * len * len);
    const double f2 = kPi / (double)(2 * m);
    const double f3 = (2.0 * m * m + 1.0) / 3.0;
    for (int i = 0; i < n; ++i) {
        const double si = std::sin(kPi * (double)(i + 1) / (double)m);
        T[(size_t)i * n + i] = f1 * (f3 - 1.0 / (si * si));
        for (int j = 0; j < i; ++j) {
            const double a = std::sin(f2 * (j - i)), b = std::sin(f2 * (j + i + 2));
            double t = f1 * (1.0 / (a * a) - 1.0 / (b * b));
            if ((i - j) % 2 == 1) t = -t;
            T[(size_t)i * n + j] = T[(size_t)j * n + i] = t;
        }
    }
}

// ------------------------------------------------------------------ water
// Standard orientation of the DPP monomer relative to its centre of mass, bohr:
// Water::SetupCharge case 2 (src/Water.cpp:216-236) with src/ttm_data.h:29-44.
static void monomer_ref(double ref[4][3])
{
    const double a2b = 1.0 / kBohr2Angs;
    const double oz = 0.0650980307, hz = -0.5207842453, hy = 0.756950327, dm = 0.25;
    const double r[4][3] = {{0, 0, oz * a2b}, {0, hy * a2b, hz * a2b}, {0, -hy * a2b, hz * a2b},
                            {0, 0, (oz - dm) * a2b}};
    std::memcpy(ref, r, sizeof(r));
}

// Water::Reposition (src/Water.cpp:695-825): molecular frame from the O->H
// vectors, Euler angles, centre of mass with masses 16/1/1.
static void monomer_project(const double *w, double conf[6])
{
    const double o[3] = {w[0], w[1], w[2]};
    double r1[3], r3[3], z[3], y[3], x[3];
    for (int a = 0; a < 3; ++a) { r1[a] = w[3 + a] - o[a]; r3[a] = w[6 + a] - o[a]; }
    for (int a = 0; a < 3; ++a) z[a] = r1[a] + r3[a];
    const double zm = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (int a = 0; a < 3; ++a) { z[a] /= zm; z[a] = -z[a]; }
    const double p = r1[0] * z[0] + r1[1] * z[1] + r1[2] * z[2];
    for (int a = 0; a < 3; ++a) y[a] = r1[a] - p * z[a];
    const double ym = std::sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
    for (int a = 0; a < 3; ++a) y[a] /= ym;
    x[0] = y[1] * z[2] - y[2] * z[1];
    x[1] = y[2] * z[0] - y[0] * z[2];
    x[2] = y[0] * z[1] - y[1] * z[0];
    double phi, theta, psi;
    theta = std::acos(z[2]);
    const double st = std::sin(theta);
    if (std::fabs(st) > 1e-5) {
        double tc = -x[2] / st;
        double ac = std::acos(tc);
        if (tc <= -1) ac = kPi;
        psi = (y[2] >= 0) ? ac : 2 * kPi - ac;
        tc = z[0] / st;
        if (tc > 1) tc = 1;
        if (tc < -1) tc = -1;
        phi = (z[1] >= 0) ? std::acos(tc) : 2 * kPi - std::acos(tc);
    } else {
        phi = 0.0;
        psi = (x[1] >= 0) ? std::acos(y[1]) : 2 * kPi - std::acos(y[1]);
    }
    conf[0] = (w[3] + w[6] + 16.0 * o[0]) / 18.0;
    conf[1] = (w[4] + w[7] + 16.0 * o[1]) / 18.0;
    conf[2] = (w[5] + w[8] + 16.0 * o[2]) / 18.0;
    conf[3] = phi;
    conf[4] = theta;
    conf[5] = psi;
}

// Water::UpdatePositions = ResetPositions + ApplyEuler (src/Water.cpp:837-943)
static void monomer_place(const double conf[6], double out[4][3])
{
    double ref[4][3];
    monomer_ref(ref);
    const double c1 = std::cos(conf[3]), c2 = std::cos(conf[4]), c3 = std::cos(conf[5]);
    const double s1 = std::sin(conf[3]), s2 = std::sin(conf[4]), s3 = std::sin(conf[5]);
    const double R[9] = {c1 * c2 * c3 - s1 * s3, -c1 * c2 * s3 - s1 * c3, c1 * s2,
                         s1 * c2 * c3 + c1 * s3, -s1 * c2 * s3 + c1 * c3, s1 * s2,
                         -s2 * c3,               s2 * s3,                 c2};
    for (int a = 0; a < 4; ++a) {
        const double p0 = conf[0] + ref[a][0], p1 = conf[1] + ref[a][1], p2 = conf[2] + ref[a][2];
        const double o0 = p0 - conf[0], o1 = p1 - conf[1], o2 = p2 - conf[2];
        out[a][0] = R[0] * o0 + R[1] * o1 + R[2] * o2 + conf[0];
        out[a][1] = R[3] * o0 + R[4] * o1 + R[5] * o2 + conf[1];
        out[a][2] = R[6] * o0 + R[7] * o1 + R[8] * o2 + conf[2];
    }
}

bool water_sites_host(int nw, const double *xyz_ang, double *sites)
{
    const double a2b = 1.0 / kBohr2Angs;
    for (int w = 0; w < nw; ++w) {
        double b[9], c[6], pos[4][3];
        for (int a = 0; a < 9; ++a) b[a] = xyz_ang[9 * w + a] * a2b;
        monomer_project(b, c);
        monomer_place(c, pos);
        for (int s = 0; s < 4; ++s)
            for (int a = 0; a < 3; ++a) {
                if (!std::isfinite(pos[s][a])) return false;
                sites[12 * w + 3 * s + a] = pos[s][a];
            }
    }
    return true;
}

// GetLists / ReportInducedDipoles ordering (src/DPP.cpp:402-529): charges on
// H1, H2, M of each water; induced dipoles on O, H1, H2.
// Energy of the neutral water cluster, WaterCluster::CalcIntermolecularPotential
// (src/DPP.cpp:1648-1700): over every pair of waters the charge-charge
// electrostatics (WaterWaterElectrostaticEnergy, :1820-1870), the induced
// dipoles of O,H1,H2 in the other water's damped charge field, halved
// (WaterWaterPolarizationEnergy, :1935-2024, damping a = 0.21), and the DPP
// exponential repulsion with the Tang-Toennies damped O-O C6 dispersion
// (WaterWaterDispersionEnergy, :2256-2303; src/ttm_data.h:76-119,
// src/Water.cpp:225-235).  out: E0, electrostatic, polarization, dispersion.
void water_energy_host(int nw, const double *sites, const double *mu, double out[4])
{
    const double a2b = 1.0 / kBohr2Angs, a3 = a2b * a2b * a2b;
    const double kcal = 6.02214179e23 * 4.35974394e-18 / 4.184 / 1000;  // hartree in kcal/mol
    const double qh = 0.5742185854046506;
    // per site O, H1, H2, M: charge, damping polarizability, polarizability
    const double q[4] = {0.0, qh, qh, -2.0 * qh};
    const double sg[4] = {0.0, 0.28 * a3, 0.28 * a3, 1.22 * a3};
    const double al[3] = {1.22 * a3, 0.28 * a3, 0.28 * a3};
    const double delta = 2.23 / a2b, c6 = -1300.0 / kcal * std::pow(a2b, 6);
    // repulsion (A, b) of the O-O, H-H and O-H pairs
    const double rep[3][2] = {{115271.77066806905 / kcal, 4.167868773196562 / a2b},
                              {119085.35806729361 / kcal, 5.581698710292979 / a2b},
                              {1803.662375255084 / kcal, 3.5354602029164868 / a2b}};
    auto dist = [](const double *a, const double *b, double d[3]) {
        d[0] = a[0] - b[0];
        d[1] = a[1] - b[1];
        d[2] = a[2] - b[2];
        return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    };
    auto damp_cd = [](double r, double ai, double aj) {
        const double A = std::pow(ai * aj, 1.0 / 6.0);
        return 1.0 - std::exp(-0.21 * std::pow(r / A, 3));
    };
    double e_es = 0, e_pol = 0, e_disp = 0, nre = 0, lj = 0;
    for (int i = 0; i + 1 < nw; ++i)
        for (int j = i + 1; j < nw; ++j) {
            const double *s1 = sites + 12 * i, *s2 = sites + 12 * j;
            double d[3], es = 0, pol = 0, disp = 0;
            for (int a = 1; a < 4; ++a)
                for (int b = 1; b < 4; ++b) es += q[a] * q[b] / std::sqrt(dist(s1 + 3 * a, s2 + 3 * b, d));
            for (int k = 0; k < 3; ++k) {
                const double *m1 = mu + 9 * i + 3 * k, *m2 = mu + 9 * j + 3 * k;
                for (int b = 1; b < 4; ++b) {  // charges of water j at dipole k of water i
                    const double r2 = dist(s1 + 3 * k, s2 + 3 * b, d), r = std::sqrt(r2);
                    pol -= q[b] * (m1[0] * d[0] + m1[1] * d[1] + m1[2] * d[2]) *
                           damp_cd(r, al[k], sg[b]) * std::pow(r2, -1.5) / 2.0;
                }
                for (int b = 1; b < 4; ++b) {  // charges of water i at dipole k of water j
                    const double r2 = dist(s2 + 3 * k, s1 + 3 * b, d), r = std::sqrt(r2);
                    pol -= q[b] * (m2[0] * d[0] + m2[1] * d[1] + m2[2] * d[2]) *
                           damp_cd(r, al[k], sg[b]) * std::pow(r2, -1.5) / 2.0;
                }
            }
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double r2 = dist(s1 + 3 * a, s2 + 3 * b, d), r = std::sqrt(r2);
                    if (a == 0 && b == 0) {
                        const double x = delta * r;
                        const double tt = 1.0 + x + std::pow(x, 2) / 2.0 + std::pow(x, 3) / 6.0 +
                                          std::pow(x, 4) / 24.0 + std::pow(x, 5) / 120.0 +
                                          std::pow(x, 6) / 720.0;
                        disp += rep[0][0] * std::exp(-rep[0][1] * r) +
                                (1 - tt * std::exp(-x)) * c6 / (r2 * r2 * r2);
                    } else {
                        const int t = (a != 0 && b != 0) ? 1 : 2;
                        disp += rep[t][0] * std::exp(-rep[t][1] * r);
                    }
                }
            e_es += es;
            e_pol += pol;
            e_disp += disp;
            nre += es + pol;
            lj += disp;
        }
    out[0] = nre + lj;
    out[1] = e_es;
    out[2] = e_pol;
    out[3] = e_disp;
}

void water_lists_host(int nw, std::vector<double> &q, std::vector<int> &iq, std::vector<int> &id)
{
    const double qh = 0.5742185854046506;  // src/ttm_data.h:14
    q.clear(); iq.clear(); id.clear();
    for (int w = 0; w < nw; ++w) {
        q.push_back(qh); iq.push_back(4 * w + 1);
        q.push_back(qh); iq.push_back(4 * w + 2);
        q.push_back(-2.0 * qh); iq.push_back(4 * w + 3);
        for (int s = 0; s < 3; ++s) id.push_back(4 * w + s);
    }
}

// ------------------------------------------------------------------ PolType 3
// ComputeInvA (src/MolPolAux.cpp:39-136): A = alpha^-1 on the diagonal 3x3
// blocks and -T_ij off the diagonal, T_ij the Thole-damped dipole tensor of
// CalcDDTensor (T = 3 f5 R R^T / r^5 - f3 I / r^3, x = a r^3 / sqrt(alpha_i
// alpha_j), f3 = 1 - e^-x, f5 = 1 - (1 + x) e^-x, a = 0.3) for every pair; then
// its inverse (the reference: dsytrf/dsytri; here by Cholesky where A is
// positive definite, else Gauss-Jordan elimination with partial pivoting).
// sum_{k < n} a[k] b[k] in eight interleaved partial sums (independent chains)
static double dot8(const double *a, const double *b, int n)
{
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int k = 0;
    for (; k + 8 <= n; k += 8)
        for (int u = 0; u < 8; ++u) s[u] += a[k + u] * b[k + u];
    for (; k < n; ++k) s[0] += a[k] * b[k];
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

bool pol3_inva_host(int np, const double *rpol, const double *alpha, int n3p, double *inva)
{
    const int n = 3 * np;
    if (np <= 0 || n3p < n) return false;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < np; ++i) {
        if (!(alpha[i] > 0.0)) return false;
        for (int c = 0; c < 3; ++c) A[(size_t)n * (3 * i + c) + 3 * i + c] = 1.0 / alpha[i];
    }
    for (int i = 0; i < np; ++i)
        for (int j = i + 1; j < np; ++j) {
            const double r[3] = {rpol[3 * i] - rpol[3 * j], rpol[3 * i + 1] - rpol[3 * j + 1],
                                 rpol[3 * i + 2] - rpol[3 * j + 2]};
            const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const double rr = std::sqrt(r2), r3 = rr * r2, r5 = r3 * r2;
            const double x = 0.3 * r3 / std::sqrt(alpha[i] * alpha[j]), ex = std::exp(-x);
            const double f3 = 1.0 - ex, f5 = 1.0 - (1.0 + x) * ex;
            for (int u = 0; u < 3; ++u)
                for (int w = 0; w < 3; ++w) {
                    const double t = 3.0 * f5 * r[u] * r[w] / r5 - (u == w ? f3 / r3 : 0.0);
                    A[(size_t)n * (3 * i + u) + 3 * j + w] = -t;
                    A[(size_t)n * (3 * j + w) + 3 * i + u] = -t;
                }
        }
    return sym_inverse(n, A, n3p, inva, 0);
}

// Inverse of the symmetric n x n matrix A (row-major, overwritten) into the
// leading n x n block of inva (pitch n3p; the rest zeroed).  method 0: Cholesky
// where A is positive definite, else Gauss-Jordan; 1: Cholesky only (false if
// A is not positive definite); 2: Gauss-Jordan only.  False on a singular or
// non-finite result.
bool sym_inverse(int n, std::vector<double> &A, int n3p, double *inva, int method)
{
    // symmetric positive definite (any physical set of sites): Cholesky A = L L^T,
    // M = L^-1, InvA = M^T M -- about n^3 / 2 multiply-adds, every inner loop a
    // dot product of contiguous rows
    if (method != 2) {
        std::vector<double> L((size_t)n * n, 0.0), Mt((size_t)n * n, 0.0);
        bool spd = true;
        for (int j = 0; j < n && spd; ++j) {
            const double *lj = &L[(size_t)n * j];
            const double d = A[(size_t)n * j + j] - dot8(lj, lj, j);
            if (!(d > 0.0)) {
                spd = false;
                break;
            }
            const double ljj = std::sqrt(d), inv = 1.0 / ljj;
            L[(size_t)n * j + j] = ljj;
            for (int i = j + 1; i < n; ++i) {
                const double *li = &L[(size_t)n * i];
                L[(size_t)n * i + j] = (A[(size_t)n * i + j] - dot8(li, lj, j)) * inv;
            }
        }
        if (spd) {
            // Mt = (L^-1)^T, upper triangular: column j of L^-1 solves L x = e_j
            std::vector<double> x(n);
            for (int j = 0; j < n; ++j) {
                for (int i = 0; i < j; ++i) x[i] = 0.0;
                for (int i = j; i < n; ++i) {
                    const double *li = &L[(size_t)n * i];
                    x[i] = ((i == j ? 1.0 : 0.0) - dot8(li + j, &x[j], i - j)) / li[i];
                }
                for (int i = j; i < n; ++i) Mt[(size_t)n * j + i] = x[i];
            }
            // InvA[a][b] = sum_{k >= max(a, b)} M[k][a] M[k][b] = Mt[a][.] . Mt[b][.]
            std::fill(inva, inva + (size_t)n3p * n3p, 0.0);
            for (int a = 0; a < n; ++a) {
                const double *ra = &Mt[(size_t)n * a];
                for (int b = 0; b <= a; ++b) {
                    const double *rb = &Mt[(size_t)n * b];
                    const double t = dot8(ra + a, rb + a, n - a);
                    inva[(size_t)n3p * a + b] = inva[(size_t)n3p * b + a] = t;
                }
            }
            for (size_t k = 0; k < (size_t)n3p * n3p; ++k)
                if (!std::isfinite(inva[k])) return false;
            return true;
        }
        if (method == 1) return false;
    }
    // otherwise Gauss-Jordan elimination with partial pivoting on [A | I]
    std::vector<double> X((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) X[(size_t)n * k + k] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[(size_t)n * r + c]) > std::fabs(A[(size_t)n * piv + c])) piv = r;
        if (A[(size_t)n * piv + c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(A[(size_t)n * c + k], A[(size_t)n * piv + k]);
                std::swap(X[(size_t)n * c + k], X[(size_t)n * piv + k]);
            }
        const double f = 1.0 / A[(size_t)n * c + c];
        double *ac = &A[(size_t)n * c], *xc = &X[(size_t)n * c];
        for (int k = 0; k < n; ++k) { ac[k] *= f; xc[k] *= f; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double m = A[(size_t)n * r + c];
            if (m == 0.0) continue;
            double *ar = &A[(size_t)n * r], *xr = &X[(size_t)n * r];
            for (int k = c; k < n; ++k) ar[k] -= m * ac[k];
            for (int k = 0; k < n; ++k) xr[k] -= m * xc[k];
        }
    }
    // dsymv("L") in SelfConsistentPolarizability reads the lower triangle only
    std::fill(inva, inva + (size_t)n3p * n3p, 0.0);
    for (int a = 0; a < n; ++a)
        for (int b = 0; b <= a; ++b)
            inva[(size_t)n3p * a + b] = inva[(size_t)n3p * b + a] = X[(size_t)n * a + b];
    for (size_t k = 0; k < (size_t)n3p * n3p; ++k)
        if (!std::isfinite(inva[k])) return false;
    return true;
}

// ------------------------------------------------------------------ potential
// Potential::SetupDPPGTOP (src/Potential.cpp:234-593) with the repulsive cores
// Set6GTORepCore / Set12GTORepCore / Set4STORepCore (:598-728).  The numeric
// tables are the reference's published model parameters.
bool potential_setup(const int *flags, const double *pp, int nsites, PotParams &P,
                     std::vector<double> &rep, std::vector<int> &irep, std::vector<int> &ipps)
{
    const int pf = flags[0];
    if (pf < 1 || pf > 4) return false;
    if (nsites % 4) return false;
    P.damp_type = flags[1];
    P.pol_type = flags[3];
    P.adiabatic = flags[4];
    if (P.pol_type != 0 && P.pol_type != 1 && P.pol_type != 3) return false;
    P.alpha = pp[0];
    const double vrep = pp[2], vrepO = (flags[2] == 1) ? pp[5] : pp[2];
    if (P.damp_type == 1) {
        P.qdamp = 1.0 / (pp[3] * pp[3]);
        P.ddamp = 1.0 / (pp[4] * pp[4]);
        P.pdamp = 1.0 / (pp[1] * pp[1]);
    } else if (P.damp_type == 2) {
        P.qdamp = std::fabs(pp[3]);
        P.ddamp = std::fabs(pp[4]);
        P.pdamp = std::fabs(pp[1]);
    } else {
        return false;
    }
    P.eps = pp[6];
    const int nw = nsites / 4;
    ipps.clear();
    if (P.pol_type == 1)
        for (int w = 0; w < nw; ++w) ipps.push_back(4 * w + 3);  // M sites
    rep.clear(); irep.clear();
    P.rep_type = (pf == 3) ? 2 : (pf == 4) ? 0 : 1;
    const double sH = 1.0 / vrep, sO = 1.0 / vrepO;
    for (int w = 0; w < nw; ++w) {
        if (pf == 1) {
            const double e6[3] = {33.87, 5.095, 1.159};
            const double c6[3] = {2.84230876640242, 4.55401933663180, 6.91243747981459};
            for (int h = 0; h < 2; ++h)
                for (int k = 0; k < 3; ++k) {
                    irep.push_back(4 * w + 1 + h);
                    rep.push_back(e6[k]);
                    rep.push_back(c6[k] * (1.0 / vrep));
                }
        } else if (pf == 2) {
            const double eo[6] = {322.037, 48.4308, 10.4206, 7.40294, 1.5762, 0.373684};
            const double co[6] = {41.69216064, 58.35946922, 35.19878892, -3.27944113, 4.44843633, 2.36213092};
            const double eh[3] = {5.447178, 0.824547, 0.183192};
            const double ch[3] = {0.69233452, 0.66686241, 0.07193908};
            for (int k = 0; k < 6; ++k) { irep.push_back(4 * w); rep.push_back(eo[k]); rep.push_back(co[k] * sO); }
            for (int h = 0; h < 2; ++h)
                for (int k = 0; k < 3; ++k) {
                    irep.push_back(4 * w + 1 + h);
                    rep.push_back(eh[k]);
                    rep.push_back(ch[k] * sH);
                }
        } else if (pf == 3) {
            irep.push_back(4 * w); rep.push_back(9.98526); rep.push_back(192.608 * sO);
            irep.push_back(4 * w); rep.push_back(1.67075); rep.push_back(13.8937 * sO);
            irep.push_back(4 * w + 1); rep.push_back(1.64121); rep.push_back(1.7643 * sH);
            irep.push_back(4 * w + 2); rep.push_back(1.64121); rep.push_back(1.7643 * sH);
        }
    }
    return true;
}

// ------------------------------------------------------------------ eigensolver
void syev_host(int n, const double *A, int lda, double *W, double *V, int ldv)
{
    std::vector<double> a((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            const double x = A[i + (size_t)lda * j];
            a[i + (size_t)n * j] = x;
            a[j + (size_t)n * i] = x;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) V[i + (size_t)ldv * j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double v = a[i + (size_t)n * j];
                tot += v * v;
                if (i != j) off += v * v;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p + (size_t)n * q];
                if (apq == 0.0) continue;
                const double app = a[p + (size_t)n * p], aqq = a[q + (size_t)n * q];
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[k + (size_t)n * p], akq = a[k + (size_t)n * q];
                    a[k + (size_t)n * p] = c * akp - s * akq;
                    a[k + (size_t)n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[p + (size_t)n * k], aqk = a[q + (size_t)n * k];
                    a[p + (size_t)n * k] = c * apk - s * aqk;
                    a[q + (size_t)n * k] = s * apk + c * aqk;
                }
                a[p + (size_t)n * q] = a[q + (size_t)n * p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double vp = V[k + (size_t)ldv * p], vq = V[k + (size_t)ldv * q];
                    V[k + (size_t)ldv * p] = c * vp - s * vq;
                    V[k + (size_t)ldv * q] = s * vp + c * vq;
                }
            }
    }
    for (int i = 0; i < n; ++i) W[i] = a[i + (size_t)n * i];
    for (int i = 0; i < n - 1; ++i) {
        int m = i;
        for (int j = i + 1; j < n; ++j)
            if (W[j] < W[m]) m = j;
        if (m != i) {
            std::swap(W[i], W[m]);
            for (int k = 0; k < n; ++k) std::swap(V[k + (size_t)ldv * i], V[k + (size_t)ldv * m]);
        }
    }
}

}  // namespace pisces
