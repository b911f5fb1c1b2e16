// ref_driver.cpp -- harness that runs the reference PISCES functions that build
// in this image (src/sine_dvr.cpp, src/Davidson.cpp, src/Potential.cpp with
// src/MolPolAux.cpp and src/DerivDDTensor.cpp) on inputs written by
// tests/golden/make_golden.py, so that the CPU restatement in pisces_oracle.c
// can be checked against the reference itself.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile.ref into oracle/_ref/;
// the reference sources are compiled where they lie and never copied.
//
// File format (little endian): a sequence of records, each an int64 element
// count followed by that many float64 values (integers are stored as float64).
//
// Commands:
//   ref_driver tkin      <in> <out>   in: [n, L]            out: [x(n)], [Tpacked]
//   ref_driver potential <in> <out>   in: flags, potpara, sites, q, iq, d, id,
//                                         alphas, ips, epc, [vpolww], pts
//                                     out: 5 energies per point
//   ref_driver davidson  <in> <out>   in: [nx,ny,nz,Lx,Ly,Lz], V, [nroots,maxsub,
//                                         maxmacro,ptol,corrflag], start vectors
//                                     out: [nconv,n_mtx], evals, vectors
//   ref_driver kin       <in> <out>   in: [type, n, para] (type 1: ho_dvr with
//                                     frequency para; 20: cm_ham on [-para/2, para/2])
//                                     out: [x(n)], [T full n*n]
//   ref_driver damp      <in> <out>   in: [r], [alpha_i], [alpha_j], [a]
//                                     out: ElectrostaticDCDamp, ElectrostaticDDDamp
//                                     (src/DPP.cpp:1798-1812) per entry
//   ref_driver input     <file> <out> the reference's TSIN + GetInputParameters +
//                                     ReadWatersGroup on an input file
//                                     out: [runtype, nWater, DVRType, Sampling,
//                                     DiagMethod, nStates, maxSub, maxIter, ptol,
//                                     istartvec], PotFlag[6], PotPara[11],
//                                     ngrid[3], gpara[3], water coordinates
//   ref_driver water     <in> <out>   in: [nW, CenterFlag, KTFlag], [O,H1,H2 xyz in
//                                     Angstrom, 9*nW]
//                                     the water model exactly as ClusterAnion runs it
//                                     (src/ClusterAnion.cpp:59-66, 165-197):
//                                     SetStructure -> GetConfiguration ->
//                                     SetConfiguration -> CalcEnergy -> GetLists ->
//                                     ReportInducedDipoles -> ReportPolSitesAndField
//                                     out: conf[6nW], sites[12nW], q, iq, d (one per
//                                     water), id, all dipoles[9nW], their sites,
//                                     alpha[3nW], ipp[3nW], Epc[9nW],
//                                     [E0, PolarizationEnergy]
//   ref_driver grad      <in> <out>   in: [nW], O,H1,H2 xyz (Angstrom), [nx,ny,nz,Lx,Ly,Lz],
//                                     psi (ngp, x fastest), flags[6], potpara[11]
//                                     the analytic gradient of ClusterAnion::GetAnalGrad
//                                     (src/ClusterAnion.cpp:322-430) after the setup of
//                                     EnergyFromConfiguration (:165-246), with the grid
//                                     loop of DVR::ComputeGradient (src/DVR.cpp:1099-1176)
//                                     over the reference's Potential::EvaluateGradient,
//                                     FinalGradient and SubtractWWGradient (src/Gradients.cpp)
//                                     out: Gel = sum psi^2 tGrad [3 nSites], tmu [3 nPol],
//                                     mu_cross_mu [(3 nPol)^2], dT_x, dT_y, dT_z, dEfield,
//                                     PolGrad, Gww (CalcGradient), Gfinal, analgrad [6 nW], [E0]
#include <mpi.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "vecdefs.h"
#include "Potential.h"
#include "sine_dvr.h"
#include "Davidson.h"
#include "tsin.h"
#include "ho_dvr.h"
#include "cm_dvr.h"
#include "Parameters.h"
#include "GetInput.h"
#include "GTO.h"
#include "MO.h"
#include "AtomCenter.h"
#include "Water.h"
#include "DPP.h"

typedef std::vector<double> rec;

static std::vector<rec> read_records(const char *path)
{
    std::vector<rec> out;
    FILE *f = fopen(path, "rb");
    if (!f) { perror(path); exit(2); }
    int64_t n;
    while (fread(&n, sizeof(n), 1, f) == 1) {
        rec r(n);
        if (n && fread(r.data(), sizeof(double), n, f) != (size_t)n) { fprintf(stderr, "short read\n"); exit(2); }
        out.push_back(r);
    }
    fclose(f);
    return out;
}

static void write_records(const char *path, const std::vector<rec> &rs)
{
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); exit(2); }
    for (const rec &r : rs) {
        int64_t n = (int64_t)r.size();
        fwrite(&n, sizeof(n), 1, f);
        if (n) fwrite(r.data(), sizeof(double), n, f);
    }
    fclose(f);
}

static std::vector<int> as_int(const rec &r)
{
    std::vector<int> v(r.size());
    for (size_t i = 0; i < r.size(); ++i) v[i] = (int)llround(r[i]);
    return v;
}

static int cmd_tkin(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    int n = (int)r[0][0];
    double L = r[0][1];
    rec x(n), t(n * (n + 1) / 2);
    sine_dvr(n, 1.0, -0.5 * L, 0.5 * L, x.data(), t.data(), 0, 0);
    write_records(out, {x, t});
    return 0;
}

static int cmd_potential(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    std::vector<int> flags = as_int(r[0]);
    const rec &potpara = r[1], &sites = r[2], &q = r[3];
    std::vector<int> iq = as_int(r[4]);
    const rec &d = r[5];
    std::vector<int> id = as_int(r[6]);
    const rec &alphas = r[7];
    std::vector<int> ips = as_int(r[8]);
    const rec &epc = r[9];
    double vpolww = r[10].empty() ? 0.0 : r[10][0];
    const rec &pts = r[11];
    iVec pf(flags.begin(), flags.end());
    rec dmu(3 * (sites.size() / 3) * 3 * (d.size() / 3) + 1, 0.0);
    rec pp(potpara);
    pp.resize(32, 0.0);
    Potential V;
    V.VpolWaterWater = vpolww;
    V.Setup(pf, (int)sites.size() / 3, sites.data(), (int)q.size(), q.data(), iq.data(),
            (int)d.size() / 3, d.data(), id.data(), (int)alphas.size(), alphas.data(),
            ips.data(), epc.data(), dmu.data(), pp.data());
    size_t npts = pts.size() / 3;
    rec e(5 * npts);
    for (size_t p = 0; p < npts; ++p) {
        V.Evaluate(&pts[3 * p]);
        V.ReportEnergies(5, &e[5 * p]);
    }
    write_records(out, {e});
    return 0;
}

// H*x with full T matrices (harness code; the reference's own
// DVR::MatrixTimesVector lives in a translation unit that needs FFTW).
static void hx(int n[3], double *T[3], const double *V, const double *x, double *y)
{
    long ng = (long)n[0] * n[1] * n[2];
    for (long g = 0; g < ng; ++g) y[g] = V[g] * x[g];
    long st[3] = {1, n[0], (long)n[0] * n[1]};
    for (int ax = 0; ax < 3; ++ax) {
        int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
        for (int b = 0; b < n[o2]; ++b)
            for (int a = 0; a < n[o1]; ++a) {
                long base = a * st[o1] + b * st[o2];
                for (int i = 0; i < n[ax]; ++i) {
                    double s = 0;
                    for (int j = 0; j < n[ax]; ++j) s += T[ax][i + n[ax] * j] * x[base + j * st[ax]];
                    y[base + i * st[ax]] += s;
                }
            }
    }
}

static int cmd_davidson(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    int n[3] = {(int)r[0][0], (int)r[0][1], (int)r[0][2]};
    double L[3] = {r[0][3], r[0][4], r[0][5]};
    const rec &V = r[1];
    std::vector<int> p = as_int(r[2]);
    int nroots = p[0], maxsub = p[1], maxmacro = p[2], ptol = p[3], corrflag = p[4];
    const rec &start = r[3];
    long ng = (long)n[0] * n[1] * n[2];
    double *T[3];
    rec Tm[3], xg[3];
    for (int a = 0; a < 3; ++a) {
        Tm[a].resize((size_t)n[a] * n[a]);
        xg[a].resize(n[a]);
        sine_dvr(n[a], 1.0, -0.5 * L[a], 0.5 * L[a], xg[a].data(), Tm[a].data(), n[a], 0);
        T[a] = Tm[a].data();
    }
    rec diag(ng);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                long g = i + (long)n[0] * (j + (long)n[1] * k);
                diag[g] = V[g] + T[0][i * (n[0] + 1)] + T[1][j * (n[1] + 1)] + T[2][k * (n[2] + 1)];
            }
    rec B(ng * maxsub, 0.0), Z(ng * maxsub, 0.0);
    std::copy(start.begin(), start.begin() + ng * nroots, B.begin());
    rec work(DavidsonWorkSize((int)ng, maxsub, nroots, corrflag));
    rec ev(nroots, 0.0);
    int inout[3] = {0, 0, 0}, nconv = 0, nmtx = 0, ido = 1;
    // driver loop in the manner of DVR::davdriver (src/davdriver.cpp:58-86)
    while (ido != 0) {
        ido = Davidson((int)ng, maxsub, nroots, maxmacro, ptol, corrflag, 0, ev.data(), nconv,
                       B.data(), Z.data(), diag.data(), work.data(), inout);
        if (ido == 1) {
            for (int v = 0; v < inout[2]; ++v) {
                hx(n, T, V.data(), &B[(inout[0] + v) * ng], &Z[(inout[1] + v) * ng]);
                ++nmtx;
            }
        }
    }
    rec meta = {(double)nconv, (double)nmtx};
    rec vecs(B.begin(), B.begin() + ng * nroots);
    write_records(out, {meta, ev, vecs});
    return 0;
}

// defined in src/DPP.cpp (declared there, not in a header)
double ElectrostaticDCDamp(double distance, double alphai, double alphaj, double a);
double ElectrostaticDDDamp(double distance, double alphai, double alphaj, double a);

static int cmd_damp(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    const size_t n = r[0].size();
    rec dc(n), dd(n);
    for (size_t i = 0; i < n; ++i) {
        dc[i] = ElectrostaticDCDamp(r[0][i], r[1][i], r[2][i], r[3][i]);
        dd[i] = ElectrostaticDDDamp(r[0][i], r[1][i], r[2][i], r[3][i]);
    }
    write_records(out, {dc, dd});
    return 0;
}

static int cmd_kin(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    const int type = (int)r[0][0], n = (int)r[0][1];
    const double para = r[0][2];
    rec x(n), t((size_t)n * n), z((size_t)n * n);
    if (type == 1) ho_dvr(n, para, x.data(), t.data(), n, z.data(), 1, 0);
    else cm_ham(n, 1.0, -0.5 * para, 0.5 * para, t.data(), x.data(), n, 0);
    write_records(out, {x, t});
    return 0;
}

static int cmd_input(const char *in, const char *out)
{
    TSIN Input;
    Input.ReadFromFile(in, 0);
    Parameters P;
    GetInputParameters(P, Input);
    dVec xyz;
    int nw = ReadWatersGroup(Input, xyz);
    if (P.nWater == 0) P.nWater = nw;
    rec a = {(double)P.runtype, (double)P.nWater, (double)P.DVRType, (double)P.Sampling,
             (double)P.DiagMethod, (double)P.nStates, (double)P.maxSub, (double)P.maxIter,
             (double)P.ptol, (double)P.istartvec};
    rec f(6), pp(11), ng(3), gp(3);
    for (int k = 0; k < 6; ++k) f[k] = P.PotFlag[k];
    for (int k = 0; k < 11; ++k) pp[k] = P.PotPara[k];
    for (int k = 0; k < 3; ++k) { ng[k] = P.ngrid[k]; gp[k] = P.gpara[k]; }
    rec w(xyz.begin(), xyz.begin() + 9 * nw);
    write_records(out, {a, f, pp, ng, gp, w});
    return 0;
}

static rec as_rec(const int *v, int n) { return rec(v, v + n); }

static int cmd_water(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    const int nw = (int)r[0][0], center = (int)r[0][1], kt = (int)r[0][2];
    const rec &xyz = r[1];
    WaterCluster W;
    W.SetStructure(nw, xyz.data(), 1, center, kt, 0);
    rec conf(6 * nw);
    W.GetConfiguration(nw, conf.data(), 1.0, 0);
    W.SetConfiguration(nw, conf.data(), 1, 0);
    const double e0 = W.CalcEnergy(0);
    const double epol = W.ReportPolarization();
    int ns = 0, nq = 0, nd = 0;
    W.ReportNoOfSites(ns, nq, nd);
    const int np = W.ReportNoOfPolSites();
    rec sites(3 * ns), q(nq), d(3 * nd), dmu(3 * ns * 3 * nd + 1, 0.0);
    std::vector<int> iq(nq), id(nd);
    W.GetLists(ns, sites.data(), nq, q.data(), iq.data(), nd, d.data(), id.data(), dmu.data(), 0);
    rec dall(9 * nd);
    std::vector<int> idall(3 * nd);
    W.ReportInducedDipoles(3 * nd, dall.data(), idall.data(), 0);
    rec alpha(np), epc(3 * np);
    std::vector<int> ipp(np);
    W.ReportPolSitesAndField(np, alpha.data(), ipp.data(), epc.data(), 0);
    write_records(out, {conf, sites, q, as_rec(iq.data(), nq), d, as_rec(id.data(), nd), dall,
                        as_rec(idall.data(), 3 * nd), alpha, as_rec(ipp.data(), np), epc,
                        rec{e0, epol}});
    return 0;
}

void BuildSuperdT(int nPolSites, const double *R, const double *alpha, double aThole,
                  double *dT_x, double *dT_y, double *dT_z);

static int cmd_grad(const char *in, const char *out)
{
    std::vector<rec> r = read_records(in);
    const int nw = (int)r[0][0];
    const rec &xyz = r[1];
    const int n[3] = {(int)r[2][0], (int)r[2][1], (int)r[2][2]};
    const double L[3] = {r[2][3], r[2][4], r[2][5]};
    const rec &psi = r[3];
    std::vector<int> flags = as_int(r[4]);
    rec pp(r[5]);
    pp.resize(32, 0.0);
    // EnergyFromConfiguration: the water model and the electron's Potential
    WaterCluster W;
    W.SetStructure(nw, xyz.data(), 1, 0, 0, 0);
    rec conf(6 * nw);
    W.GetConfiguration(nw, conf.data(), 1.0, 0);
    W.SetConfiguration(nw, conf.data(), 1, 0);
    const double e0 = W.CalcEnergy(0);
    Potential V;
    V.VpolWaterWater = W.ReportPolarization();
    int ns = 0, nq = 0, nd = 0;
    W.ReportNoOfSites(ns, nq, nd);
    const int np = W.ReportNoOfPolSites();
    rec sites(3 * ns), q(nq), d(9 * nd), dmu(3 * ns * 3 * nd + 1, 0.0);
    std::vector<int> iq(nq), id(3 * nd);
    W.GetLists(ns, sites.data(), nq, q.data(), iq.data(), nd, d.data(), id.data(), dmu.data(), 0);
    W.ReportInducedDipoles(3 * nd, d.data(), id.data(), 0);
    rec alpha(np), epc(3 * np);
    std::vector<int> ipp(np);
    W.ReportPolSitesAndField(np, alpha.data(), ipp.data(), epc.data(), 0);
    iVec pf(flags.begin(), flags.end());
    V.Setup(pf, ns, sites.data(), nq, q.data(), iq.data(), 3 * nd, d.data(), id.data(), np,
            alpha.data(), ipp.data(), epc.data(), dmu.data(), pp.data());
    // GetAnalGrad
    rec dEfield((size_t)3 * ns * nq * 3), dTensor((size_t)3 * ns * 3 * nd * 3);
    rec dTx((size_t)9 * np * np), dTy((size_t)9 * np * np), dTz((size_t)9 * np * np);
    rec grad(3 * ns, 0.0), polgrad(3 * ns, 0.0);
    W.SetConfiguration(nw, conf.data(), 1, 0);
    W.Calc_dInducedDipoles(0, dmu.data(), dEfield.data(), dTensor.data());
    W.CalcGradient(3 * ns, grad.data(), polgrad.data(), 0);
    const rec gww(grad);
    W.ReportPolSitesAndField(np, alpha.data(), ipp.data(), epc.data(), 0);
    rec rpps(3 * np);
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < 3; ++k) rpps[3 * i + k] = sites[3 * ipp[i] + k];
    BuildSuperdT(np, rpps.data(), alpha.data(), 0.3, dTx.data(), dTy.data(), dTz.data());
    // DVR::ComputeGradient: the grid loop (x fastest) over EvaluateGradient
    rec xg[3], Tm[3];
    for (int a = 0; a < 3; ++a) {
        xg[a].resize(n[a]);
        Tm[a].resize((size_t)n[a] * n[a]);
        sine_dvr(n[a], 1.0, -0.5 * L[a], 0.5 * L[a], xg[a].data(), Tm[a].data(), n[a], 0);
    }
    const int n3 = 3 * np;
    dVec tGrad(3 * ns), tmu(n3), mcm((size_t)n3 * n3, 0.0);
    rec gel(3 * ns, 0.0);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const long g = i + (long)n[0] * (j + (long)n[1] * k);
                const double pt[3] = {xg[0][i], xg[1][j], xg[2][k]};
                std::fill(tGrad.begin(), tGrad.end(), 0.0);
                V.EvaluateGradient(pt, tGrad, tmu, mcm, psi[g], W);
                for (int c = 0; c < 3 * ns; ++c) gel[c] += psi[g] * psi[g] * tGrad[c];
            }
    rec loc(gel);
    V.FinalGradient(np, loc.data(), &tmu[0], &mcm[0], dTx.data(), dTy.data(), dTz.data(),
                    dEfield.data());
    for (int c = 0; c < 3 * ns; ++c) grad[c] += loc[c];
    V.SubtractWWGradient(ns, polgrad.data(), grad.data());
    rec ana(6 * nw);
    W.ConvertF(grad.data(), ana.data());
    write_records(out, {gel, rec(tmu.begin(), tmu.end()), rec(mcm.begin(), mcm.end()), dTx, dTy,
                        dTz, dEfield, polgrad, gww, grad, ana, rec{e0}});
    return 0;
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int rc = 1;
    if (argc == 4) {
        if (!strcmp(argv[1], "tkin")) rc = cmd_tkin(argv[2], argv[3]);
        else if (!strcmp(argv[1], "potential")) rc = cmd_potential(argv[2], argv[3]);
        else if (!strcmp(argv[1], "davidson")) rc = cmd_davidson(argv[2], argv[3]);
        else if (!strcmp(argv[1], "input")) rc = cmd_input(argv[2], argv[3]);
        else if (!strcmp(argv[1], "kin")) rc = cmd_kin(argv[2], argv[3]);
        else if (!strcmp(argv[1], "damp")) rc = cmd_damp(argv[2], argv[3]);
        else if (!strcmp(argv[1], "water")) rc = cmd_water(argv[2], argv[3]);
        else if (!strcmp(argv[1], "grad")) rc = cmd_grad(argv[2], argv[3]);
    } else {
        fprintf(stderr, "usage: ref_driver {tkin|potential|davidson} <in> <out>\n");
    }
    MPI_Finalize();
    return rc;
}
