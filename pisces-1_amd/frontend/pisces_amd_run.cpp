// pisces_amd_run -- the reference's input file in, electron binding energies out.
//
//   pisces_amd_run <input> [--parse-only] [--precision fp64|fp32|bf16x2] [--device N]
//
// The drop-in for `pisces <input>` with runtype = 1 (src/pisces.cpp:112-130,
// ClusterAnion::SinglePoint, src/ClusterAnion.cpp:92-94 and :296-320): read the
// input groups, build the water model and the electron's potential on the GPU,
// solve for the lowest nStates states and print the binding energies in the
// reference's format.
//
// Input reading restates TSIN (src/tsin.cpp): lines beginning with '#' and
// empty lines are dropped; a group is the lines after the first line that
// contains "Begin" + the group name (cut to 15 characters) up to the first line
// that contains "End" + that name; a key is found by substring search inside
// its group and read as "<token> = value" (arrays: "key = a, b, c,").  Missing
// scalar keys take the defaults of GetInputParameters (src/GetInput.cpp:23-186);
// a missing array key is an error, as in the reference.  The Waters group holds
// O, H, H lines in Angstrom (ReadWatersGroup, src/GetInput.cpp:256-320).
//
// Output follows ClusterAnion::EnergyFromConfiguration and SinglePoint
// (src/ClusterAnion.cpp:165-320, 70-99): the neutral cluster's energy E0
// (WaterCluster::CalcEnergy), the energy block (Eneutral, EBE, Etotal in meV),
// the binding energies per state and DVR::ExpectationValues
// (src/DVR.cpp:2329-2449).  One difference, stated where it occurs: the second
// and later start vectors come from a seeded generator (the reference calls the
// unseeded rand()).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "pisces_hip.h"

namespace {

constexpr int kSigSigns = 15;   // significant characters of a group name (src/tsin.h)
constexpr int kLineLength = 256;
constexpr double kAu2Mev = 27211.38386;  // src/constants.h

struct InputError {
    std::string msg;
};

class InputFile {
public:
    explicit InputFile(const char *path)
    {
        FILE *f = std::fopen(path, "r");
        if (!f) throw InputError{std::string("cannot open ") + path};
        char buf[kLineLength];
        int lineno = 0;
        while (std::fgets(buf, kLineLength, f)) {
            const size_t l = std::strlen(buf);
            ++lineno;
            // a line that does not fit the buffer would be split into two lines
            if (l == kLineLength - 1 && buf[l - 1] != '\n' && !std::feof(f)) {
                std::fclose(f);
                throw InputError{std::string(path) + ": line " + std::to_string(lineno) +
                                 " is longer than " + std::to_string(kLineLength - 2) +
                                 " characters"};
            }
            if (buf[0] == '#' || l == 1) continue;
            lines_.emplace_back(buf);
        }
        std::fclose(f);
    }

    // lines strictly between Begin<group> and End<group>; -1 if no Begin line
    int find_group(const char *group, int *end) const
    {
        const std::string b = "Begin" + std::string(group).substr(0, kSigSigns);
        const std::string e = "End" + std::string(group).substr(0, kSigSigns);
        int i = 0;
        while (i < (int)lines_.size() && lines_[i].find(b) == std::string::npos) ++i;
        if (i == (int)lines_.size()) return -1;
        int j = i + 1;
        while (j < (int)lines_.size() && lines_[j].find(e) == std::string::npos) ++j;
        *end = j;
        return i + 1;
    }

    std::vector<std::string> group(const char *group) const
    {
        int end = 0;
        const int b = find_group(group, &end);
        if (b < 0) throw InputError{std::string("group ") + group + " not found"};
        return std::vector<std::string>(lines_.begin() + b, lines_.begin() + end);
    }

    // the text from the key's first occurrence inside the group, or nullptr
    const char *key_text(const char *group, const char *key, int *line) const
    {
        int end = 0;
        const int b = find_group(group, &end);
        if (b < 0) return nullptr;
        for (int i = b; i < end; ++i) {
            const size_t p = lines_[i].find(key);
            if (p != std::string::npos) {
                *line = i;
                return lines_[i].c_str() + p;
            }
        }
        return nullptr;
    }

    int get_int(const char *group, const char *key, int def) const
    {
        int li = 0;
        const char *s = key_text(group, key, &li);
        if (!s) return def;
        int v = 0;
        if (std::sscanf(s, "%*s = %i", &v) != 1) throw bad_value(group, key, li);
        return v;
    }

    double get_double(const char *group, const char *key, double def) const
    {
        int li = 0;
        const char *s = key_text(group, key, &li);
        if (!s) return def;
        double v = 0;
        if (std::sscanf(s, "%*s = %lf", &v) != 1) throw bad_value(group, key, li);
        return v;
    }

    // "key = v0, v1, ..., " (src/tsin.cpp:400-546)
    template <typename T>
    void get_array(const char *group, const char *key, T *fill, int n) const
    {
        int li = 0;
        const char *s = key_text(group, key, &li);
        if (!s) throw InputError{std::string("key ") + key + " not found in group " + group};
        const char *fmt0 = std::is_same<T, int>::value ? "%*s = %i," : "%*s = %lf,";
        const char *fmtk = std::is_same<T, int>::value ? ", %i," : ", %lf,";
        if (std::sscanf(s, fmt0, &fill[0]) != 1) throw bad_value(group, key, li);
        for (int k = 1; k < n; ++k) {
            ++s;  // s sits on the previous comma (or the key), never on the terminator
            while (*s && *s != ',') ++s;
            if (!*s) throw bad_value(group, key, li);
            if (std::sscanf(s, fmtk, &fill[k]) != 1) throw bad_value(group, key, li);
        }
    }

private:
    InputError bad_value(const char *group, const char *key, int li) const
    {
        return InputError{std::string("cannot read ") + key + " in group " + group + " from: " +
                          lines_[li]};
    }
    std::vector<std::string> lines_;
};

// GetInputParameters (src/GetInput.cpp:23-186): the keys of a single point
struct Params {
    int runtype = 0, nElectron = 1, nWater = 0, KTFlag = 0, CenterFlag = 0;
    int PotFlag[6] = {0, 0, 0, 0, 0, 0};
    double PotPara[11] = {0};
    int DVRType = 0, Sampling = 1, ngrid[3] = {0, 0, 0};
    double gpara[3] = {0, 0, 0};
    int DiagMethod = 2, nStates = 1, maxSub = 20, maxIter = 100, ptol = 5, istartvec = 1;
};

Params read_params(const InputFile &in)
{
    Params P;
    P.runtype = in.get_int("Job", "runtype", 0);
    P.nElectron = in.get_int("Job", "nElectron", 1);
    P.nWater = in.get_int("WaterModel", "NoOfWaters", 0);
    P.KTFlag = in.get_int("WaterModel", "KTCharges", 0);
    P.CenterFlag = in.get_int("WaterModel", "CoMOrigin", 0);
    if (P.nElectron > 0) {
        const char *g = "ElectronPotential";
        P.PotFlag[0] = in.get_int(g, "Potential", 3);
        P.PotFlag[1] = in.get_int(g, "DampType", 2);
        P.PotFlag[2] = in.get_int(g, "SigmaOFlag", 0);
        P.PotFlag[3] = in.get_int(g, "Polarization", 1);
        P.PotFlag[4] = in.get_int(g, "AdiabaticPolPot", 0);
        P.PotFlag[5] = in.get_int(g, "InternalParam", 0);
        if (P.PotFlag[0] < 1 || P.PotFlag[0] > 4)
            throw InputError{"Potential " + std::to_string(P.PotFlag[0]) +
                             ": only the DPP water potentials 1-4 are supported"};
        P.PotPara[0] = in.get_double(g, "Alpha", 9.745);
        P.PotPara[1] = in.get_double(g, "PolDamping", 2.536);
        P.PotPara[2] = in.get_double(g, "RepCoreScale", 4.4837);
        P.PotPara[3] = in.get_double(g, "CoulombDamping", 1.7724);
        P.PotPara[4] = in.get_double(g, "DipoleDamping", 1.7724);
        P.PotPara[5] = in.get_double(g, "ORepCoreScale", 4.4837);
        P.PotPara[6] = in.get_double(g, "EpsDrude", 0.5 / std::sqrt(P.PotPara[0]));
        P.PotPara[10] = in.get_double(g, "Rtol", 20.0);
        if (P.PotFlag[5] == 1) {
            // the internal parameter sets per polarization model (src/GetInput.cpp:86-116)
            P.PotFlag[0] = 3;
            P.PotFlag[1] = 2;
            static const double pol[4] = {5.011, 5.1543, 5.005, 4.84};
            static const double rep[4] = {6.6396, 6.6717, 6.6382, 6.602};
            const int m = P.PotFlag[3];
            if (m >= 1 && m <= 4) {
                P.PotPara[1] = pol[m - 1];
                P.PotPara[4] = P.PotPara[1];
                P.PotPara[2] = rep[m - 1];
            }
        }
    }
    P.DVRType = in.get_int("GridDef", "DVRType", 0);
    P.Sampling = in.get_int("GridDef", "Sampling", 1);
    in.get_array("GridDef", "NoOfGridPoints", P.ngrid, 3);
    in.get_array("GridDef", P.DVRType == 1 ? "Omega0" : "Length", P.gpara, 3);
    P.DiagMethod = in.get_int("Diag", "Method", 2);
    P.nStates = in.get_int("Diag", "nStates", 1);
    P.maxSub = in.get_int("Diag", "maxSubspace", 20);
    P.maxIter = in.get_int("Diag", "maxIter", 100);
    P.ptol = in.get_int("Diag", "pTol", 5);
    P.istartvec = in.get_int("Diag", "StartVector", 1);
    return P;
}

// ReadWatersGroup (src/GetInput.cpp:256-320): O, H, H per water, Angstrom
std::vector<double> read_waters(const InputFile &in)
{
    const std::vector<std::string> L = in.group("Waters");
    if (L.size() % 3 != 0) throw InputError{"Waters: the number of lines is not a multiple of 3"};
    std::vector<double> xyz;
    for (size_t i = 0; i < L.size(); ++i) {
        char sym[64] = {0};
        double x, y, z;
        if (std::sscanf(L[i].c_str(), " %63s %lf %lf %lf", sym, &x, &y, &z) != 4)
            throw InputError{"Waters: cannot read line " + std::to_string(i) + ": " + L[i]};
        const bool oxygen = i % 3 == 0;
        if (oxygen ? (sym[0] != '8' && sym[0] != 'O') : (sym[0] != '1' && sym[0] != 'H'))
            throw InputError{"Waters line " + std::to_string(i) + (oxygen ? ": O or 8" : ": H or 1") +
                             " is expected"};
        xyz.push_back(x);
        xyz.push_back(y);
        xyz.push_back(z);
    }
    return xyz;
}

void print_json(const Params &P, const std::vector<double> &xyz)
{
    std::printf("{\"runtype\": %d, \"nWater\": %d, \"KTCharges\": %d, \"CoMOrigin\": %d, ", P.runtype,
                P.nWater, P.KTFlag, P.CenterFlag);
    std::printf("\"PotFlag\": [%d, %d, %d, %d, %d, %d], \"PotPara\": [", P.PotFlag[0], P.PotFlag[1],
                P.PotFlag[2], P.PotFlag[3], P.PotFlag[4], P.PotFlag[5]);
    for (int k = 0; k < 11; ++k) std::printf("%s%.17g", k ? ", " : "", P.PotPara[k]);
    std::printf("], \"DVRType\": %d, \"Sampling\": %d, \"ngrid\": [%d, %d, %d], ", P.DVRType,
                P.Sampling, P.ngrid[0], P.ngrid[1], P.ngrid[2]);
    std::printf("\"gpara\": [%.17g, %.17g, %.17g], ", P.gpara[0], P.gpara[1], P.gpara[2]);
    std::printf("\"DiagMethod\": %d, \"nStates\": %d, \"maxSub\": %d, \"maxIter\": %d, \"ptol\": %d, "
                "\"StartVector\": %d, \"waters\": [",
                P.DiagMethod, P.nStates, P.maxSub, P.maxIter, P.ptol, P.istartvec);
    for (size_t k = 0; k < xyz.size(); ++k) std::printf("%s%.17g", k ? ", " : "", xyz[k]);
    std::printf("]}\n");
}

// DVR::ExpectationValues (src/DVR.cpp:2329-2449) for the converged states:
// normalisation, <|r|> = |<r>|, sqrt(<r^2>) and the spread sqrt(<r^2> - <|r|>^2)
// in Angstrom, and the transition dipoles <n|r|0> in atomic units.
void expectation_values(const int n[3], const double *grid, const double *wf, int nconv)
{
    const double *xg = grid, *yg = grid + n[0], *zg = grid + n[0] + n[1];
    const long long ngp = (long long)n[0] * n[1] * n[2];
    std::vector<double> norm(nconv, 0.0), xyz(3 * nconv, 0.0), r2(nconv, 0.0),
        tdip(3 * (size_t)std::max(nconv - 1, 0), 0.0);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const long long g = i + (long long)n[0] * (j + (long long)n[1] * k);
                const double q[3] = {xg[i], yg[j], zg[k]};
                for (int s = 0; s < nconv; ++s) {
                    const double w = wf[s * ngp + g], rho = w * w;
                    norm[s] += rho;
                    for (int a = 0; a < 3; ++a) xyz[3 * s + a] += rho * q[a];
                    r2[s] += rho * (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                    if (s > 0)
                        for (int a = 0; a < 3; ++a) tdip[3 * (s - 1) + a] += w * q[a] * wf[g];
                }
            }
    const double b2a = 0.52917720859;
    std::printf("\nState     <|r|>     sqrt(<r^2>)   variance  (all in Angstrom)\n");
    for (int s = 0; s < nconv; ++s) {
        if (std::fabs(norm[s] - 1.0) > 1e-8)
            std::printf("Warning: normaliziation integral of state %d is not 1.0, but %g\n"
                        "This chould not happen.\n", s, norm[s]);
        std::printf("normalization of state%d is %g\n", s, norm[s]);
        double r = 0;
        for (int a = 0; a < 3; ++a) r += xyz[3 * s + a] * xyz[3 * s + a];
        r = std::sqrt(r);
        std::printf(" %3i   %10.5f  %10.5f  %10.5f\n", s, b2a * r, b2a * std::sqrt(r2[s]),
                    b2a * std::sqrt(r2[s] - r * r));
    }
    std::printf("\n");
    if (nconv > 1) {
        std::printf("Transition dipoles d^2 and d=(<n|x|0>, <n|y|0>, <n|z|0>) (all in au)\n");
        for (int s = 1; s < nconv; ++s) {
            const double *d = &tdip[3 * (s - 1)];
            std::printf(" %3i      %10.5f     (%10.5f,  %10.5f,  %10.5f)\n", s,
                        d[0] * d[0] + d[1] * d[1] + d[2] * d[2], d[0], d[1], d[2]);
        }
        std::printf("\n");
    }
}

int fail(const std::string &msg)
{
    std::fprintf(stderr, "pisces_amd_run: %s\n", msg.c_str());
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *path = nullptr;
    bool parse_only = false;
    int precision = PISCES_FP64, device = 0;
    for (int a = 1; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--parse-only")) parse_only = true;
        else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) device = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--precision") && a + 1 < argc) {
            const char *p = argv[++a];
            precision = !std::strcmp(p, "fp32") ? PISCES_FP32
                        : !std::strcmp(p, "bf16x2") ? PISCES_BF16X2 : PISCES_FP64;
        } else path = argv[a];
    }
    if (!path) {
        std::printf("usage: pisces_amd_run input-file [--parse-only] [--precision fp64|fp32|bf16x2]\n");
        return 1;
    }
    Params P;
    std::vector<double> xyz;
    try {
        InputFile in(path);
        P = read_params(in);
        xyz = read_waters(in);
        const int nw = (int)(xyz.size() / 9);
        if (P.nWater == 0) P.nWater = nw;
        else if (nw != P.nWater)
            throw InputError{"inconsistent no of waters in Waters and WaterModel groups"};
    } catch (const InputError &e) {
        return fail(e.msg);
    }
    if (parse_only) {
        print_json(P, xyz);
        return 0;
    }
    if (P.runtype != 1) return fail("runtype " + std::to_string(P.runtype) + ": only single points (1) run here");
    if (P.nElectron != 1) return fail("nElectron must be 1");
    if ((P.DVRType != 0 && P.DVRType != 1 && P.DVRType != 3 && P.DVRType != 20) || P.Sampling != 1)
        return fail("DVRType 0 (sine), 1 (oscillator), 3 (FFT) or 20 (Colbert-Miller), with Sampling 1");
    if (P.KTFlag != 0 || P.CenterFlag != 0) return fail("KTCharges and CoMOrigin must be 0");
    // DVR::Diagonalize (src/DVR.cpp:462-471): Method 2 Davidson, 3 Jacobi-Davidson, 4/5 none
    int corrflag;
    switch (P.DiagMethod) {
    case 2: corrflag = 1; break;
    case 3: corrflag = 2; break;
    case 4:
    case 5: corrflag = 0; break;
    default: return fail("Diag Method " + std::to_string(P.DiagMethod) + " is not supported (2-5 are)");
    }
    if (P.istartvec != 1) return fail("StartVector 1 (particle in a box) is the one supported here");

    // Polarization 6, the dual-grid method (src/ClusterAnion.cpp:207-299): a
    // solve on the coarse grid with self-consistent polarization, then the
    // fine grid's potential with the coarse one carried over far from the
    // sites, started from the coarse ground state
    const bool dual = P.PotFlag[3] == 6;
    int flags[6];
    std::copy(P.PotFlag, P.PotFlag + 6, flags);
    if (dual) flags[3] = 3;
    const double L[3] = {P.gpara[0], P.gpara[1], P.gpara[2]};
    int rc;
    int nc[3] = {0, 0, 0};
    std::vector<double> coarse_v, coarse_wf;
    if (dual) {
        if (P.DVRType != 0) return fail("the dual-grid method (Polarization 6) runs on the sine DVR");
        double Lc[3];
        if (pisces_dual_grid_coarse(P.ngrid, L, nc, Lc))
            return fail(std::string("dual grid: ") + pisces_last_error());
        pisces_ctx *cc = nullptr;
        if (pisces_create(&cc, nc, Lc, device)) return fail(std::string("coarse grid: ") + pisces_last_error());
        rc = pisces_build_cluster_potential(cc, P.nWater, xyz.data(), flags, P.PotPara, 0, nullptr,
                                            nullptr, 2.9, nullptr, nullptr);
        const long long ncg = (long long)nc[0] * nc[1] * nc[2];
        std::vector<double> cv((size_t)ncg * P.nStates), cev(P.nStates);
        pisces_piab(cc, cv.data());
        std::mt19937_64 cgen(54321);
        std::uniform_real_distribution<double> cu(-1.0, 1.0);
        for (size_t k = (size_t)ncg; k < cv.size(); ++k) cv[k] = cu(cgen);
        pisces_stats cst{};
        if (rc >= 0)
            rc = pisces_solve(cc, P.nStates, P.maxSub, P.maxIter, P.ptol, corrflag, precision,
                              cv.data(), cev.data(), &cst);
        coarse_v.resize((size_t)ncg);
        if (rc >= 0) pisces_get_potential(cc, coarse_v.data());
        pisces_destroy(cc);
        if (rc < 0) return fail(std::string("coarse grid: ") + pisces_last_error());
        std::printf("  ***  Energies For pre-calculations    ***  \n");
        std::printf("       EBE      = %g meV\n", cev[0] * kAu2Mev);
        std::printf("  coarse grid %d^3 on %.6g bohr: %d matvecs\n", nc[0], Lc[0], cst.n_matvec);
        coarse_wf.assign(cv.begin(), cv.begin() + ncg);
    }
    pisces_ctx *ctx = nullptr;
    rc = pisces_create_ex(&ctx, P.ngrid, L, device, P.DVRType);
    if (rc) return fail(std::string("pisces_create_ex: ") + pisces_last_error());
    std::vector<double> sites((size_t)12 * P.nWater), mu((size_t)9 * P.nWater);
    rc = pisces_build_cluster_potential(ctx, P.nWater, xyz.data(), flags, P.PotPara, 0, nullptr,
                                        nullptr, 2.9, sites.data(), mu.data());
    if (rc >= 0 && dual && pisces_dual_grid_potential(ctx, nc, coarse_v.data(), P.PotPara[10]) < 0)
        rc = -1;
    if (rc < 0) {
        pisces_destroy(ctx);
        return fail(std::string("potential: ") + pisces_last_error());
    }
    if (rc > 0) std::printf("Warning: the induced dipoles of the water model did not converge\n");
    // WaterCluster::CalcEnergy (src/ClusterAnion.cpp:169-175)
    double ew[4];
    pisces_water_energy(P.nWater, sites.data(), mu.data(), ew);
    const double E0 = ew[0];
    std::printf("ElectrostaticEnergy = %g\n", ew[1]);
    std::printf("Energy of the neutral cluster E0 = %g\n", E0);
    const long long ngp = (long long)P.ngrid[0] * P.ngrid[1] * P.ngrid[2];
    std::vector<double> vecs((size_t)ngp * P.nStates), ev(P.nStates);
    // start: the half-box particle-in-a-box function (or, for the dual grid,
    // the coarse ground state carried over: SVFlag 3), then uniform [-1, 1)
    // vectors (src/DVR.cpp:369-513)
    if (dual) pisces_dual_grid_extend(ctx, nc, coarse_wf.data(), 0, vecs.data());
    else pisces_piab(ctx, vecs.data());
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (size_t k = (size_t)ngp; k < vecs.size(); ++k) vecs[k] = u(gen);
    pisces_stats st{};
    rc = pisces_solve(ctx, P.nStates, P.maxSub, P.maxIter, P.ptol, corrflag, precision, vecs.data(),
                      ev.data(), &st);
    std::vector<double> grid((size_t)P.ngrid[0] + P.ngrid[1] + P.ngrid[2]);
    pisces_grid(ctx, grid.data(), nullptr);
    pisces_destroy(ctx);
    if (rc < 0) return fail(std::string("solve: ") + pisces_last_error());
    const int nconv = rc;
    if (nconv < 1) {
        std::printf("GetClusterEnergy:: Convergence failure.\n");
        return 1;
    }
    // the energy block (src/ClusterAnion.cpp:303-315)
    const double EBE = ev[0];
    std::printf("                                              \n");
    std::printf("  ------------------------------------------  \n");
    std::printf("  ***  Energies For The Current Geometry ***  \n");
    std::printf("  ------------------------------------------  \n");
    std::printf("       Eneutral = %g meV\n", E0 * kAu2Mev);
    std::printf("       EBE      = %g meV\n", EBE * kAu2Mev);
    std::printf("       Etotal   = %g meV\n", (EBE + E0) * kAu2Mev);
    std::printf("  ------------------------------------------  \n");
    std::printf("  **************  END ENERGY ***************\n");
    std::printf("                                              \n");
    std::printf("\nElectron binding energies:\n");
    for (int s = 0; s < P.nStates; ++s)
        std::printf("State %-3i  %12.8f Hartree =  %10.2f meV\n", s, ev[s], ev[s] * kAu2Mev);
    if (nconv < P.nStates) {
        std::printf("single-point:: Only %d States of %d converged.\n", nconv, P.nStates);
        return 1;
    }
    expectation_values(P.ngrid, grid.data(), vecs.data(), nconv);
    std::printf("  matvecs %d, iterations %d, restarts %d, %.3f s (EBE %.12f Eh, E0 %.17g Eh)\n",
                st.n_matvec, st.n_iter, st.n_macro, st.seconds, EBE, E0);
    return 0;
}
