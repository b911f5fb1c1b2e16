// dual.hip -- the coarse -> fine transfer of the dual-grid method.
//
// ClusterAnion's dual-grid single point (src/ClusterAnion.cpp:207-299) first
// solves on a coarse grid -- every second fine point for odd n, every third
// (shifted by checkM / 2) for even n, cubic grids only -- and then carries the
// coarse ground state (ExtendWfDouble / ExtendWfThree, src/DVR.cpp:1446-1722)
// and, with Polarization 6, the coarse potential far from every site
// (InterpolVdouble / InterpolVtriple, src/DVR.cpp:1240-1421, used by
// ComputePotential at :776-830) over to the fine grid.  One thread per fine
// point: a coarse value where the point lies on the coarse lattice, else the
// weighted sum of the 2, 4 or 8 coarse corners around it -- midpoint averages
// for double spacing, the reference's inverse-distance weights (six-digit
// constants; the wavefunction's table differs from the potential's in three
// entries, as the reference's does) for triple spacing, where points outside
// the coarse grid's hull are set to 0.
#include "common.hpp"
#include "context.hpp"

#include <cmath>
#include <vector>

namespace pisces {

namespace {

// weights of the corners (a, b, c), index a + 2 b + 4 c, for the residues
// (rx, ry, rz) of a fine point in the coarse cell (triple spacing); kind 0 the
// wavefunction (ExtendWfThree), 1 the potential (InterpolVtriple)
void triple_table(int kind, double *tab /* 27 x 8 */)
{
    const double w1 = 0.707107 / 1.955087, w2 = 0.447214 / 1.955087, w3 = 0.353553 / 1.955087;
    const double g1 = 0.57735 / 3.09077, g2 = 0.408248 / 3.09077, g3 = 0.33333 / 3.09077,
                 g4 = 0.288675 / 3.09077;
    for (int rz = 0; rz < 3; ++rz)
        for (int ry = 0; ry < 3; ++ry)
            for (int rx = 0; rx < 3; ++rx) {
                double *w = tab + 8 * (rx + 3 * ry + 9 * rz);
                for (int c = 0; c < 8; ++c) w[c] = 0.0;
                const int r[3] = {rx, ry, rz};
                const int moving = (rx != 0) + (ry != 0) + (rz != 0);
                if (moving == 0) {
                    w[0] = 1.0;
                } else if (moving == 1) {
                    const int ax = rx ? 0 : (ry ? 1 : 2);
                    w[0] = r[ax] == 1 ? 2.0 / 3.0 : 1.0 / 3.0;
                    w[1 << ax] = r[ax] == 1 ? 1.0 / 3.0 : 2.0 / 3.0;
                } else if (moving == 2) {
                    // axes u < v; the nearest corner gets w1, the far one w3
                    const int u = rx ? 0 : 1, v = rz ? 2 : 1;
                    const int ru = r[u], rv = r[v];
                    double c00, cu, cv, cuv;
                    if (ru == 1 && rv == 1) { c00 = w1; cu = w2; cv = w2; cuv = w3; }
                    else if (ru == 1) { c00 = w2; cu = w3; cv = w1; cuv = w2; }
                    else if (rv == 1) { c00 = w2; cu = w1; cv = w3; cuv = w2; }
                    else { c00 = w3; cu = w2; cv = w2; cuv = w1; }
                    // ExtendWfThree, plane (x, z), residues (2, 1): w1 on the
                    // corner (0, c + 1) (src/DVR.cpp:1570)
                    if (kind == 0 && u == 0 && v == 2 && ru == 2 && rv == 1) cv = w1;
                    w[0] = c00;
                    w[1 << u] = cu;
                    w[1 << v] = cv;
                    w[(1 << u) + (1 << v)] = cuv;
                } else {
                    // 1 / sqrt(d2) by the squared distance in fine spacings
                    for (int c = 0; c < 8; ++c) {
                        int d2 = 0;
                        for (int a = 0; a < 3; ++a) {
                            const int d = ((c >> a) & 1) ? 3 - r[a] : r[a];
                            d2 += d * d;
                        }
                        w[c] = d2 == 3 ? g1 : d2 == 6 ? g2 : d2 == 9 ? g3 : g4;
                    }
                    // ExtendWfThree, residues (2, 1, 2): the corners (1,0,0)
                    // and (1,1,0) swapped (src/DVR.cpp:1615-1620)
                    if (kind == 0 && rx == 2 && ry == 1 && rz == 2) {
                        const double t = w[1];
                        w[1] = w[3];
                        w[3] = t;
                    }
                }
            }
}

struct Table {
    double w[27 * 8];
};

// fine[n^3] from coarse[nc^3]; mode 2 (double spacing) or 3 (triple, coarse
// point p at fine index 3 p + sh, points past [sh, sh + lim) zero).  With
// sites (n_sites > 0) only the points on the coarse lattice or farther than
// rtol from every site are written (Polarization 6); the others keep their
// exact value.
__global__ void dual_kernel(const double *__restrict__ coarse, int n, int nc, int mode, int sh,
                            int lim, Table T, const double *__restrict__ gx,
                            const double *__restrict__ sites, int n_sites, double rtol,
                            double *__restrict__ fine)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long ng = (long long)n * n * n;
    if (g >= ng) return;
    const int i = (int)(g % n), j = (int)((g / n) % n), k = (int)(g / ((long long)n * n));
    const int f[3] = {i, j, k};
    int p[3], r[3];
    bool out = false, lattice = true;
    for (int a = 0; a < 3; ++a) {
        const int q = f[a] - (mode == 3 ? sh : 0);
        if (mode == 3 && (q < 0 || q >= lim)) out = true;
        p[a] = q >= 0 ? q / mode : 0;
        r[a] = q >= 0 ? q - mode * p[a] : 0;
        lattice = lattice && r[a] == 0;
    }
    if (n_sites > 0 && !out && !lattice) {
        // Potential::MinDistCheck (src/Potential.cpp:96-111): from 50 bohr down
        const double x = gx[i], y = gx[j], z = gx[k];
        double mind = 50.0;
        for (int s = 0; s < n_sites; ++s) {
            const double dx = x - sites[3 * s], dy = y - sites[3 * s + 1], dz = z - sites[3 * s + 2];
            mind = fmin(mind, sqrt(dx * dx + dy * dy + dz * dz));
        }
        if (!(mind > rtol)) return;  // near a site: the exact value stays
    }
    double v = 0.0;
    if (!out) {
        if (mode == 2) {
            int m = 0;
            for (int c = 0; c < 8; ++c) {
                const int a = c & 1, b = (c >> 1) & 1, e = (c >> 2) & 1;
                if ((a && !r[0]) || (b && !r[1]) || (e && !r[2])) continue;
                v += coarse[(p[0] + a) + (long long)nc * ((p[1] + b) + (long long)nc * (p[2] + e))];
                ++m;
            }
            v /= (double)m;
        } else {
            const double *w = T.w + 8 * (r[0] + 3 * r[1] + 9 * r[2]);
            for (int c = 0; c < 8; ++c) {
                if (w[c] == 0.0) continue;
                const int a = c & 1, b = (c >> 1) & 1, e = (c >> 2) & 1;
                v += coarse[(p[0] + a) + (long long)nc * ((p[1] + b) + (long long)nc * (p[2] + e))] * w[c];
            }
        }
    }
    fine[g] = v;
}

}  // namespace

bool dual_coarse_grid(int n, double L, int &nc, double &Lc)
{
    if (n < 4) return false;
    const double h = L / (n + 1);
    if (n % 2 == 1) {
        nc = (n + 1) / 2;
        Lc = nc * 2 * h + 2 * h;
    } else {
        nc = (n / 2 + 1) / 3 * 2;
        Lc = nc * 3 * h + 3 * h;
    }
    return nc >= 2;
}

void dual_transfer(pisces_ctx *c, int nc, const double *d_coarse, int kind, const double *sites,
                   int n_sites, double rtol, double *d_fine)
{
    const int n = c->n[0];
    PC_REQUIRE(c->n[1] == n && c->n[2] == n && c->L[1] == c->L[0] && c->L[2] == c->L[0],
               PISCES_ERR_UNSUPPORTED, "the dual-grid transfer needs a cubic grid");
    int ncx;
    double Lc;
    PC_REQUIRE(dual_coarse_grid(n, c->L[0], ncx, Lc) && ncx == nc, PISCES_ERR_ARG,
               "the coarse grid is not the dual grid of this context's grid");
    const int mode = n % 2 ? 2 : 3;
    const int check = n - (nc * 3 / 2 - 1) * 2;
    Table T{};
    if (mode == 3) triple_table(kind, T.w);
    dbuf ds;
    if (n_sites > 0) {
        ds.alloc(sizeof(double) * 3 * n_sites);
        PC_HIP(hipMemcpyAsync(ds.p, sites, sizeof(double) * 3 * n_sites, hipMemcpyHostToDevice,
                              c->stream));
    }
    const long long ng = (long long)n * n * n;
    hipLaunchKernelGGL(dual_kernel, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, c->stream,
                       d_coarse, n, nc, mode, check / 2, n - check, T, c->dx[0].as<double>(),
                       ds.as<double>(), n_sites, rtol, d_fine);
    PC_HIP(hipGetLastError());
    c->sync();  // ds is released on return
}

}  // namespace pisces
