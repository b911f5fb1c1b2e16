// solve_block.cpp -- single-instance solves on the block-Davidson engine.
#include <chrono>
#include <cstring>

#include "context.hpp"
#include "engine.hpp"

namespace pisces {

BdGrid grid_of(pisces_ctx *c)
{
    BdGrid g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = c->n[a];
        g.ax[a] = c->ax[a].m;
    }
    g.N = c->ngp;
    g.V64 = c->dV64.as<double>();
    g.V32 = c->dV32.as<float>();
    c->wdiag.ensure(sizeof(double) * c->ngp);
    diag_h(c->dV64.as<double>(), c->dT64[0].as<double>(), c->dT64[1].as<double>(),
           c->dT64[2].as<double>(), c->n[0], c->n[1], c->n[2], c->wdiag.as<double>(), c->stream);
    g.D = c->wdiag.as<double>();
    return g;
}

int solve_block(pisces_ctx *c, int nroots, int block, int maxsub, int max_iter, int ptol,
                int precond, int precision, double *vecs, double *evals, pisces_stats *st)
{
    PC_REQUIRE(c->has_v, PISCES_ERR_STATE, "no potential set");
    auto t0 = std::chrono::steady_clock::now();
    BdEngine eng(c->device, c->stream, precision, block, maxsub, ptol, max_iter, precond);
    const BdGrid g = grid_of(c);
    const int slot = eng.add(g, nroots, vecs, nroots, 7u, nullptr);
    while (eng.step() > 0) {
    }
    const BdResult &r = eng.result(slot);
    eng.vectors(slot, vecs);
    if (r.status == nroots)
        for (int i = 0; i < nroots; ++i) evals[i] = r.evals[i];
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->nconv = r.status;
        st->n_matvec = r.n_mv64;
        st->n_matvec_lo = r.n_mvlo;
        st->n_iter = r.n_iter;
        st->n_macro = r.n_restart;
        st->n_orth_data = eng.data_rounds(slot);
        st->seconds = std::chrono::duration<double>(t1 - t0).count();
        for (int i = 0; i < nroots && i < 64; ++i) st->resid[i] = r.resid[i];
    }
    return r.status;
}

int solve_mixed(pisces_ctx *c, int nroots, int maxsub, int maxmacro, int ptol, int corrflag,
                int precision, double *vecs, double *evals, pisces_stats *st)
{
    PC_REQUIRE(corrflag == 1, PISCES_ERR_UNSUPPORTED,
               "the low-precision forms use the Davidson correction (corrflag 1)");
    const int block = nroots;
    PC_REQUIRE(maxsub >= 3 * block, PISCES_ERR_ARG, "maxsub must be at least 3*nroots");
    const int per_macro = (maxsub - block) / block + 1;
    return solve_block(c, nroots, block, maxsub, maxmacro * per_macro, ptol, 1, precision, vecs,
                       evals, st);
}

}  // namespace pisces

extern "C" int pisces_solve_block(pisces_ctx *c, int nroots, int block, int maxsub, int max_iter,
                                  int ptol, int precond, int precision, double *vecs,
                                  double *evals, pisces_stats *st)
{
    if (!c || !vecs || !evals) return PISCES_ERR_ARG;
    if (nroots < 1 || block < nroots || precision < PISCES_FP64 || precision > PISCES_BF16X2)
        return PISCES_ERR_ARG;
    if (precond != 1 && precond != 3) return PISCES_ERR_ARG;
    return pisces::guarded_call([&] {
        PC_HIP(hipSetDevice(c->device));
        return pisces::solve_block(c, nroots, block, maxsub, max_iter, ptol, precond, precision,
                                   vecs, evals, st);
    });
}
