// rc.cpp -- reverse-communication Davidson with the calling convention of the
// reference's Davidson() (src/Davidson.h:2-15, src/Davidson.cpp:100-217): the
// caller loops, and on return 1 computes Z[inout[1]*ndim ..] = H B[inout[0]*ndim ..]
// for inout[2] vectors by its own means; return 0 ends the solve with the
// converged vectors in B[0..nroots) and the eigenvalues in evals.
//
// The solve itself is the FP64 reference-mirroring Davidson of solver_fp64.cpp
// with its block operations on the GPU of the context; it runs on a worker
// thread that hands each matrix-vector request to the caller (copying the
// vectors into the caller's B) and waits for the next call, which carries the
// products in Z.  State lives in the context (re-entrant across contexts), not
// in function statics as in the reference.
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "context.hpp"

using namespace pisces;

namespace {

struct RcState {
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    enum Phase { RUNNING, NEED_MV, MV_DONE, FINISHED } phase = RUNNING;
    bool abort = false;
    long long ndim = 0;
    int nroots = 0;
    int col0 = 0, nvec = 0;
    double *B = nullptr, *Z = nullptr;  // caller arrays (valid during the caller's call)
    std::vector<double> vecs, evals, diag;
    int result = 0;
    std::string error;
    ~RcState()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            abort = true;
        }
        cv.notify_all();
        if (worker.joinable()) worker.join();
    }
};

struct rc_abort {};

void worker_main(pisces_ctx *c, RcState *st, int maxsub, int maxmacro, int tol, int corrflag)
{
    Fp64Hooks hooks;
    hooks.host_diag = st->diag.data();
    hooks.matvec = [c, st](const double *X, double *Y, int col0, int nvec) {
        const long long N = st->ndim;
        // vectors to the caller's B at the subspace position of the request
        PC_HIP(hipMemcpyAsync(st->B + N * col0, X, sizeof(double) * N * nvec,
                              hipMemcpyDeviceToHost, c->stream));
        PC_HIP(hipStreamSynchronize(c->stream));
        std::unique_lock<std::mutex> lk(st->mu);
        st->col0 = col0;
        st->nvec = nvec;
        st->phase = RcState::NEED_MV;
        st->cv.notify_all();
        st->cv.wait(lk, [st] { return st->phase == RcState::MV_DONE || st->abort; });
        if (st->abort) throw rc_abort{};
        st->phase = RcState::RUNNING;
        PC_HIP(hipMemcpyAsync(Y, st->Z + N * col0, sizeof(double) * N * nvec,
                              hipMemcpyHostToDevice, c->stream));
        PC_HIP(hipStreamSynchronize(c->stream));
    };
    int rc = 0;
    try {
        rc = guarded_call([&] {
            return solve_fp64(c, st->nroots, maxsub, maxmacro, tol, corrflag, st->vecs.data(),
                              st->evals.data(), nullptr, &hooks);
        });
    } catch (const rc_abort &) {
        rc = PISCES_ERR_STATE;
    }
    std::lock_guard<std::mutex> lk(st->mu);
    st->result = rc;
    if (rc < 0) st->error = pisces_last_error();
    st->phase = RcState::FINISHED;
    st->cv.notify_all();
}

}  // namespace

extern "C" {

int pisces_davidson_work_size(int ndim, int maxsub, int nroots, int corrflag)
{
    // the state lives in the context; davwork is not used
    if (ndim <= 0 || maxsub <= nroots || nroots <= 0 || corrflag < 0 || corrflag > 2)
        return PISCES_ERR_ARG;
    return 1;
}

int pisces_davidson(pisces_ctx *c, int ndim, int maxsub, int nroots, int maxmacro, int tol,
                    int corrflag, int verbose, double *evals, int *nConv, double *B, double *Z,
                    const double *diag, double *davwork, int *inout)
{
    static_cast<void>(verbose);
    static_cast<void>(davwork);
    if (!c || !evals || !nConv || !B || !Z || !inout) return PISCES_ERR_ARG;
    if (ndim != c->ngp) {
        set_last_error("ndim must equal the context's grid size");
        return PISCES_ERR_ARG;
    }
    auto st = std::static_pointer_cast<RcState>(c->rc);
    if (!st) {
        // first call: start vectors in B[0..nroots), diag(H) from the caller
        if (!diag || nroots < 1 || nroots > 16 || maxsub <= nroots || maxsub > 64 ||
            corrflag < 0 || corrflag > 2)
            return PISCES_ERR_ARG;
        st = std::make_shared<RcState>();
        st->ndim = ndim;
        st->nroots = nroots;
        st->B = B;
        st->Z = Z;
        st->vecs.assign(B, B + (size_t)ndim * nroots);
        st->evals.assign(nroots, 0.0);
        st->diag.assign(diag, diag + ndim);
        c->rc = st;
        st->worker = std::thread(worker_main, c, st.get(), maxsub, maxmacro, tol, corrflag);
    } else {
        // return from the caller's matrix-vector products
        std::lock_guard<std::mutex> lk(st->mu);
        if (st->phase != RcState::NEED_MV) return PISCES_ERR_STATE;
        st->B = B;
        st->Z = Z;
        st->phase = RcState::MV_DONE;
        st->cv.notify_all();
    }
    std::unique_lock<std::mutex> lk(st->mu);
    st->cv.wait(lk, [&] {
        return st->phase == RcState::NEED_MV || st->phase == RcState::FINISHED;
    });
    if (st->phase == RcState::NEED_MV) {
        inout[0] = st->col0;
        inout[1] = st->col0;
        inout[2] = st->nvec;
        return 1;
    }
    lk.unlock();
    st->worker.join();
    const int rc = st->result;
    if (rc >= 0) {
        std::memcpy(B, st->vecs.data(), sizeof(double) * (size_t)ndim * st->nroots);
        // like the reference, the eigenvalues are set only when every root converged
        if (rc == st->nroots) std::memcpy(evals, st->evals.data(), sizeof(double) * st->nroots);
        *nConv = rc;
        inout[0] = rc;
    } else {
        set_last_error(st->error);
    }
    c->rc.reset();
    return rc < 0 ? rc : 0;
}

}  // extern "C"
