// server.cpp -- continuous-batching eigensolver service on one GPU.
//
// Requests (cluster geometries with a grid and a number of states) queue on
// the host.  Every pisces_server_step() admits queued requests while slots and
// the HBM budget allow (cost = the instance's device footprint), builds their
// potentials on the GPU (rigid-water projection on the host, FP64 induced
// dipoles and the grid potential on the device), adds them to the batched
// block-Davidson engine, advances every live instance by one Davidson step and
// retires the finished ones.  Requests of a chain warm-start from the chain's
// previous converged Ritz vectors (kept on the device).  Grids of the same
// shape share their kinetic matrices and eigenbases.
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "blockops.hpp"
#include "context.hpp"
#include "engine.hpp"
#include "prof.hpp"

using namespace pisces;

namespace {

using clk = std::chrono::steady_clock;

struct GridShape {
    int n[3];
    double L[3];
    long long N;
    std::vector<double> x[3];
    dbuf dx[3], dT64[3];
    AxisStore ax[3];
    std::vector<double> piab;
};

struct InstRes {
    std::shared_ptr<GridShape> shape;
    dbuf V64, V32, D;
};

struct Pending {
    long long id;
    pisces_request req;
    std::vector<double> xyz, ion_xyz, ion_q;
    clk::time_point t_submit;
};

struct Live {
    long long id;
    int slot;
    size_t bytes;
    clk::time_point t_submit;
    long long chain;
    int nroots;
    long long N;
    int want_vectors;
    std::shared_ptr<InstRes> res;
};

}  // namespace

struct pisces_server {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t cap = 0, used = 0;
    int max_slots = 0, precision = 0, maxsub = 32, ptol = 7, block = 8;
    std::unique_ptr<BdEngine> eng;
    std::map<std::tuple<int, int, int, double, double, double>, std::shared_ptr<GridShape>> shapes;
    std::deque<Pending> queue;
    std::vector<Live> live;
    std::deque<pisces_result> done;
    std::map<long long, std::pair<dbuf, long long>> warm;  // chain -> vectors, N
    std::map<long long, std::vector<double>> vecs;        // finished request -> vectors
    std::deque<long long> vec_order;
    long long next_id = 0;
    std::vector<dbuf> pool;  // potential buffers of retired instances
    dbuf take(size_t bytes)
    {
        for (size_t i = 0; i < pool.size(); ++i)
            if (pool[i].bytes == bytes) {
                dbuf b = std::move(pool[i]);
                pool.erase(pool.begin() + (long)i);
                return b;
            }
        return dbuf(bytes);
    }
    // scratch of the potential build
    dbuf s_sites, s_mu, s_epc, s_tens, s_info, s_q, s_iq, s_id, s_rep, s_irep, s_ipps, s_ion, s_ionq;
    ~pisces_server()
    {
        if (stream) static_cast<void>(hipStreamSynchronize(stream));
        eng.reset();
        if (stream) static_cast<void>(hipStreamDestroy(stream));
    }
};

namespace {

void upload(dbuf &b, const void *h, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    b.ensure(bytes);
    PC_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
}

std::shared_ptr<GridShape> shape_of(pisces_server *sv, const pisces_request &r)
{
    auto key = std::make_tuple(r.n[0], r.n[1], r.n[2], r.L[0], r.L[1], r.L[2]);
    auto it = sv->shapes.find(key);
    if (it != sv->shapes.end()) return it->second;
    auto g = std::make_shared<GridShape>();
    g->N = 1;
    for (int a = 0; a < 3; ++a) {
        g->n[a] = r.n[a];
        g->L[a] = r.L[a];
        g->N *= r.n[a];
        g->x[a].resize(r.n[a]);
        std::vector<double> T((size_t)r.n[a] * r.n[a]);
        sine_dvr_full(r.n[a], 1.0, -0.5 * r.L[a], 0.5 * r.L[a], g->x[a].data(), T.data());
        upload(g->dx[a], g->x[a].data(), sizeof(double) * r.n[a], sv->stream);
        upload(g->dT64[a], T.data(), sizeof(double) * T.size(), sv->stream);
        g->ax[a].build(r.n[a], T.data(), sv->stream);  // waits for the uploads
    }
    // half-box particle-in-a-box start vector (DVR::ParticleInAnDBoxWf,
    // src/DVR.cpp:1725-1805)
    const double pi = 3.14159265358979;
    g->piab.resize(g->N);
    double x0[3], Lb[3];
    for (int a = 0; a < 3; ++a) {
        x0[a] = g->x[a][0];
        Lb[a] = g->x[a][g->n[a] - 1] - x0[a];
    }
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                const int ii[3] = {i, j, k};
                double w = 1.0;
                for (int a = 0; a < 3; ++a) {
                    const double q = g->x[a][ii[a]];
                    if (q < x0[a] / 2 || q > -x0[a] / 2) w = 0.0;
                    else w *= std::sin(pi / Lb[a] * 2.0 * (q - x0[a] / 2.0));
                }
                g->piab[i + (long long)g->n[0] * (j + (long long)g->n[1] * k)] = w;
            }
    sv->shapes[key] = g;
    return g;
}

size_t footprint(const pisces_server *sv, long long N)
{
    return sv->eng->bytes_per_instance(N) + (size_t)N * (8 + 4 + 8);
}

// potential of one request into res (V64, V32, D), on the server stream
void build_potential(pisces_server *sv, const Pending &p, InstRes &res)
{
    const GridShape &g = *res.shape;
    const int nw = p.req.nw;
    std::vector<double> sites((size_t)12 * nw);
    PC_REQUIRE(water_sites_host(nw, p.xyz.data(), sites.data()), PISCES_ERR_NUMERIC,
               "water projection produced non-finite coordinates");
    const int ns = 3 * nw;
    upload(sv->s_sites, sites.data(), sizeof(double) * sites.size(), sv->stream);
    sv->s_mu.ensure(sizeof(double) * 3 * ns);
    sv->s_epc.ensure(sizeof(double) * 3 * ns);
    sv->s_tens.ensure(sizeof(double) * 6 * (size_t)ns * ns);
    sv->s_info.ensure(sizeof(int) * 2);
    {
        ProfScope ps(PK_SCF, sv->stream, 0.0);
        water_scf_launch(nw, sv->s_sites.as<double>(), sv->s_mu.as<double>(),
                         sv->s_epc.as<double>(), sv->s_tens.as<double>(), sv->s_info.as<int>(),
                         sv->stream);
    }
    std::vector<double> q;
    std::vector<int> iq, id;
    water_lists_host(nw, q, iq, id);
    // default DPP potential: Potential 3, DampType 2, Polarization 1
    // (src/GetInput.cpp:43-61)
    const int flags[6] = {3, 2, 0, 1, 0, 0};
    const double pp[11] = {9.745, 2.536, 4.4837, 1.7724, 1.7724, 4.4837, 0.5 / std::sqrt(9.745),
                           0, 0, 0, 20.0};
    PotParams P{};
    std::vector<double> rep;
    std::vector<int> irep, ipps;
    PC_REQUIRE(potential_setup(flags, pp, 4 * nw, P, rep, irep, ipps), PISCES_ERR_ARG, "setup");
    P.ion_r = 2.9;  // Bloomfield r_Na, src/NaCl.cpp:95
    upload(sv->s_q, q.data(), sizeof(double) * q.size(), sv->stream);
    upload(sv->s_iq, iq.data(), sizeof(int) * iq.size(), sv->stream);
    upload(sv->s_id, id.data(), sizeof(int) * id.size(), sv->stream);
    upload(sv->s_rep, rep.data(), sizeof(double) * rep.size(), sv->stream);
    upload(sv->s_irep, irep.data(), sizeof(int) * irep.size(), sv->stream);
    upload(sv->s_ipps, ipps.data(), sizeof(int) * ipps.size(), sv->stream);
    if (p.req.nion > 0) {
        upload(sv->s_ion, p.ion_xyz.data(), sizeof(double) * p.ion_xyz.size(), sv->stream);
        upload(sv->s_ionq, p.ion_q.data(), sizeof(double) * p.ion_q.size(), sv->stream);
    }
    PotDev D{4 * nw, (int)q.size(), (int)id.size(), (int)irep.size(), (int)ipps.size(),
             p.req.nion, sv->s_sites.as<double>(), sv->s_q.as<double>(), sv->s_iq.as<int>(),
             sv->s_mu.as<double>(), sv->s_id.as<int>(), sv->s_rep.as<double>(),
             sv->s_irep.as<int>(), sv->s_ipps.as<int>(), sv->s_ion.as<double>(),
             sv->s_ionq.as<double>()};
    // V written in FP64 and FP32, then read once more with diag H written
    ProfScope ps(PK_POT, sv->stream, (double)g.N * (8 + 4 + 8 + 8));
    potential_grid_launch(P, D, g.n[0], g.n[1], g.n[2], g.dx[0].as<double>(), g.dx[1].as<double>(),
                          g.dx[2].as<double>(), res.V64.as<double>(), res.V32.as<float>(), sv->stream);
    diag_h(res.V64.as<double>(), g.dT64[0].as<double>(), g.dT64[1].as<double>(),
           g.dT64[2].as<double>(), g.n[0], g.n[1], g.n[2], res.D.as<double>(), sv->stream);
    // the host vectors above are read by the async uploads: wait before they go
    PC_HIP(hipStreamSynchronize(sv->stream));
}

void admit(pisces_server *sv)
{
    while (!sv->queue.empty() && (int)sv->live.size() < sv->max_slots) {
        Pending &p = sv->queue.front();
        long long N = (long long)p.req.n[0] * p.req.n[1] * p.req.n[2];
        const size_t fp = footprint(sv, N);
        if (!sv->live.empty() && sv->used + fp > sv->cap) break;  // wait for memory
        auto res = std::make_shared<InstRes>();
        res->shape = shape_of(sv, p.req);
        res->V64 = sv->take(sizeof(double) * N);
        res->V32 = sv->take(sizeof(float) * N);
        res->D = sv->take(sizeof(double) * N);
        build_potential(sv, p, *res);
        BdGrid g;
        for (int a = 0; a < 3; ++a) {
            g.n[a] = res->shape->n[a];
            g.ax[a] = res->shape->ax[a].m;
        }
        g.N = N;
        g.V64 = res->V64.as<double>();
        g.V32 = res->V32.as<float>();
        g.D = res->D.as<double>();
        int slot;
        auto w = sv->warm.find(p.req.chain);
        if (p.req.warm && p.req.chain >= 0 && w != sv->warm.end() && w->second.second == N) {
            slot = sv->eng->add(g, p.req.nroots, w->second.first.as<double>(), p.req.nroots,
                                (unsigned)p.id, nullptr, true);
        } else {
            slot = sv->eng->add(g, p.req.nroots, res->shape->piab.data(), 1, (unsigned)p.id,
                                nullptr, false);
        }
        sv->live.push_back(Live{p.id, slot, fp, p.t_submit, p.req.chain, p.req.nroots, N,
                                p.req.want_vectors, res});
        sv->used += fp;
        sv->queue.pop_front();
    }
}

void retire(pisces_server *sv)
{
    for (size_t i = 0; i < sv->live.size();) {
        Live &L = sv->live[i];
        if (!sv->eng->finished(L.slot)) {
            ++i;
            continue;
        }
        const BdResult &r = sv->eng->result(L.slot);
        pisces_result out{};
        out.id = L.id;
        out.status = r.status;
        out.nroots = L.nroots;
        for (int k = 0; k < L.nroots && k < 8; ++k) {
            out.evals[k] = k < (int)r.evals.size() ? r.evals[k] : 0.0;
            out.resid[k] = k < (int)r.resid.size() ? r.resid[k] : 0.0;
        }
        out.n_iter = r.n_iter;
        out.n_matvec = r.n_mv64 + r.n_mvlo;
        out.latency_s = std::chrono::duration<double>(clk::now() - L.t_submit).count();
        // warm-start snapshot of the chain, kept on the device
        if (L.chain >= 0 && r.status == L.nroots) {
            auto &w = sv->warm[L.chain];
            w.first.ensure(sizeof(double) * L.N * L.nroots);
            w.second = L.N;
            PC_HIP(hipMemcpyAsync(w.first.p, sv->eng->vectors_device(L.slot),
                                  sizeof(double) * L.N * L.nroots, hipMemcpyDeviceToDevice,
                                  sv->stream));
        }
        // keep the vectors of the last few finished requests that asked for them
        if (L.want_vectors) {
            std::vector<double> v((size_t)L.N * L.nroots);
            sv->eng->vectors(L.slot, v.data());
            sv->vecs[L.id] = std::move(v);
            sv->vec_order.push_back(L.id);
            while (sv->vec_order.size() > 16) {
                sv->vecs.erase(sv->vec_order.front());
                sv->vec_order.pop_front();
            }
        }
        sv->eng->release(L.slot);
        for (dbuf *b : {&L.res->V64, &L.res->V32, &L.res->D}) sv->pool.push_back(std::move(*b));
        sv->used -= L.bytes;
        sv->done.push_back(out);
        sv->live.erase(sv->live.begin() + (long)i);
    }
    // pooled buffers of retired instances count against the budget too
    size_t spool = 0;
    for (const dbuf &b : sv->pool) spool += b.bytes;
    const size_t pool = sv->eng->pool_bytes() + spool;
    if (sv->used + pool > sv->cap) {
        PC_HIP(hipStreamSynchronize(sv->stream));
        while (!sv->pool.empty() && sv->used + sv->eng->pool_bytes() + spool > sv->cap) {
            spool -= sv->pool.front().bytes;
            sv->pool.erase(sv->pool.begin());
        }
        const size_t room = sv->cap > sv->used + spool ? sv->cap - sv->used - spool : 0;
        sv->eng->trim_pool(room);
    }
}

}  // namespace

extern "C" {

int pisces_server_create(pisces_server **out, int device, size_t hbm_cap_bytes, int max_slots,
                         int precision, int maxsub, int ptol)
{
    if (!out || max_slots <= 0 || hbm_cap_bytes == 0) return PISCES_ERR_ARG;
    if (precision < PISCES_FP64 || precision > PISCES_BF16X2) return PISCES_ERR_ARG;
    *out = nullptr;
    auto *sv = new (std::nothrow) pisces_server();
    if (!sv) return PISCES_ERR_NOMEM;
    const int rc = guarded_call([&] {
        sv->device = device;
        PC_HIP(hipSetDevice(device));
        PC_HIP(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
        sv->cap = hbm_cap_bytes;
        sv->max_slots = max_slots;
        sv->precision = precision;
        sv->maxsub = maxsub;
        sv->ptol = ptol;
        sv->block = 8;
        sv->eng = std::make_unique<BdEngine>(device, sv->stream, precision, sv->block, maxsub, ptol,
                                             600, PISCES_PRECOND_KINETIC);
        return 0;
    });
    if (rc) {
        delete sv;
        return rc;
    }
    *out = sv;
    return PISCES_OK;
}

void pisces_server_destroy(pisces_server *sv) { delete sv; }

long long pisces_server_submit(pisces_server *sv, const pisces_request *req)
{
    if (!sv || !req || req->nw <= 0 || !req->xyz || req->nroots < 1 || req->nroots > 8 ||
        req->nion < 0 || (req->nion > 0 && (!req->ion_xyz || !req->ion_q)))
        return PISCES_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (req->n[a] < 4 || req->n[a] > 128 || !(req->L[a] > 0)) return PISCES_ERR_ARG;
    const long long N = (long long)req->n[0] * req->n[1] * req->n[2];
    if (footprint(sv, N) > sv->cap) return PISCES_ERR_NOMEM;  // can never be admitted
    Pending p;
    p.id = sv->next_id++;
    p.req = *req;
    p.xyz.assign(req->xyz, req->xyz + 9 * req->nw);
    if (req->nion > 0) {
        p.ion_xyz.assign(req->ion_xyz, req->ion_xyz + 3 * req->nion);
        p.ion_q.assign(req->ion_q, req->ion_q + req->nion);
    }
    p.req.xyz = nullptr;
    p.req.ion_xyz = nullptr;
    p.req.ion_q = nullptr;
    p.t_submit = clk::now();
    sv->queue.push_back(std::move(p));
    return sv->queue.back().id;
}

int pisces_server_step(pisces_server *sv)
{
    if (!sv) return PISCES_ERR_ARG;
    return guarded_call([&] {
        PC_HIP(hipSetDevice(sv->device));
        admit(sv);
        if (!sv->live.empty()) sv->eng->step();
        retire(sv);
        return (int)sv->live.size();
    });
}

int pisces_server_poll(pisces_server *sv, pisces_result *out, int max)
{
    if (!sv || (max > 0 && !out)) return PISCES_ERR_ARG;
    int n = 0;
    while (n < max && !sv->done.empty()) {
        out[n++] = sv->done.front();
        sv->done.pop_front();
    }
    return n;
}

size_t pisces_server_footprint(const pisces_server *sv, const pisces_request *req)
{
    if (!sv || !req) return 0;
    return footprint(sv, (long long)req->n[0] * req->n[1] * req->n[2]);
}

int pisces_server_vectors(pisces_server *sv, long long id, double *v)
{
    if (!sv || !v) return PISCES_ERR_ARG;
    auto it = sv->vecs.find(id);
    if (it == sv->vecs.end()) return PISCES_ERR_STATE;
    std::memcpy(v, it->second.data(), sizeof(double) * it->second.size());
    return PISCES_OK;
}

int pisces_server_pending(const pisces_server *sv)
{
    if (!sv) return PISCES_ERR_ARG;
    return (int)sv->queue.size();
}

size_t pisces_server_used_bytes(const pisces_server *sv) { return sv ? sv->used : 0; }

}  // extern "C"
