// server.cpp -- continuous-batching eigensolver service on one GPU.
//
// Requests (cluster geometries with a grid and a number of states) queue on
// the host.  Every pisces_server_step() admits queued requests while slots and
// the HBM budget allow (cost = the instance's device footprint), builds their
// potentials on the GPU (rigid-water projection on the host, FP64 induced
// dipoles and the grid potential on the device), adds them to the batched
// block-Davidson engine, advances every live instance by one Davidson step and
// retires the finished ones.  Instances are spread over lanes (three by
// default), each an engine on its own stream: the server queues the first half of a step on
// every lane before it waits on any, so one lane's latency-bound work (the
// projected eigensolves, small factor kernels, the host's per-instance
// decisions) runs beside the other lanes' memory-bound streams (on a batch of
// 32 (H2O)24 64^3 solves: 515 solves/s on two lanes, 538 on three or four).  Requests of a chain warm-start from the chain's
// previous converged Ritz vectors (kept on the device).  Grids of the same
// shape share their kinetic matrices and eigenbases.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "blockops.hpp"
#include <cstdlib>
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
    dbuf dpiab;  // the start vector on the device
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
    int e;  // engine (block size class)
    int l;  // lane
};

// Converged vectors of a chain's last request.  `ev` is recorded after every
// copy into or out of them, on the stream of that copy, and every such copy
// first waits for it: the lanes that write and read them have their own
// streams.
struct WarmVec {
    dbuf v;
    long long N = 0;
    hipEvent_t ev = nullptr;
};

}  // namespace

struct pisces_server {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t cap = 0, used = 0;
    int max_slots = 0, precision = 0, maxsub = 32, ptol = 7;
    // engines by block size: requests with up to block[0] states (4 unless
    // PISCES_SERVER_BLOCK says otherwise) and the rest (block 8), each class
    // on `lanes` lanes (3 unless PISCES_SERVER_LANES says 1 .. kMaxLanes);
    // engines are created on first use
    static constexpr int kMaxLanes = 4;
    int block[2] = {4, 8};
    int lanes = 3;
    hipStream_t lane_s[kMaxLanes] = {};
    std::unique_ptr<BdEngine> eng[2][kMaxLanes];  // [block class][lane]
    // PISCES_HOST_TIMING=1: host wall time of the parts of pisces_server_step()
    // (admission, of it the potential build, the two step halves, retirement),
    // printed to stderr when the server is destroyed
    bool htime = false;
    double ht[5] = {0, 0, 0, 0, 0};
    long long hsteps = 0;
    int engine_of(int nroots) const { return nroots <= block[0] ? 0 : 1; }
    BdEngine &engine(int e, int l)
    {
        if (!eng[e][l]) {
            const int ms = std::max(maxsub, 3 * block[e]);
            eng[e][l] = std::make_unique<BdEngine>(device, lane_s[l], precision, block[e], ms, ptol,
                                                   600, PISCES_PRECOND_KINETIC);
        }
        return *eng[e][l];
    }
    size_t engine_pool_bytes() const
    {
        size_t b = 0;
        for (const auto &c : eng)
            for (const auto &e : c)
                if (e) b += e->pool_bytes();
        return b;
    }
    std::map<std::tuple<int, int, int, double, double, double>, std::shared_ptr<GridShape>> shapes;
    std::deque<Pending> queue;
    std::vector<Live> live;
    std::deque<pisces_result> done;
    std::map<long long, WarmVec> warm;  // chain -> converged vectors
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
    // scratch of the potential build: uploaded tables (+ pinned staging), work
    dbuf s_up, s_wk;
    unsigned char *h_up = nullptr;
    size_t h_up_cap = 0;
    ~pisces_server()
    {
        if (htime)
            std::fprintf(stderr,
                         "[server] host ms over %lld steps: admit %.1f (potentials %.1f), "
                         "step_a %.1f, step_b %.1f, retire %.1f\n",
                         hsteps, ht[0], ht[1], ht[2], ht[3], ht[4]);
        if (stream) static_cast<void>(hipStreamSynchronize(stream));
        for (auto &c : eng)
            for (auto &e : c) e.reset();
        for (auto &kv : warm)
            if (kv.second.ev) static_cast<void>(hipEventDestroy(kv.second.ev));
        if (h_up) static_cast<void>(hipHostFree(h_up));
        for (hipStream_t ls : lane_s)
            if (ls) static_cast<void>(hipStreamDestroy(ls));
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
    // read by the lanes' streams after build_potentials() has waited for this one
    upload(g->dpiab, g->piab.data(), sizeof(double) * g->N, sv->stream);
    sv->shapes[key] = g;
    return g;
}

size_t footprint(pisces_server *sv, long long N, int nroots)
{
    return sv->engine(sv->engine_of(nroots), 0).bytes_per_instance(N) + (size_t)N * (8 + 4 + 8);
}

// Host-side tables of one admitted request's potential build.
struct Prep {
    Pending *p;
    std::shared_ptr<InstRes> res;
    std::vector<double> sites, q, rep;
    std::vector<int> iq, id, irep, ipps;
    size_t o_sites, o_q, o_iq, o_id, o_rep, o_irep, o_ipps, o_ion, o_ionq;  // upload offsets
    size_t o_mu, o_epc, o_tens, o_info;                                   // work offsets
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Potentials of a batch of admitted requests into their InstRes (V64, V32, D):
// rigid-water projection and tables on the host, one upload, one induced-dipole
// launch and one grid-potential launch for the whole batch (one workgroup per
// cluster / one grid row of workgroups per cluster), then diag H.
void build_potentials(pisces_server *sv, std::vector<Prep> &batch)
{
    const int nb = (int)batch.size();
    // default DPP potential: Potential 3, DampType 2, Polarization 1
    // (src/GetInput.cpp:43-61)
    const int flags[6] = {3, 2, 0, 1, 0, 0};
    const double pp[11] = {9.745, 2.536, 4.4837, 1.7724, 1.7724, 4.4837, 0.5 / std::sqrt(9.745),
                           0, 0, 0, 20.0};
    PotParams P{};
    size_t up = align256(sizeof(ScfTask) * nb) + align256(sizeof(PotTask) * nb), wk = 0;
    for (Prep &r : batch) {
        const int nw = r.p->req.nw;
        PC_REQUIRE(nw <= water_scf_max_waters() && 4 * nw <= potential_max_sites(),
                   PISCES_ERR_UNSUPPORTED, "too many waters");
        r.sites.resize((size_t)12 * nw);
        PC_REQUIRE(water_sites_host(nw, r.p->xyz.data(), r.sites.data()), PISCES_ERR_NUMERIC,
                   "water projection produced non-finite coordinates");
        water_lists_host(nw, r.q, r.iq, r.id);
        PC_REQUIRE(potential_setup(flags, pp, 4 * nw, P, r.rep, r.irep, r.ipps), PISCES_ERR_ARG,
                   "potential setup");
        const int ns = 3 * nw, nion = r.p->req.nion;
        auto put = [&](size_t bytes) { const size_t o = up; up += align256(bytes); return o; };
        r.o_sites = put(sizeof(double) * r.sites.size());
        r.o_q = put(sizeof(double) * r.q.size());
        r.o_iq = put(sizeof(int) * r.iq.size());
        r.o_id = put(sizeof(int) * r.id.size());
        r.o_rep = put(sizeof(double) * r.rep.size());
        r.o_irep = put(sizeof(int) * r.irep.size());
        r.o_ipps = put(sizeof(int) * r.ipps.size());
        r.o_ion = put(sizeof(double) * 3 * nion);
        r.o_ionq = put(sizeof(double) * nion);
        auto wput = [&](size_t bytes) { const size_t o = wk; wk += align256(bytes); return o; };
        r.o_mu = wput(sizeof(double) * 3 * ns);
        r.o_epc = wput(sizeof(double) * 3 * ns);
        r.o_tens = wput(sizeof(double) * 6 * (size_t)ns * ns);
        r.o_info = wput(sizeof(int) * 2);
    }
    P.ion_r = 2.9;  // Bloomfield r_Na, src/NaCl.cpp:95
    sv->s_up.ensure(up);
    sv->s_wk.ensure(wk);
    if (sv->h_up_cap < up) {
        if (sv->h_up) PC_HIP(hipHostFree(sv->h_up));
        sv->h_up = nullptr;
        sv->h_up_cap = up + up / 2;
        PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&sv->h_up), sv->h_up_cap, hipHostMallocDefault));
    }
    unsigned char *h = sv->h_up;
    unsigned char *d = sv->s_up.as<unsigned char>(), *w = sv->s_wk.as<unsigned char>();
    ScfTask *scf = reinterpret_cast<ScfTask *>(h);
    PotTask *pot = reinterpret_cast<PotTask *>(h + align256(sizeof(ScfTask) * nb));
    long long max_ng = 0;
    for (int i = 0; i < nb; ++i) {
        Prep &r = batch[i];
        auto cp = [&](size_t off, const void *src, size_t bytes) {
            if (bytes) std::memcpy(h + off, src, bytes);
        };
        const int nw = r.p->req.nw, nion = r.p->req.nion;
        cp(r.o_sites, r.sites.data(), sizeof(double) * r.sites.size());
        cp(r.o_q, r.q.data(), sizeof(double) * r.q.size());
        cp(r.o_iq, r.iq.data(), sizeof(int) * r.iq.size());
        cp(r.o_id, r.id.data(), sizeof(int) * r.id.size());
        cp(r.o_rep, r.rep.data(), sizeof(double) * r.rep.size());
        cp(r.o_irep, r.irep.data(), sizeof(int) * r.irep.size());
        cp(r.o_ipps, r.ipps.data(), sizeof(int) * r.ipps.size());
        cp(r.o_ion, r.p->ion_xyz.data(), sizeof(double) * 3 * nion);
        cp(r.o_ionq, r.p->ion_q.data(), sizeof(double) * nion);
        double *mu = reinterpret_cast<double *>(w + r.o_mu);
        scf[i] = ScfTask{nw, reinterpret_cast<const double *>(d + r.o_sites), mu,
                         reinterpret_cast<double *>(w + r.o_epc),
                         reinterpret_cast<double *>(w + r.o_tens),
                         reinterpret_cast<int *>(w + r.o_info)};
        const GridShape &g = *r.res->shape;
        PotTask t{};
        t.D = PotDev{4 * nw, (int)r.q.size(), (int)r.id.size(), (int)r.irep.size(),
                     (int)r.ipps.size(), nion, reinterpret_cast<const double *>(d + r.o_sites),
                     reinterpret_cast<const double *>(d + r.o_q),
                     reinterpret_cast<const int *>(d + r.o_iq), mu,
                     reinterpret_cast<const int *>(d + r.o_id),
                     reinterpret_cast<const double *>(d + r.o_rep),
                     reinterpret_cast<const int *>(d + r.o_irep),
                     reinterpret_cast<const int *>(d + r.o_ipps),
                     reinterpret_cast<const double *>(d + r.o_ion),
                     reinterpret_cast<const double *>(d + r.o_ionq)};
        t.nx = g.n[0];
        t.ny = g.n[1];
        t.nz = g.n[2];
        t.gx = g.dx[0].as<double>();
        t.gy = g.dx[1].as<double>();
        t.gz = g.dx[2].as<double>();
        t.v = r.res->V64.as<double>();
        t.v32 = r.res->V32.as<float>();
        pot[i] = t;
        max_ng = std::max(max_ng, g.N);
    }
    PC_HIP(hipMemcpyAsync(d, h, up, hipMemcpyHostToDevice, sv->stream));
    {
        ProfScope ps(PK_SCF, sv->stream, 0.0);
        water_scf_batched_launch(reinterpret_cast<const ScfTask *>(d), nb, sv->stream);
    }
    {
        double bytes = 0;  // V written in FP64 and FP32, read once more with diag H written
        for (Prep &r : batch) bytes += (double)r.res->shape->N * (8 + 4 + 8 + 8);
        ProfScope ps(PK_POT, sv->stream, bytes);
        potential_grid_batched_launch(
            P, reinterpret_cast<const PotTask *>(d + align256(sizeof(ScfTask) * nb)), nb, max_ng,
            sv->stream);
        for (Prep &r : batch) {
            const GridShape &g = *r.res->shape;
            diag_h(r.res->V64.as<double>(), g.dT64[0].as<double>(), g.dT64[1].as<double>(),
                   g.dT64[2].as<double>(), g.n[0], g.n[1], g.n[2], r.res->D.as<double>(),
                   sv->stream);
        }
    }
    // the pinned staging buffer is rewritten by the next batch: wait for the copy
    PC_HIP(hipStreamSynchronize(sv->stream));
}

void admit(pisces_server *sv)
{
    std::vector<Prep> batch;
    size_t used = sv->used;
    int live = (int)sv->live.size();
    for (auto it = sv->queue.begin(); it != sv->queue.end() && live < sv->max_slots; ++it) {
        const long long N = (long long)it->req.n[0] * it->req.n[1] * it->req.n[2];
        const size_t fp = footprint(sv, N, it->req.nroots);
        if (live > 0 && used + fp > sv->cap) break;  // wait for memory (first come, first served)
        Prep r;
        r.p = &*it;
        r.res = std::make_shared<InstRes>();
        r.res->shape = shape_of(sv, it->req);
        r.res->V64 = sv->take(sizeof(double) * N);
        r.res->V32 = sv->take(sizeof(float) * N);
        r.res->D = sv->take(sizeof(double) * N);
        batch.push_back(std::move(r));
        used += fp;
        ++live;
    }
    if (batch.empty()) return;
    const auto tp = clk::now();
    build_potentials(sv, batch);
    if (sv->htime) sv->ht[1] += std::chrono::duration<double, std::milli>(clk::now() - tp).count();
    for (Prep &r : batch) {
        Pending &p = *r.p;
        const long long N = (long long)p.req.n[0] * p.req.n[1] * p.req.n[2];
        const size_t fp = footprint(sv, N, p.req.nroots);
        const int e = sv->engine_of(p.req.nroots);
        // the lane of this class with the fewest live instances
        int cnt[pisces_server::kMaxLanes] = {};
        for (const Live &L : sv->live)
            if (L.e == e) ++cnt[L.l];
        int l = 0;
        for (int q = 1; q < sv->lanes; ++q)
            if (cnt[q] < cnt[l]) l = q;
        BdEngine &eng = sv->engine(e, l);
        auto res = r.res;
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
        if (p.req.warm && p.req.chain >= 0 && w != sv->warm.end() && w->second.N == N) {
            WarmVec &wv = w->second;
            PC_HIP(hipStreamWaitEvent(eng.stream(), wv.ev, 0));
            slot = eng.add(g, p.req.nroots, wv.v.as<double>(), p.req.nroots, (unsigned)p.id,
                           nullptr, true);
            PC_HIP(hipEventRecord(wv.ev, eng.stream()));
        } else {
            slot = eng.add(g, p.req.nroots, res->shape->dpiab.as<double>(), 1, (unsigned)p.id,
                           nullptr, true);
        }
        sv->live.push_back(Live{p.id, slot, fp, p.t_submit, p.req.chain, p.req.nroots, N,
                                p.req.want_vectors, res, e, l});
        sv->used += fp;
    }
    for (size_t i = 0; i < batch.size(); ++i) sv->queue.pop_front();
}

void retire(pisces_server *sv)
{
    for (size_t i = 0; i < sv->live.size();) {
        Live &L = sv->live[i];
        BdEngine &eng = sv->engine(L.e, L.l);
        if (!eng.finished(L.slot)) {
            ++i;
            continue;
        }
        const BdResult &r = eng.result(L.slot);
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
            WarmVec &w = sv->warm[L.chain];
            const size_t need = sizeof(double) * L.N * L.nroots;
            if (!w.ev) {
                PC_HIP(hipEventCreateWithFlags(&w.ev, hipEventDisableTiming));
            } else if (need > w.v.bytes) {
                PC_HIP(hipEventSynchronize(w.ev));  // the old buffer is freed
            } else {
                PC_HIP(hipStreamWaitEvent(eng.stream(), w.ev, 0));
            }
            w.v.ensure(need);
            w.N = L.N;
            PC_HIP(hipMemcpyAsync(w.v.p, eng.vectors_device(L.slot), need,
                                  hipMemcpyDeviceToDevice, eng.stream()));
            PC_HIP(hipEventRecord(w.ev, eng.stream()));
        }
        // keep the vectors of the last few finished requests that asked for them
        if (L.want_vectors) {
            std::vector<double> v((size_t)L.N * L.nroots);
            eng.vectors(L.slot, v.data());
            sv->vecs[L.id] = std::move(v);
            sv->vec_order.push_back(L.id);
            while (sv->vec_order.size() > 64) {
                sv->vecs.erase(sv->vec_order.front());
                sv->vec_order.pop_front();
            }
        }
        eng.release(L.slot);
        for (dbuf *b : {&L.res->V64, &L.res->V32, &L.res->D}) sv->pool.push_back(std::move(*b));
        sv->used -= L.bytes;
        sv->done.push_back(out);
        sv->live.erase(sv->live.begin() + (long)i);
    }
    // pooled buffers of retired instances count against the budget too
    size_t spool = 0;
    for (const dbuf &b : sv->pool) spool += b.bytes;
    const size_t pool = sv->engine_pool_bytes() + spool;
    if (sv->used + pool > sv->cap) {
        PC_HIP(hipStreamSynchronize(sv->stream));
        while (!sv->pool.empty() && sv->used + sv->engine_pool_bytes() + spool > sv->cap) {
            spool -= sv->pool.front().bytes;
            sv->pool.erase(sv->pool.begin());
        }
        size_t room = sv->cap > sv->used + spool ? sv->cap - sv->used - spool : 0;
        for (auto &c : sv->eng)
            for (auto &e : c)
                if (e) {
                    e->trim_pool(room);
                    room = room > e->pool_bytes() ? room - e->pool_bytes() : 0;
                }
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
        const char *htv = std::getenv("PISCES_HOST_TIMING");
        sv->htime = htv && htv[0] == '1';
        const char *ln = std::getenv("PISCES_SERVER_LANES");
        if (ln && std::atoi(ln) >= 1) sv->lanes = std::min(std::atoi(ln), pisces_server::kMaxLanes);
        for (int l = 0; l < sv->lanes; ++l)
            PC_HIP(hipStreamCreateWithFlags(&sv->lane_s[l], hipStreamNonBlocking));
        sv->cap = hbm_cap_bytes;
        sv->max_slots = max_slots;
        sv->precision = precision;
        sv->maxsub = maxsub;
        sv->ptol = ptol;
        // block 4 for up to 4 states (PISCES_SERVER_BLOCK overrides): on (H2O)24
        // 64^3 4-state solves it takes ~40 steps against ~32 with block 8, but
        // each step moves about half the bytes -- 284 vs 205 solves/s
        const char *be = std::getenv("PISCES_SERVER_BLOCK");
        if (be && std::atoi(be) > 0) sv->block[0] = std::min(8, std::atoi(be));
        PC_REQUIRE(3 * sv->block[0] <= maxsub, PISCES_ERR_ARG,
                   "maxsub must be at least 3x the block");
        sv->engine(0, 0);
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
    if (footprint(sv, N, req->nroots) > sv->cap) return PISCES_ERR_NOMEM;  // can never be admitted
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
        const auto t0 = clk::now();
        admit(sv);
        const auto t1 = clk::now();
        bool any[2][pisces_server::kMaxLanes] = {};
        for (const Live &L : sv->live) any[L.e][L.l] = true;
        for (int e = 0; e < 2; ++e)
            for (int l = 0; l < sv->lanes; ++l)
                if (any[e][l]) sv->engine(e, l).step_a();
        const auto t2 = clk::now();
        for (int e = 0; e < 2; ++e)
            for (int l = 0; l < sv->lanes; ++l)
                if (any[e][l]) sv->engine(e, l).step_b();
        const auto t3 = clk::now();
        retire(sv);
        if (sv->htime) {
            using ms = std::chrono::duration<double, std::milli>;
            sv->ht[0] += ms(t1 - t0).count();
            sv->ht[2] += ms(t2 - t1).count();
            sv->ht[3] += ms(t3 - t2).count();
            sv->ht[4] += ms(clk::now() - t3).count();
            ++sv->hsteps;
        }
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
    if (!sv || !req || req->nroots < 1 || req->nroots > 8) return 0;
    return footprint(const_cast<pisces_server *>(sv), (long long)req->n[0] * req->n[1] * req->n[2],
                     req->nroots);
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
