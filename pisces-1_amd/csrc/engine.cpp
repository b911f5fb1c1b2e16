// engine.cpp -- host control of the batched block-Davidson engine.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "blockops.hpp"

namespace pisces {

namespace {
// per-instance small-array layout (doubles)
constexpr int oG = 0;        // projected matrix, ld kBdMaxM
constexpr int oY = 4096;     // eigenvectors, m x k (ld m)
constexpr int oTh = 5120;    // theta (k)
constexpr int oRn = 5136;    // ||R||^2 (k)
constexpr int oH = 5152;     // CGS coefficients ((k + mc) x nt)
constexpr int oWc = 6176;    // CholQR factor (k x k)
constexpr int oGt = 6432;    // Gram of the new block (k x k)
constexpr int oKeep = 6688;  // int keep mask of all columns (as ints)
constexpr int oKeepT = 6752; // int keep mask of the new block
constexpr int oGu = 6784;    // Gram of U at restart (k x k)
constexpr int oWu = 7040;    // restart factor (k x k)
constexpr int kSmall = 8192;
constexpr int kRegions = 24;
constexpr int kCpb = 4;      // Gram chunks per workgroup
}  // namespace

struct BdEngine::Inst {
    BdGrid g;
    int nroots = 0, k = 0, mcmax = 0;
    dbuf X, HX, U, HU, R, T, T2, C, HC, small;
    int mc = 0, nnew = 0, nt = 0;
    int src[kBdMaxK];
    bool restart = true, init = true, refresh = false, fresh = false, done = false;
    bool released = false;
    BdResult res;
    double theta[kBdMaxK], rn[kBdMaxK];
    std::shared_ptr<void> keepalive;
    double *sm() const { return small.as<double>(); }
    int *keep() const { return reinterpret_cast<int *>(small.as<double>() + oKeep); }
    int *keepT() const { return reinterpret_cast<int *>(small.as<double>() + oKeepT); }
};

BdEngine::BdEngine(int device, hipStream_t s, int precision, int block, int maxsub, int ptol,
                   int max_iter)
    : dev_(device), s_(s), prec_(precision), k_(block), maxsub_(maxsub), ptol_(ptol),
      max_iter_(max_iter)
{
    PC_REQUIRE(block >= 1 && block <= kBdMaxK, PISCES_ERR_ARG, "block size must be 1..16");
    PC_REQUIRE(maxsub >= 2 * block && maxsub <= kBdMaxM, PISCES_ERR_ARG,
               "need 2*block <= maxsub <= 64");
    PC_REQUIRE(precision >= PISCES_FP64 && precision <= PISCES_BF16X2, PISCES_ERR_ARG,
               "unknown precision");
}

BdEngine::~BdEngine()
{
    if (s_) static_cast<void>(hipStreamSynchronize(s_));
    if (htask_) static_cast<void>(hipHostFree(htask_));
    if (hsmall_) static_cast<void>(hipHostFree(hsmall_));
}

size_t BdEngine::bytes_per_instance(long long N) const
{
    const size_t e = (prec_ == PISCES_FP64) ? 8 : 4;
    const size_t mcmax = (size_t)(maxsub_ - k_);
    return (size_t)N * (7 * k_ * 8 + 2 * mcmax * e) + kSmall * 8;
}

int BdEngine::live() const
{
    int n = 0;
    for (auto &p : inst_)
        if (p && !p->done) ++n;
    return n;
}

void BdEngine::ensure_tasks(size_t per_region)
{
    if (per_region <= region_bytes_) return;
    PC_HIP(hipStreamSynchronize(s_));
    per_region = std::max(per_region, (size_t)4096);
    if (htask_) PC_HIP(hipHostFree(htask_));
    PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&htask_), per_region * kRegions, hipHostMallocDefault));
    dtask_.alloc(per_region * kRegions);
    region_bytes_ = per_region;
}

void BdEngine::grow_part(size_t bytes)
{
    if (bytes <= dpart_.bytes) return;
    PC_HIP(hipStreamSynchronize(s_));  // the old buffer may still be read
    dpart_.alloc(bytes + bytes / 4);
}

template <typename T>
T *BdEngine::stage(int region, int count, T **dptr)
{
    PC_REQUIRE((size_t)count * sizeof(T) <= region_bytes_, PISCES_ERR_ARG, "task region overflow");
    *dptr = reinterpret_cast<T *>(dtask_.as<unsigned char>() + (size_t)region * region_bytes_);
    return reinterpret_cast<T *>(htask_ + (size_t)region * region_bytes_);
}

namespace {
template <typename T>
const T *upload(BdEngine *, hipStream_t s, T *h, T *d, int count)
{
    if (count > 0) PC_HIP(hipMemcpyAsync(d, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    return d;
}
}  // namespace

int BdEngine::add(const BdGrid &g, int nroots, const double *start, int nstart, unsigned seed,
                  std::shared_ptr<void> keepalive)
{
    PC_REQUIRE(nroots >= 1 && nroots <= k_, PISCES_ERR_ARG, "nroots must be 1..block");
    PC_REQUIRE(nstart >= 0 && nstart <= k_, PISCES_ERR_ARG, "too many start vectors");
    auto in = std::make_unique<Inst>();
    in->g = g;
    in->nroots = nroots;
    in->k = k_;
    in->mcmax = maxsub_ - k_;
    in->keepalive = std::move(keepalive);
    const long long N = g.N;
    const size_t e = (prec_ == PISCES_FP64) ? 8 : 4;
    in->X.alloc(8 * N * k_);
    in->HX.alloc(8 * N * k_);
    in->U.alloc(8 * N * k_);
    in->HU.alloc(8 * N * k_);
    in->R.alloc(8 * N * k_);
    in->T.alloc(8 * N * k_);
    in->T2.alloc(8 * N * k_);
    in->C.alloc(e * N * in->mcmax);
    in->HC.alloc(e * N * in->mcmax);
    in->small.alloc(8 * kSmall);
    // start vectors in U (restart path orthonormalises them): given ones, then
    // seeded uniform [-1, 1] fill for the rest of the block
    if (nstart > 0)
        PC_HIP(hipMemcpyAsync(in->U.p, start, 8 * N * nstart, hipMemcpyHostToDevice, s_));
    if (nstart < k_)
        fill_uniform(in->U.as<double>() + N * nstart, N * (k_ - nstart), true, seed, s_);
    std::vector<int> ones(kBdMaxM + kBdMaxK, 1);
    PC_HIP(hipMemcpyAsync(in->keep(), ones.data(), sizeof(int) * (kBdMaxM + kBdMaxK),
                          hipMemcpyHostToDevice, s_));
    PC_HIP(hipStreamSynchronize(s_));
    in->res.nroots = nroots;
    for (size_t i = 0; i < inst_.size(); ++i)
        if (!inst_[i]) {
            inst_[i] = std::move(in);
            return (int)i;
        }
    inst_.push_back(std::move(in));
    return (int)inst_.size() - 1;
}

bool BdEngine::finished(int slot) const { return inst_.at(slot)->done; }
const BdResult &BdEngine::result(int slot) const { return inst_.at(slot)->res; }

void BdEngine::vectors(int slot, double *out)
{
    Inst *in = inst_.at(slot).get();
    PC_HIP(hipMemcpyAsync(out, in->U.p, 8 * in->g.N * in->nroots, hipMemcpyDeviceToHost, s_));
    PC_HIP(hipStreamSynchronize(s_));
}

void BdEngine::release(int slot)
{
    PC_HIP(hipStreamSynchronize(s_));
    inst_.at(slot).reset();
}

// X = orthonormalised U (two Cholesky-QR passes), HX = HU W (or marked for an
// FP64 recompute).  U holds the start vectors on the first call.
void BdEngine::run_restarts(std::vector<Inst *> &v)
{
    if (v.empty()) return;
    const int n = (int)v.size();
    for (int pass = 0; pass < 2; ++pass) {
        GramTask *dg;
        GramTask *hg = stage<GramTask>(pass * 3 + 0, n, &dg);
        CholTask *dc;
        CholTask *hc = stage<CholTask>(pass * 3 + 1, n, &dc);
        int max_nch = 0;
        long long part = 0;
        for (int i = 0; i < n; ++i) {
            Inst *in = v[i];
            const double *src = (pass == 0) ? in->U.as<double>() : in->X.as<double>();
            GramTask &t = hg[i];
            t = GramTask{src, in->g.N, k_, 0, src, in->g.N, k_, 0, in->g.N,
                         in->sm() + oGu, k_, part, 0, kCpb};
            t.nch = (int)((in->g.N + (long long)kBdChunk * kCpb - 1) / ((long long)kBdChunk * kCpb));
            part += (long long)k_ * k_ * t.nch;
            max_nch = std::max(max_nch, t.nch);
            hc[i] = CholTask{in->sm() + oGu, k_, 0, in->sm() + oWu, pass ? in->keep() : nullptr,
                             in->keep(), 1e-12};
        }
        grow_part(sizeof(double) * part);
        upload(this, s_, hg, dg, n);
        upload(this, s_, hc, dc, n);
        bd_gram(dg, n, max_nch, k_ * k_, 2 * k_, dpart_.as<double>(), s_);
        bd_cholinv(dc, n, s_);
        CombTask *dm;
        CombTask *hm = stage<CombTask>(pass * 3 + 2, 2 * n, &dm);
        int nm = 0;
        long long maxN = 0;
        for (int i = 0; i < n; ++i) {
            Inst *in = v[i];
            const double *src = (pass == 0) ? in->U.as<double>() : in->X.as<double>();
            hm[nm++] = CombTask{src, in->g.N, k_, 0, in->sm() + oWu, k_, nullptr, 0, 0, 0, nullptr,
                                0, in->X.p, in->g.N, 0, k_, 0, in->g.N};
            if (!in->init) {
                const double *hsrc = (pass == 0) ? in->HU.as<double>() : in->HX.as<double>();
                hm[nm++] = CombTask{hsrc, in->g.N, k_, 0, in->sm() + oWu, k_, nullptr, 0, 0, 0,
                                    nullptr, 0, in->HX.p, in->g.N, 0, k_, 0, in->g.N};
            }
            maxN = std::max(maxN, in->g.N);
        }
        upload(this, s_, hm, dm, nm);
        bd_combine(dm, nm, maxN, s_);
    }
    for (Inst *in : v) {
        in->refresh = in->init || prec_ != PISCES_FP64;
        in->init = false;
        in->restart = false;
        in->fresh = true;
        in->mc = 0;
        in->nnew = 0;
        in->res.n_restart++;
    }
}

int BdEngine::step()
{
    std::vector<Inst *> live;
    int maxn = 0;
    for (auto &p : inst_)
        if (p && !p->done) {
            live.push_back(p.get());
            maxn = std::max({maxn, p->g.n[0], p->g.n[1], p->g.n[2]});
        }
    if (live.empty()) return 0;
    const int n = (int)live.size();
    const size_t tmax = std::max({sizeof(GramTask), sizeof(CombTask), sizeof(PrecTask),
                                  sizeof(HxJob), sizeof(EigTask), sizeof(ResTask),
                                  sizeof(CholTask)});
    ensure_tasks(tmax * 2 * (n + 1) + 256);
    const double tol = std::pow(10.0, -ptol_);
    const bool lo = prec_ != PISCES_FP64;
    const int e32 = lo ? 1 : 0;

    // 1. restarts (and the first orthonormalisation of the start block)
    std::vector<Inst *> rs;
    for (Inst *in : live)
        if (in->restart) rs.push_back(in);
    run_restarts(rs);

    // 2. FP64 H X where X changed without an exact HX (start, low-precision restart)
    // 3. H C for the corrections appended in the previous step
    for (int phase = 0; phase < 2; ++phase) {
        HxJob *dj;
        HxJob *hj = stage<HxJob>(6 + phase, n, &dj);
        int nj = 0, pa = 0, pb = 0, mx = 0;
        for (Inst *in : live) {
            HxJob j{};
            j.nx = in->g.n[0]; j.ny = in->g.n[1]; j.nz = in->g.n[2];
            if (phase == 0) {
                if (!in->refresh) continue;
                j.nvec = k_;
                for (int a = 0; a < 3; ++a) j.T[a] = in->g.T64[a];
                j.V = in->g.V64;
                j.X = in->X.p; j.Y = in->HX.p;
                in->res.n_mv64 += k_;
                in->refresh = false;
            } else {
                if (in->nnew <= 0) continue;
                j.nvec = in->nnew;
                const size_t e = lo ? 4 : 8;
                for (int a = 0; a < 3; ++a) j.T[a] = lo ? in->g.T32[a] : in->g.T64[a];
                j.V = lo ? (const void *)in->g.V32 : (const void *)in->g.V64;
                j.X = in->C.as<unsigned char>() + e * in->g.N * (in->mc - in->nnew);
                j.Y = in->HC.as<unsigned char>() + e * in->g.N * (in->mc - in->nnew);
                if (lo) in->res.n_mvlo += in->nnew;
                else in->res.n_mv64 += in->nnew;
            }
            j.ld = in->g.N;
            hj[nj++] = j;
            pa = std::max(pa, j.nz * j.nvec);
            pb = std::max(pb, j.ny * j.nvec);
            mx = std::max({mx, j.nx, j.ny, j.nz});
        }
        if (nj) {
            upload(this, s_, hj, dj, nj);
            hx_launch(dj, nj, pa, pb, mx, phase == 0 ? PISCES_FP64 : prec_, s_);
        }
    }

    // 4. new columns of G
    {
        GramTask *dg;
        GramTask *hg = stage<GramTask>(8, 2 * n, &dg);
        int ng = 0, max_nch = 0, max_mamb = 0, max_cols = 0;
        long long part = 0;
        auto addg = [&](const void *A, int ma, int a32, const void *B, int mb, int b32, long long N,
                        double *out, int out_ld) {
            GramTask t{A, N, ma, a32, B, N, mb, b32, N, out, out_ld, part, 0, kCpb};
            t.nch = (int)((N + (long long)kBdChunk * kCpb - 1) / ((long long)kBdChunk * kCpb));
            part += (long long)ma * mb * t.nch;
            max_nch = std::max(max_nch, t.nch);
            max_mamb = std::max(max_mamb, ma * mb);
            max_cols = std::max(max_cols, ma + mb);
            hg[ng++] = t;
        };
        const size_t e = lo ? 4 : 8;
        for (Inst *in : live) {
            const long long N = in->g.N;
            if (in->fresh)
                addg(in->X.p, k_, 0, in->HX.p, k_, 0, N, in->sm() + oG, kBdMaxM);
            if (in->nnew > 0) {
                const int c0 = in->mc - in->nnew;
                const void *Cn = in->C.as<unsigned char>() + e * N * c0;
                const void *HCn = in->HC.as<unsigned char>() + e * N * c0;
                double *col = in->sm() + oG + (size_t)kBdMaxM * (k_ + c0);
                // rows of X: HX^T C_new (FP64 HX keeps these entries FP64-exact)
                addg(in->HX.p, k_, 0, Cn, in->nnew, e32, N, col, kBdMaxM);
                // rows of C (including the new ones): C^T HC_new
                addg(in->C.p, in->mc, e32, HCn, in->nnew, e32, N, col + k_, kBdMaxM);
            }
        }
        if (ng) {
            grow_part(sizeof(double) * part);
            upload(this, s_, hg, dg, ng);
            bd_gram(dg, ng, max_nch, max_mamb, max_cols, dpart_.as<double>(), s_);
        }
    }

    // 5. lowest k eigenpairs of G
    {
        EigTask *de;
        EigTask *he = stage<EigTask>(9, n, &de);
        for (int i = 0; i < n; ++i) {
            Inst *in = live[i];
            he[i] = EigTask{in->sm() + oG, kBdMaxM, k_ + in->mc, k_, in->sm() + oTh, in->sm() + oY,
                            in->keep(), 0, 0};
        }
        upload(this, s_, he, de, n);
        bd_eig(de, n, s_);
    }

    // 6. Ritz vectors U = [X|C] y and HU = [HX|HC] y
    {
        CombTask *dm;
        CombTask *hm = stage<CombTask>(10, 2 * n, &dm);
        long long maxN = 0;
        for (int i = 0; i < n; ++i) {
            Inst *in = live[i];
            const long long N = in->g.N;
            const int m = k_ + in->mc;
            const double *Y = in->sm() + oY;
            hm[2 * i] = CombTask{in->X.p, N, k_, 0, Y, m, in->C.p, N, in->mc, e32, Y + k_, m,
                                 in->U.p, N, 0, k_, 0, N};
            hm[2 * i + 1] = CombTask{in->HX.p, N, k_, 0, Y, m, in->HC.p, N, in->mc, e32, Y + k_, m,
                                     in->HU.p, N, 0, k_, 0, N};
            maxN = std::max(maxN, N);
        }
        upload(this, s_, hm, dm, 2 * n);
        bd_combine(dm, 2 * n, maxN, s_);
    }

    // 7. residuals R = HU - theta U and their norms
    if (dgather_.bytes < sizeof(double) * 2 * kBdMaxK * (size_t)n) {
        PC_HIP(hipStreamSynchronize(s_));
        dgather_.alloc(sizeof(double) * 2 * kBdMaxK * (size_t)std::max(n, 64));
    }
    {
        ResTask *dr;
        ResTask *hr = stage<ResTask>(11, n, &dr);
        int max_nch = 0;
        long long part = 0;
        for (int i = 0; i < n; ++i) {
            Inst *in = live[i];
            const long long N = in->g.N;
            ResTask t{in->U.as<double>(), in->HU.as<double>(), in->R.as<double>(), N, k_, 0,
                      in->sm() + oTh, N, part, 0, 0, in->sm() + oRn,
                      dgather_.as<double>() + (size_t)2 * kBdMaxK * i};
            t.nch = (int)((N + kBdResChunk - 1) / kBdResChunk);
            part += (long long)k_ * t.nch;
            max_nch = std::max(max_nch, t.nch);
            hr[i] = t;
        }
        grow_part(sizeof(double) * part);
        upload(this, s_, hr, dr, n);
        bd_resid(dr, n, max_nch, dpart_.as<double>(), s_);
    }
    // one transfer of theta and ||R||^2 for every live instance
    if (hsmall_cap_ < (size_t)n * 2 * kBdMaxK) {
        PC_HIP(hipStreamSynchronize(s_));
        if (hsmall_) PC_HIP(hipHostFree(hsmall_));
        hsmall_cap_ = (size_t)std::max(n, 64) * 2 * kBdMaxK;
        PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&hsmall_), sizeof(double) * hsmall_cap_,
                             hipHostMallocDefault));
    }
    PC_HIP(hipMemcpyAsync(hsmall_, dgather_.p, sizeof(double) * 2 * kBdMaxK * n,
                          hipMemcpyDeviceToHost, s_));
    PC_HIP(hipStreamSynchronize(s_));

    // 8. control
    std::vector<Inst *> corr;
    for (int i = 0; i < n; ++i) {
        Inst *in = live[i];
        const double *h = hsmall_ + (size_t)2 * kBdMaxK * i;
        for (int r = 0; r < k_; ++r) {
            in->theta[r] = h[r];
            in->rn[r] = std::sqrt(std::max(0.0, h[16 + r]));
        }
        in->res.n_iter++;
        bool wanted = true;
        for (int r = 0; r < in->nroots; ++r)
            if (!(in->rn[r] < tol)) wanted = false;
        const bool was_fresh = in->fresh;
        in->fresh = false;
        in->nnew = 0;
        in->nt = 0;
        if (wanted && (!lo || was_fresh)) {
            in->done = true;
            in->res.status = in->nroots;
            in->res.evals.assign(in->theta, in->theta + in->nroots);
            in->res.resid.assign(in->rn, in->rn + in->nroots);
            continue;
        }
        if (in->res.n_iter >= max_iter_) {
            int nc = 0;
            for (int r = 0; r < in->nroots; ++r) nc += (in->rn[r] < tol);
            in->done = true;
            in->res.status = nc;
            in->res.evals.assign(in->theta, in->theta + in->nroots);
            in->res.resid.assign(in->rn, in->rn + in->nroots);
            continue;
        }
        if (wanted) {  // low precision: confirm on an FP64-fresh HX
            in->restart = true;
            continue;
        }
        int cnt = 0;
        for (int r = 0; r < k_; ++r)
            if (!(in->rn[r] < tol)) in->src[cnt++] = r;
        if (cnt == 0 || in->mc + cnt > in->mcmax) {
            in->restart = true;
            continue;
        }
        in->nt = cnt;
        corr.push_back(in);
    }

    // 9-11. corrections: precondition, orthogonalise, store in C
    if (!corr.empty()) {
        const int nc = (int)corr.size();
        long long maxN = 0;
        {
            PrecTask *dp;
            PrecTask *hp = stage<PrecTask>(12, nc, &dp);
            for (int i = 0; i < nc; ++i) {
                Inst *in = corr[i];
                PrecTask t{};
                t.R = in->R.as<double>(); t.ld = in->g.N; t.theta = in->sm() + oTh; t.D = in->g.D;
                t.T = in->T.as<double>(); t.N = in->g.N; t.nt = in->nt;
                std::memcpy(t.src, in->src, sizeof(in->src));
                hp[i] = t;
                maxN = std::max(maxN, in->g.N);
            }
            upload(this, s_, hp, dp, nc);
            bd_precond(dp, nc, maxN, s_);
        }
        const size_t e = lo ? 4 : 8;
        for (int pass = 0; pass < 2; ++pass) {
            GramTask *dg;
            GramTask *hg = stage<GramTask>(13 + 2 * pass, 2 * nc, &dg);
            CombTask *dm;
            CombTask *hm = stage<CombTask>(14 + 2 * pass, nc, &dm);
            int ng = 0, max_nch = 0, max_mamb = 0, max_cols = 0;
            long long part = 0;
            for (int i = 0; i < nc; ++i) {
                Inst *in = corr[i];
                const long long N = in->g.N;
                double *h = in->sm() + oH;
                const int m = k_ + in->mc;
                auto mk = [&](const void *A, int ma, int a32, double *out) {
                    GramTask t{A, N, ma, a32, in->T.p, N, in->nt, 0, N, out, m, part, 0, kCpb};
                    t.nch = (int)((N + (long long)kBdChunk * kCpb - 1) / ((long long)kBdChunk * kCpb));
                    part += (long long)ma * in->nt * t.nch;
                    max_nch = std::max(max_nch, t.nch);
                    max_mamb = std::max(max_mamb, ma * in->nt);
                    max_cols = std::max(max_cols, ma + in->nt);
                    hg[ng++] = t;
                };
                mk(in->X.p, k_, 0, h);
                if (in->mc > 0) mk(in->C.p, in->mc, e32, h + k_);
                hm[i] = CombTask{in->X.p, N, k_, 0, h, m, in->C.p, N, in->mc, e32, h + k_, m,
                                 in->T.p, N, 0, in->nt, 1, N};
            }
            grow_part(sizeof(double) * part);
            upload(this, s_, hg, dg, ng);
            bd_gram(dg, ng, max_nch, max_mamb, max_cols, dpart_.as<double>(), s_);
            upload(this, s_, hm, dm, nc);
            bd_combine(dm, nc, maxN, s_);
        }
        // Cholesky QR of the new block, twice: T -> T2 -> C[mc .. mc+nt)
        for (int pass = 0; pass < 2; ++pass) {
            GramTask *dg;
            GramTask *hg = stage<GramTask>(17 + 3 * pass, nc, &dg);
            CholTask *dc;
            CholTask *hc = stage<CholTask>(18 + 3 * pass, nc, &dc);
            CombTask *dm;
            CombTask *hm = stage<CombTask>(19 + 3 * pass, nc, &dm);
            int max_nch = 0, max_mamb = 0, max_cols = 0;
            long long part = 0;
            for (int i = 0; i < nc; ++i) {
                Inst *in = corr[i];
                const long long N = in->g.N;
                const void *src = pass == 0 ? in->T.p : in->T2.p;
                GramTask t{src, N, in->nt, 0, src, N, in->nt, 0, N, in->sm() + oGt, in->nt, part, 0, kCpb};
                t.nch = (int)((N + (long long)kBdChunk * kCpb - 1) / ((long long)kBdChunk * kCpb));
                part += (long long)in->nt * in->nt * t.nch;
                max_nch = std::max(max_nch, t.nch);
                max_mamb = std::max(max_mamb, in->nt * in->nt);
                max_cols = std::max(max_cols, 2 * in->nt);
                hg[i] = t;
                hc[i] = CholTask{in->sm() + oGt, in->nt, 0, in->sm() + oWc,
                                 pass ? in->keepT() : nullptr, in->keepT(), 1e-10};
                void *dst = pass == 0 ? in->T2.p
                                      : (void *)(in->C.as<unsigned char>() + e * N * in->mc);
                hm[i] = CombTask{src, N, in->nt, 0, in->sm() + oWc, in->nt, nullptr, 0, 0, 0,
                                 nullptr, 0, dst, N, pass == 0 ? 0 : e32, in->nt, 0, N};
            }
            grow_part(sizeof(double) * part);
            upload(this, s_, hg, dg, nc);
            upload(this, s_, hc, dc, nc);
            upload(this, s_, hm, dm, nc);
            bd_gram(dg, nc, max_nch, max_mamb, max_cols, dpart_.as<double>(), s_);
            bd_cholinv(dc, nc, s_);
            bd_combine(dm, nc, maxN, s_);
        }
        for (Inst *in : corr) {
            PC_HIP(hipMemcpyAsync(in->keep() + k_ + in->mc, in->keepT(), sizeof(int) * in->nt,
                                  hipMemcpyDeviceToDevice, s_));
            in->mc += in->nt;
            in->nnew = in->nt;
        }
    }
    return this->live();
}

}  // namespace pisces
