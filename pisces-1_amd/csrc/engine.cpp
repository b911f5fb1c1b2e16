// engine.cpp -- host control of the batched block-Davidson engine.
//
// A step is two launch phases separated by one device-to-host transfer:
//   A: restarts (keeping the last search directions, "GD+k"), FP64 refresh of
//      H X, H C of the new corrections, new columns of G, eigenpairs, Ritz
//      vectors, residuals and their norms; a step that starts with a full
//      correction basis restarts at the end of its phase A instead (see
//      "late restarts" below);
//   B: corrections for the unconverged columns (precondition, orthogonalise,
//      append).
// All task descriptors of a phase are written to one pinned arena and sent
// with a single copy before the phase's first launch; the two phases use
// separate arenas so that a phase never overwrites descriptors still in flight.
#include "engine.hpp"
#include "prof.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "blockops.hpp"
#include "fdm.hpp"

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
constexpr int oKeep = 6688;  // int keep mask of all columns
constexpr int oKeepT = 6752; // int keep mask of the new block
constexpr int oGu = 6784;    // Gram of U at restart (k x k)
constexpr int oWu = 7040;    // restart factor (k x k)
constexpr int oQp = 7552;    // restart search directions in [X|C] coefficients (m x k)
constexpr int oI = 8576;     // identity (k x k)
constexpr int oF = 8832;     // orthogonalisation coefficients on [X|C] ((k + mc) x nt)
constexpr int oHG = 9856;    // [Xc|C|T]^T T of an orthogonalisation round ((k + mc + nt) x nt)
constexpr int oVf = oHG + (kBdMaxM + kBdMaxK) * kBdMaxK;  // eigenvectors of the last solve (m x m, ld kBdMaxM)
constexpr int oS = oVf + kBdMaxM * kBdMaxM;  // S = V^T V of the stored basis [Xc | C] (ld kBdMaxM)
constexpr int kSmall = oS + kBdMaxM * kBdMaxM;
static_assert(oKeepT * 8 + (2 * kBdMaxK + 3) * 4 <= oGu * 8, "new-block mask, flags and count fit before oGu");
// |W - I|, |F| bound under which the second orthogonalisation round's
// combination of FP32 columns is left out: it would move an element by about
// (m + nt) * 1e-9 of the column scale, well below FP32 rounding (6e-8)
constexpr double kSkipTol = 1e-9;

}  // namespace

struct BdEngine::Inst {
    BdGrid g;
    int nroots = 0, k = 0, mcmax = 0;
    // C holds [Xc | corrections]: Xc = X in the corrections' element type (kept
    // with X at every restart), so that V = [Xc | C] and a new block T written
    // behind the corrections are one contiguous column block
    dbuf X, HX, U, HU, R, T2, C, HC, small;
    size_t es = 8;  // element size of C, HC
    int mc = 0, nnew = 0, gnew = 0, nt = 0;
    int src[kBdMaxK];
    bool restart = true, init = true, refresh = false, fresh = false, done = false;
    bool have_prev = false;
    bool confirm = false;  // restart requested to confirm convergence on an FP64-fresh HX
    bool vec_x = false;    // the current Ritz block is X (a late restart), not U
    bool late = false;     // this step ends in a late restart
    bool u_valid = false;  // U holds the Ritz block of the last step
    bool late_rf = false;  // the late restart refreshes H X in FP64
    int eig_m = 0;         // columns of the basis the last eigensolve diagonalised (0: basis changed)
    bool s_alg = false;    // the last corrections' columns of S came from orth2 (no data path)
    BdResult res;
    double theta[kBdMaxK], rn[kBdMaxK];
    std::shared_ptr<void> keepalive;
    double *sm() const { return small.as<double>(); }
    void *xc() const { return C.p; }
    void *corr(int c) const { return C.as<unsigned char>() + es * g.N * (size_t)(k + c); }
    int *keep() const { return reinterpret_cast<int *>(small.as<double>() + oKeep); }
    int *keepT() const { return reinterpret_cast<int *>(small.as<double>() + oKeepT); }
    // flags of the second orthogonalisation round (after the new block's mask):
    // its combination is a no-op / both rounds came from one Gram and S
    int *skip2() const { return keepT() + 2 * kBdMaxK; }
    int *alg() const { return keepT() + 2 * kBdMaxK + 1; }
    // steps whose second orthogonalisation round was left to the stored data
    int *ndata() const { return keepT() + 2 * kBdMaxK + 2; }
    double *Sv() const { return small.as<double>() + oS; }
};

// ------------------------------------------------------------------ arena
struct BdEngine::Arena {
    unsigned char *h = nullptr;
    dbuf d;
    size_t cap = 0, used = 0;
    ~Arena()
    {
        if (h) static_cast<void>(hipHostFree(h));
    }
    void reserve(size_t bytes, hipStream_t s)
    {
        if (bytes <= cap) return;
        PC_HIP(hipStreamSynchronize(s));
        if (h) PC_HIP(hipHostFree(h));
        h = nullptr;
        cap = bytes + bytes / 2;
        PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&h), cap, hipHostMallocDefault));
        d.alloc(cap);
    }
    template <typename T> T *take(int count, T **dp)
    {
        used = (used + 63) & ~(size_t)63;
        PC_REQUIRE(used + sizeof(T) * (size_t)count <= cap, PISCES_ERR_ARG, "task arena overflow");
        T *hp = reinterpret_cast<T *>(h + used);
        *dp = reinterpret_cast<T *>(d.as<unsigned char>() + used);
        used += sizeof(T) * (size_t)count;
        return hp;
    }
    void flush(hipStream_t s)
    {
        if (used) PC_HIP(hipMemcpyAsync(d.p, h, used, hipMemcpyHostToDevice, s));
    }
};

BdEngine::BdEngine(int device, hipStream_t s, int precision, int block, int maxsub, int ptol,
                   int max_iter, int precond)
    : dev_(device), s_(s), prec_(precision), k_(block), maxsub_(maxsub), ptol_(ptol),
      max_iter_(max_iter), precond_(precond)
{
    PC_REQUIRE(precond == 1 || precond == 3, PISCES_ERR_ARG,
               "preconditioner must be 1 (diagonal) or 3 (kinetic)");
    PC_REQUIRE(block >= 1 && block <= kBdMaxK, PISCES_ERR_ARG, "block size must be 1..16");
    PC_REQUIRE(maxsub >= 3 * block && maxsub <= kBdMaxM, PISCES_ERR_ARG,
               "need 3*block <= maxsub <= 64");
    PC_REQUIRE(precision >= PISCES_FP64 && precision <= PISCES_BF16X2, PISCES_ERR_ARG,
               "unknown precision");
    arena_[0] = std::make_unique<Arena>();
    arena_[1] = std::make_unique<Arena>();
    {
        // device copy of every new instance's keep mask and identity
        static_assert(sizeof(int) * (kBdMaxM + kBdMaxK) <= kInitEye, "keep mask size");
        std::vector<unsigned char> h(kInitEye + sizeof(double) * kBdMaxK * kBdMaxK, 0);
        int *ones = reinterpret_cast<int *>(h.data());
        for (int i = 0; i < kBdMaxM + kBdMaxK; ++i) ones[i] = 1;
        double *eye = reinterpret_cast<double *>(h.data() + kInitEye);
        for (int i = 0; i < k_; ++i) eye[(size_t)i * k_ + i] = 1.0;
        dinit_.alloc(h.size());
        // on the engine's stream: the library never touches the null stream,
        // which would hold one of the device's few hardware queues for the
        // rest of the process and leave two engine lanes sharing another
        PC_HIP(hipMemcpyAsync(dinit_.p, h.data(), h.size(), hipMemcpyHostToDevice, s_));
        PC_HIP(hipStreamSynchronize(s_));
    }
    const char *tr = std::getenv("PISCES_TRACE");
    trace_ = tr && tr[0] == '1';
    // PISCES_REFRESH_BELOW=<r>: late restarts of the low precision forms skip
    // the FP64 refresh of H X while every wanted residual exceeds r (0: never skip)
    const char *rb = std::getenv("PISCES_REFRESH_BELOW");
    if (rb) refresh_below_ = std::atof(rb) > 0 ? std::atof(rb) : 1e300;
    // The second orthogonalisation round of the FP32-stored forms comes from S
    // (no pass over the data) unless the first round's squared amplification
    // |T|^2 / |T - V h0|^2 of a kept column exceeds amp2_max: the algebraic
    // round's error grows with that amplification (1.1-3.8 on the (H2O)24
    // bench, which then never takes the data path), and an instance above it
    // runs the round on its data (the launches are there and skip the others
    // on the device).  The default 1e3: a zero potential's six lowest states
    // (a split triplet, split-BF16 products) stalled with 1e5 and 1e6 and
    // converged in 17 steps, one of them on the data, with 1e3 and 1e4.
    // PISCES_AMP2_MAX=<a> sets the threshold (1: always on the data; 0: never,
    // whatever the amplification).
    const char *am = std::getenv("PISCES_AMP2_MAX");
    const double a2 = am ? std::atof(am) : 1e3;
    data_round2_ = a2 > 0;
    amp2_max_ = a2 > 0 ? a2 : 1e300;
    gram_load_config();
}

BdEngine::~BdEngine()
{
    if (s_) static_cast<void>(hipStreamSynchronize(s_));
    if (hsmall_) static_cast<void>(hipHostFree(hsmall_));
}

size_t BdEngine::bytes_per_instance(long long N) const
{
    const size_t e = (prec_ == PISCES_FP64) ? 8 : 4;
    const size_t mcmax = (size_t)(maxsub_ - k_);
    return (size_t)N * (6 * k_ * 8 + (2 * mcmax + k_) * e) + kSmall * 8;
}

int BdEngine::live() const
{
    int n = 0;
    for (auto &p : inst_)
        if (p && !p->done) ++n;
    return n;
}

void BdEngine::grow_part(size_t bytes)
{
    if (bytes <= dpart_.bytes) return;
    PC_HIP(hipStreamSynchronize(s_));  // the old buffer may still be read
    dpart_.alloc(bytes + bytes / 4);
}

dbuf BdEngine::take_buf(size_t bytes)
{
    for (size_t i = 0; i < pool_.size(); ++i)
        if (pool_[i].bytes == bytes) {
            dbuf b = std::move(pool_[i]);
            pool_.erase(pool_.begin() + (long)i);
            return b;
        }
    return dbuf(bytes);
}

int BdEngine::add(const BdGrid &g, int nroots, const double *start, int nstart, unsigned seed,
                  std::shared_ptr<void> keepalive, bool start_on_device)
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
    for (dbuf *b : {&in->X, &in->HX, &in->U, &in->HU, &in->R, &in->T2})
        *b = take_buf(8 * N * k_);
    in->es = e;
    in->C = take_buf(e * N * (in->mcmax + k_));
    in->HC = take_buf(e * N * in->mcmax);
    in->small = take_buf(8 * kSmall);
    // start vectors in U (the first restart orthonormalises them): the given
    // ones, then a seeded uniform [-1, 1] fill for the rest of the block
    if (nstart > 0)
        PC_HIP(hipMemcpyAsync(in->U.p, start, 8 * N * nstart,
                              start_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s_));
    if (nstart < k_)
        fill_uniform(in->U.as<double>() + N * nstart, N * (k_ - nstart), true, seed, s_);
    // keep mask (all ones) and the k x k identity from their device copy
    PC_HIP(hipMemcpyAsync(in->keep(), dinit_.p, sizeof(int) * (kBdMaxM + kBdMaxK),
                          hipMemcpyDeviceToDevice, s_));
    PC_HIP(hipMemcpyAsync(in->sm() + oI, dinit_.as<unsigned char>() + kInitEye,
                          sizeof(double) * k_ * k_, hipMemcpyDeviceToDevice, s_));
    PC_HIP(hipMemsetAsync(in->ndata(), 0, sizeof(int), s_));
    // the caller's host start vectors may go out of scope after the return
    if (nstart > 0 && !start_on_device) PC_HIP(hipStreamSynchronize(s_));
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

int BdEngine::data_rounds(int slot)
{
    int v = 0;
    PC_HIP(hipMemcpyAsync(&v, inst_.at(slot)->ndata(), sizeof(int), hipMemcpyDeviceToHost, s_));
    PC_HIP(hipStreamSynchronize(s_));
    return v;
}

void BdEngine::vectors(int slot, double *out)
{
    Inst *in = inst_.at(slot).get();
    const dbuf &src = in->vec_x ? in->X : in->U;
    PC_HIP(hipMemcpyAsync(out, src.p, 8 * in->g.N * in->nroots, hipMemcpyDeviceToHost, s_));
    PC_HIP(hipStreamSynchronize(s_));
}

size_t BdEngine::pool_bytes() const
{
    size_t t = 0;
    for (const dbuf &b : pool_) t += b.bytes;
    return t;
}

void BdEngine::trim_pool(size_t max_bytes)
{
    size_t t = pool_bytes();
    if (t <= max_bytes) return;
    PC_HIP(hipStreamSynchronize(s_));
    while (t > max_bytes && !pool_.empty()) {
        t -= pool_.front().bytes;
        pool_.erase(pool_.begin());
    }
}

const double *BdEngine::vectors_device(int slot) const { return inst_.at(slot)->U.as<double>(); }

void BdEngine::release(int slot)
{
    // buffers go to the pool (no hipFree: it would stall the device); they are
    // reused by later add() calls, in stream order
    Inst *in = inst_.at(slot).get();
    for (dbuf *b : {&in->X, &in->HX, &in->U, &in->HU, &in->R, &in->T2,
                    &in->C, &in->HC, &in->small})
        if (b->p) pool_.push_back(std::move(*b));
    inst_.at(slot).reset();
}

// ------------------------------------------------------------------ launches
namespace {


struct GramBatch {
    GramTask *h = nullptr, *d = nullptr;
    int n = 0, max_blocks = 0, max_blocks_narrow = 0, max_mamb = 0;
    long long part = 0;
    bool reduce = true;  // false: every task's consumer sums the partials itself
    // tasks are sized (row chunks, workgroups, partial offsets) in finish(),
    // for the Gram configuration in use
    void add(const void *A, int ma, int a32, const void *B, int mb, int b32, long long N,
             double *out, int out_ld)
    {
        PC_REQUIRE(ma >= 1 && ma <= kBdMaxM && mb >= 1 && mb <= kGramMaxB, PISCES_ERR_ARG,
                   "Gram task exceeds the kernel's column limits");
        h[n++] = GramTask{A, N, ma, a32, B, N, mb, b32, N, out, out_ld, 0, 0, 0, 0};
        max_mamb = std::max(max_mamb, ma * mb);
    }
    void finish()
    {
        part = 0;
        max_blocks = max_blocks_narrow = 0;
        for (int i = 0; i < n; ++i) {
            GramTask &t = h[i];
            t.part_off = part;
            t.kind = gram_narrow(t.mb) ? 1 : 0;
            if (t.kind) {
                t.nch = (int)((t.N + kGRowsNarrow - 1) / kGRowsNarrow);
                t.cpb = (t.ma + kGNarrowA - 1) / kGNarrowA;
                max_blocks_narrow = std::max(max_blocks_narrow, t.nch * t.cpb);
            } else {
                t.nch = gram_chunks(t.N);
                t.cpb = gram_blocks_per_chunk(t.ma, t.mb);
                max_blocks = std::max(max_blocks, t.nch * t.cpb);
            }
            part += (long long)t.ma * t.mb * t.nch;
        }
    }
    // bytes read once: B is not counted again where its columns lie inside A
    double bytes() const
    {
        double b = 0;
        for (int i = 0; i < n; ++i) {
            const GramTask &t = h[i];
            if (t.skip) continue;  // conservative: usually skipped on the device
            const size_t ea = t.a32 ? 4 : 8;
            b += (double)t.N * t.ma * ea;
            const unsigned char *a0 = static_cast<const unsigned char *>(t.A);
            const unsigned char *bp = static_cast<const unsigned char *>(t.B);
            const bool inside = bp >= a0 && bp < a0 + ea * (size_t)t.lda * t.ma;
            if (!inside) b += (double)t.N * t.mb * (t.b32 ? 4 : 8);
        }
        return b;
    }
    // nexp: tasks the device is expected to skip (second-round Grams after
    // orth2); a launch of only those is timed with the small factor kernels
    int nexp = 0;
    bool all_skip() const { return nexp == n; }
    void launch(double *dp, hipStream_t s) const
    {
        if (!n) return;
        ProfScope ps(all_skip() ? PK_CHOL : PK_GRAM, s, prof_enabled() ? bytes() : 0.0);
        bd_gram(d, n, max_blocks, max_blocks_narrow, max_mamb, dp, s, reduce);
    }
};

struct CombBatch {
    CombTask *h = nullptr, *d = nullptr;
    int n = 0, maxmb = 0;
    long long maxN = 0;
    void add(const CombTask &t)
    {
        h[n++] = t;
        maxN = std::max(maxN, t.N);
        maxmb = std::max(maxmb, t.mb);
    }
    double bytes() const
    {
        double b = 0;
        for (int i = 0; i < n; ++i) {
            const CombTask &t = h[i];
            if (t.skip) continue;  // conservative: usually skipped on the device
            b += (double)t.N * (t.ma1 * (t.a1_32 ? 4 : 8) + t.ma2 * (t.a2_32 ? 4 : 8) +
                                t.ma3 * (t.a3_32 ? 4 : 8));
            b += (double)t.N * t.mb * (t.y32 ? 4 : 8) * (t.mode ? 2 : 1);
        }
        return b;
    }
    int nexp = 0;  // as GramBatch::nexp (second-round combinations after orth2)
    bool all_skip() const { return nexp == n; }
    void launch(hipStream_t s) const
    {
        if (!n) return;
        ProfScope ps(all_skip() ? PK_CHOL : PK_COMB, s, prof_enabled() ? bytes() : 0.0);
        bd_combine(d, n, maxN, maxmb, s);
    }
};

struct HxBatch {
    HxJob *h = nullptr, *d = nullptr;
    int *plan_h = nullptr, *plan_d = nullptr;  // one-launch split-BF16 form
    int n = 0, pa = 0, pb = 0, mx = 0;
    void add(const HxJob &j)
    {
        h[n++] = j;
        pa = std::max(pa, j.nz * j.nvec);
        pb = std::max(pb, j.ny * j.nvec);
        mx = std::max({mx, j.nx, j.ny, j.nz});
    }
    // X read and Y written once per vector, V read once
    double bytes(size_t e) const
    {
        double b = 0;
        for (int i = 0; i < n; ++i) b += (double)(2 * h[i].nvec + 1) * h[i].nx * h[i].ny * h[i].nz * e;
        return b;
    }
    // The split-BF16 form runs both passes in one XCD-ordered launch; its plan
    // goes into the arena with the jobs (before the arena's upload).
    template <typename Ar> void plan(Ar &A, int precision)
    {
        if (!n || precision != PISCES_BF16X2 || !hx_fused_enabled()) return;
        plan_h = A.template take<int>(hx_fused_plan_ints(h, n), &plan_d);
        hx_fused_plan(h, n, plan_h);
    }
    void launch(int precision, hipStream_t s) const
    {
        if (!n) return;
        const size_t e = precision == PISCES_FP64 ? 8 : 4;
        ProfScope ps(precision == PISCES_FP64 ? PK_HX64 : PK_HXLO, s, prof_enabled() ? bytes(e) : 0.0);
        if (plan_d) hx_fused_launch(h, d, n, plan_h, plan_d, mx, s);
        else hx_launch(d, n, pa, pb, mx, precision, s);
    }
};

CombTask comb1(const void *A, long long N, int ma, int a32, const double *W, int ldw, void *Y,
               int y32, int mb, int mode)
{
    return CombTask{A, N, ma, a32, W, ldw, nullptr, 0, 0, 0, nullptr, 0, Y, N, y32, mb, mode, N};
}

}  // namespace

int BdEngine::step()
{
    step_a();
    return step_b();
}

void BdEngine::step_a()
{
    mid_ = false;
    act_.clear();
    rs_.clear();
    lt_.clear();
    for (auto &p : inst_)
        if (p && !p->done) act_.push_back(p.get());
    if (act_.empty()) return;
    const std::vector<Inst *> &act = act_;
    const int n = (int)act.size();
    const bool lo = prec_ != PISCES_FP64;
    const int e32 = lo ? 1 : 0;
    const size_t e = lo ? 4 : 8;
    // residuals and raw corrections in FP32 in the low-precision forms with the
    // kinetic preconditioner (only the split-BF16 preconditioner reads them)
    const int r32 = (lo && precond_ == 3) ? 1 : 0;
    const size_t tmax = std::max({sizeof(GramTask), sizeof(CombTask), sizeof(PrecTask),
                                  sizeof(HxJob), sizeof(EigTask), sizeof(RitzTask), sizeof(OrthTask),
                                  sizeof(CholTask), sizeof(FdmTask), sizeof(Orth2Task)});
    // at most ~44 descriptors per instance and phase, plus alignment slack and
    // the plan of a one-launch H*X (13 ints per vector) or preconditioner (20)
    const size_t per = tmax * 56 + 64 * 48 + sizeof(int) * 21 * kBdMaxK;
    arena_[0]->reserve(per * (size_t)(n + 1), s_);
    arena_[1]->reserve(per * (size_t)(n + 1), s_);
    if (dgather_.bytes < sizeof(double) * 2 * kBdMaxK * (size_t)n) {
        PC_HIP(hipStreamSynchronize(s_));
        dgather_.alloc(sizeof(double) * 2 * kBdMaxK * (size_t)std::max(n, 64));
    }
    if (hsmall_cap_ < (size_t)n * 2 * kBdMaxK) {
        PC_HIP(hipStreamSynchronize(s_));
        if (hsmall_) PC_HIP(hipHostFree(hsmall_));
        hsmall_ = nullptr;
        hsmall_cap_ = (size_t)std::max(n, 64) * 2 * kBdMaxK;
        PC_HIP(hipHostMalloc(reinterpret_cast<void **>(&hsmall_), sizeof(double) * hsmall_cap_,
                             hipHostMallocDefault));
    }

    // ============================ phase A: descriptors
    Arena &A = *arena_[0];
    A.used = 0;
    std::vector<Inst *> &rs = rs_, &lt = lt_;
    for (Inst *in : act) {
        in->late = false;
        if (in->restart) rs.push_back(in);
        // late restarts: with a full correction basis any unconverged column
        // overflows it, so this step ends in a restart whatever its residuals
        else if (in->mc >= in->mcmax) {
            in->late = true;
            lt.push_back(in);
            // low precision forms: refresh H X in FP64 once the wanted
            // residuals come within refresh_below_ of convergence; above it the
            // HU W recurrence (split-BF16 products, ~1e-5 relative) is far
            // more accurate than the residuals need
            double rmin = 1e300;
            for (int r = 0; r < in->nroots; ++r) rmin = std::min(rmin, in->rn[r]);
            in->late_rf = lo && !(rmin > refresh_below_);
        }
    }
    const int nr = (int)rs.size();
    // restart: X = orth(U) by Cholesky-QR, HX = HU W.  The Ritz block U is
    // orthonormal up to the rounding of the stored basis (FP32 corrections:
    // ~1e-7), so one pass leaves ~1e-14; the random start block of a new
    // instance takes a second pass.  HX is not formed where the FP64 refresh
    // below recomputes it.
    // U = X Y_X + C Y_C of the last step where its Ritz step did not store U
    CombBatch ur;
    ur.h = A.take<CombTask>(nr, &ur.d);
    for (Inst *in : rs)
        if (!in->init && !in->u_valid) {
            const long long N = in->g.N;
            const int m = k_ + in->mc;
            const double *Y = in->sm() + oY;
            ur.add(CombTask{in->X.p, N, k_, 0, Y, m, in->corr(0), N, in->mc, e32, Y + k_, m, in->U.p,
                            N, 0, k_, 0, N});
        }
    GramBatch rg[2];
    CholTask *rch[2] = {nullptr, nullptr}, *rcd[2] = {nullptr, nullptr};
    CombBatch rc[2];
    int nch2[2] = {0, 0};
    for (int pass = 0; pass < 2; ++pass) {
        rg[pass].h = A.take<GramTask>(nr, &rg[pass].d);
        rch[pass] = A.take<CholTask>(nr, &rcd[pass]);
        rc[pass].h = A.take<CombTask>(2 * nr, &rc[pass].d);
        for (int i = 0; i < nr; ++i) {
            Inst *in = rs[i];
            if (pass == 1 && !in->init) continue;
            const long long N = in->g.N;
            const double *src = pass == 0 ? in->U.as<double>() : in->X.as<double>();
            rg[pass].add(src, k_, 0, src, k_, 0, N, in->sm() + oGu, k_);
            rch[pass][nch2[pass]++] = CholTask{in->sm() + oGu, k_, 0, in->sm() + oWu,
                                               pass ? in->keep() : nullptr, in->keep(), 1e-12,
                                               nullptr};
            CombTask cx = comb1(src, N, k_, 0, in->sm() + oWu, k_, in->X.p, 0, k_, 0);
            cx.Y2 = in->xc();
            cx.y2_32 = e32;
            rc[pass].add(cx);
            const bool refresh = in->init || lo || in->confirm;
            if (!refresh) rc[pass].add(comb1(in->HU.as<double>(), N, k_, 0, in->sm() + oWu, k_,
                                             in->HX.p, 0, k_, 0));
        }
    }
    // GD+k: the search directions of the last Ritz step -- the part of the Ritz
    // block that came from the corrections, P = C Y_C, orthonormalised against
    // the Ritz block in the coefficients of the old basis -- open the new
    // correction basis C[0..k), with H P from [HX|HC] by the same coefficients.
    // Built from the old X, C before the restart overwrites X.
    std::vector<Inst *> gk;
    for (Inst *in : rs)
        if (!in->init && in->have_prev) gk.push_back(in);
    const int ngk = (int)gk.size();
    CoefTask *hcf = nullptr, *dcf = nullptr;
    hcf = A.take<CoefTask>(ngk, &dcf);
    CombBatch pc;
    pc.h = A.take<CombTask>(2 * ngk, &pc.d);
    for (int i = 0; i < ngk; ++i) {
        Inst *in = gk[i];
        const long long N = in->g.N;
        const int m = k_ + in->mc;  // basis of the last Ritz step
        double *Q = in->sm() + oQp;
        hcf[i] = CoefTask{in->sm() + oY, m, k_, Q, in->keep() + k_};
        // in place: C[0..k) = [X|C] Q, HC[0..k) = [HX|HC] Q
        pc.add(CombTask{in->X.p, N, k_, 0, Q, m, in->corr(0), N, in->mc, e32, Q + k_, m,
                        in->corr(0), N, e32, k_, 0, N});
        pc.add(CombTask{in->HX.p, N, k_, 0, Q, m, in->HC.p, N, in->mc, e32, Q + k_, m, in->HC.p,
                        N, e32, k_, 0, N});
    }
    for (Inst *in : rs) {
        in->refresh = in->init || lo || in->confirm;
        in->confirm = false;
        in->mc = 0;
        in->nnew = 0;
        in->gnew = 0;
        in->fresh = true;
        in->res.n_restart++;
    }
    for (Inst *in : gk) {
        in->mc = k_;
        in->gnew = k_;
    }
    // FP64 refresh of H X, and H C of last step's corrections
    HxBatch hr, hc;
    hr.h = A.take<HxJob>(n, &hr.d);
    hc.h = A.take<HxJob>(n, &hc.d);
    for (Inst *in : act) {
        HxJob j{};
        j.nx = in->g.n[0]; j.ny = in->g.n[1]; j.nz = in->g.n[2];
        j.ld = in->g.N;
        if (in->refresh) {
            j.nvec = k_;
            for (int a = 0; a < 3; ++a) j.ax[a] = in->g.ax[a];
            j.V = in->g.V64;
            j.X = in->X.p;
            j.Y = in->HX.p;
            hr.add(j);
            in->res.n_mv64 += k_;
            in->refresh = false;
        }
        if (in->nnew > 0) {
            j.nvec = in->nnew;
            for (int a = 0; a < 3; ++a) j.ax[a] = in->g.ax[a];
            j.V = lo ? (const void *)in->g.V32 : (const void *)in->g.V64;
            j.X = in->corr(in->mc - in->nnew);
            j.Y = in->HC.as<unsigned char>() + e * in->g.N * (in->mc - in->nnew);
            hc.add(j);
            if (lo) in->res.n_mvlo += in->nnew;
            else in->res.n_mv64 += in->nnew;
        }
    }
    hc.plan(A, prec_);
    // new columns of G
    GramBatch gg;
    int ngg = 0;
    for (Inst *in : act)
        ngg += 1 + 3 * ((in->gnew + kGramMaxB - 1) / kGramMaxB) +
               (k_ + in->mc + kGramMaxB - 1) / kGramMaxB;
    gg.h = A.take<GramTask>(ngg, &gg.d);
    for (Inst *in : act) {
        const long long N = in->g.N;
        if (in->fresh) gg.add(in->X.p, k_, 0, in->HX.p, k_, 0, N, in->sm() + oG, kBdMaxM);
        // new columns in pieces of at most kGramMaxB (after a late restart they
        // are the GD+k block and the new corrections together)
        for (int c0 = in->mc - in->gnew; c0 < in->mc; c0 += kGramMaxB) {
            const int nb = std::min(kGramMaxB, in->mc - c0);
            const void *Cn = in->corr(c0);
            const void *HCn = in->HC.as<unsigned char>() + e * N * c0;
            double *col = in->sm() + oG + (size_t)kBdMaxM * (k_ + c0);
            gg.add(in->HX.p, k_, 0, Cn, nb, e32, N, col, kBdMaxM);
            gg.add(in->corr(0), in->mc, e32, HCn, nb, e32, N, col + k_, kBdMaxM);
        }
        // low precision forms: S = V^T V of the stored basis V = [Xc | C], in
        // full after a restart, else its new columns -- which the last step's
        // orth2 already wrote where it ran (*alg = 1: the tasks do nothing)
        if (lo && (in->fresh || !in->s_alg)) {
            const int mv = k_ + in->mc;
            const int b0 = in->fresh ? 0 : k_ + in->mc - in->gnew;
            for (int c0 = b0; c0 < mv; c0 += kGramMaxB) {
                const int nb = std::min(kGramMaxB, mv - c0);
                const void *Vc = static_cast<const unsigned char *>(in->xc()) + e * N * (size_t)c0;
                gg.add(in->xc(), mv, e32, Vc, nb, e32, N, in->Sv() + (size_t)kBdMaxM * c0, kBdMaxM);
                if (!in->fresh) gg.h[gg.n - 1].skip = in->alg();
            }
        }
    }
    EigTask *he, *de;
    he = A.take<EigTask>(n, &de);
    RitzTask *hrz, *drz;
    hrz = A.take<RitzTask>(n, &drz);
    int res_nch = 0;
    long long res_part = 0;
    double ritz_bytes = 0;
    for (int i = 0; i < n; ++i) {
        Inst *in = act[i];
        const long long N = in->g.N;
        const int m = k_ + in->mc;
        // warm start from the last solve where the basis has only grown since
        const int mwarm = in->restart ? 0 : std::min(in->eig_m, m);
        he[i] = EigTask{in->sm() + oG, kBdMaxM, m, k_, in->sm() + oTh, in->sm() + oY, in->keep(),
                        mwarm, 0, in->sm() + oVf};
        in->eig_m = in->late ? 0 : m;
        const double *Y = in->sm() + oY;
        const bool late = in->late;
        // U is read by a restart (a late one at the end of this step, or a
        // regular one that may follow) and returned at the end; HU only by the
        // FP64 form's H X recurrence.  A regular restart without a stored U
        // rebuilds it from X, C and the eigenvectors.
        const int store = !lo ? (kRitzU | kRitzHU)
                              : ((late || in->fresh || in->res.n_iter + 1 >= max_iter_) ? kRitzU : 0) |
                                    ((late && !in->late_rf) ? kRitzHU : 0);
        in->u_valid = (store & kRitzU) != 0;
        RitzTask t{in->X.as<double>(), in->HX.as<double>(), in->corr(0), in->HC.p, k_, in->mc, e32,
                   r32, Y, in->sm() + oTh, in->U.as<double>(), in->HU.as<double>(), in->R.p, N,
                   res_part, (int)((N + ritz_rows(k_) - 1) / ritz_rows(k_)), store, in->sm() + oRn,
                   late ? nullptr : dgather_.as<double>() + (size_t)2 * kBdMaxK * i};
        hrz[i] = t;
        // X, HX, C, HC read; R and the stored ones of U, HU written
        ritz_bytes += (double)N * (16.0 * k_ + 2.0 * e * in->mc + (r32 ? 4 : 8) * k_ +
                                   (((store & kRitzU) ? 8.0 : 0.0) + ((store & kRitzHU) ? 8.0 : 0.0)) * k_);
        res_part += (long long)k_ * t.nch;
        res_nch = std::max(res_nch, t.nch);
    }

    // late restarts, after this step's Ritz block U, HU: the GD+k directions
    // from this step's eigenvectors and the current X, C (before they are
    // overwritten), X = U W by one Cholesky-QR pass, HX refreshed in FP64 (low
    // precision forms) or HU W, then G_XX = X^T H X and the fresh residuals
    // R = HX - X diag(G_XX) with their norms.  In exact arithmetic X spans the
    // Ritz block of the restarted basis [X | P] (R is orthogonal to P), so the
    // step goes on to corrections with FP64-fresh residuals instead of
    // spending a step on re-deriving the same Ritz pairs.
    const int nl = (int)lt.size();
    CoefTask *lcf = nullptr, *dlcf = nullptr;
    lcf = A.take<CoefTask>(nl, &dlcf);
    CholTask *lch = nullptr, *dlch = nullptr;
    lch = A.take<CholTask>(nl, &dlch);
    ResTask *lrs = nullptr, *dlrs = nullptr;
    lrs = A.take<ResTask>(nl, &dlrs);
    GramBatch lrg, lgx;
    lrg.h = A.take<GramTask>(nl, &lrg.d);
    lgx.h = A.take<GramTask>(nl * (1 + (2 * k_ + kGramMaxB - 1) / kGramMaxB), &lgx.d);
    CombBatch lpc, lrc;
    lpc.h = A.take<CombTask>(2 * nl, &lpc.d);
    lrc.h = A.take<CombTask>(2 * nl, &lrc.d);
    HxBatch lhr;
    lhr.h = A.take<HxJob>(nl, &lhr.d);
    long long lres_part = res_part;
    int lres_nch = 0;
    double lres_bytes = 0;
    for (int i = 0; i < nl; ++i) {
        Inst *in = lt[i];
        const long long N = in->g.N;
        const int m = k_ + in->mc;
        double *Q = in->sm() + oQp;
        lcf[i] = CoefTask{in->sm() + oY, m, k_, Q, in->keep() + k_};
        lpc.add(CombTask{in->X.p, N, k_, 0, Q, m, in->corr(0), N, in->mc, e32, Q + k_, m,
                         in->corr(0), N, e32, k_, 0, N});
        lpc.add(CombTask{in->HX.p, N, k_, 0, Q, m, in->HC.p, N, in->mc, e32, Q + k_, m, in->HC.p,
                         N, e32, k_, 0, N});
        lrg.add(in->U.p, k_, 0, in->U.p, k_, 0, N, in->sm() + oGu, k_);
        lch[i] = CholTask{in->sm() + oGu, k_, 0, in->sm() + oWu, nullptr, in->keep(), 1e-12, nullptr};
        CombTask cx = comb1(in->U.p, N, k_, 0, in->sm() + oWu, k_, in->X.p, 0, k_, 0);
        cx.Y2 = in->xc();
        cx.y2_32 = e32;
        lrc.add(cx);
        if (!in->late_rf) lrc.add(comb1(in->HU.p, N, k_, 0, in->sm() + oWu, k_, in->HX.p, 0, k_, 0));
        if (in->late_rf) {
            HxJob j{};
            j.nx = in->g.n[0]; j.ny = in->g.n[1]; j.nz = in->g.n[2];
            j.ld = N;
            j.nvec = k_;
            for (int a = 0; a < 3; ++a) j.ax[a] = in->g.ax[a];
            j.V = in->g.V64;
            j.X = in->X.p;
            j.Y = in->HX.p;
            lhr.add(j);
            in->res.n_mv64 += k_;
        }
        lgx.add(in->X.p, k_, 0, in->HX.p, k_, 0, N, in->sm() + oG, kBdMaxM);
        if (lo)  // S of the restarted basis [Xc | P]
            for (int c0 = 0; c0 < 2 * k_; c0 += kGramMaxB) {
                const int nb = std::min(kGramMaxB, 2 * k_ - c0);
                const void *Vc = static_cast<const unsigned char *>(in->xc()) + e * N * (size_t)c0;
                lgx.add(in->xc(), 2 * k_, e32, Vc, nb, e32, N, in->Sv() + (size_t)kBdMaxM * c0,
                        kBdMaxM);
            }
        const int nch = (int)((N + kBdResChunk - 1) / kBdResChunk);
        lrs[i] = ResTask{in->X.as<double>(), in->HX.as<double>(), in->R.p, N, k_, r32,
                         in->sm() + oG, N, lres_part, nch, kBdMaxM + 1, in->sm() + oRn,
                         dgather_.as<double>() + (size_t)2 * kBdMaxK * (size_t)(std::find(act.begin(), act.end(), in) - act.begin())};
        lres_part += (long long)k_ * nch;
        lres_nch = std::max(lres_nch, nch);
        lres_bytes += (double)N * k_ * (16.0 + (r32 ? 4 : 8));
    }

    // ============================ phase A: launches
    for (int pass = 0; pass < 2; ++pass) rg[pass].finish();
    gg.finish();
    lrg.finish();
    lgx.finish();
    const long long part = std::max({rg[0].part, rg[1].part, gg.part, res_part, lrg.part, lgx.part,
                                     lres_part});
    grow_part(sizeof(double) * (size_t)std::max(part, 1LL));
    A.flush(s_);
    double *dp = dpart_.as<double>();
    ur.launch(s_);
    if (ngk) {
        bd_coef(dcf, ngk, s_);
        pc.launch(s_);
    }
    for (int pass = 0; pass < 2; ++pass) {
        rg[pass].launch(dp, s_);
        if (nch2[pass]) {
            ProfScope ps(PK_CHOL, s_, 0.0);
            bd_cholinv(rcd[pass], nch2[pass], s_);
        }
        rc[pass].launch(s_);
    }
    hr.launch(PISCES_FP64, s_);
    hc.launch(prec_, s_);
    gg.launch(dp, s_);
    {
        ProfScope ps(PK_EIG, s_, 0.0);
        bd_eig(de, n, s_);
    }
    {
        ProfScope ps(PK_RESID, s_, ritz_bytes);
        bd_ritz(drz, n, res_nch, k_, dp, s_);
    }
    if (nl) {
        bd_coef(dlcf, nl, s_);
        lpc.launch(s_);
        lrg.launch(dp, s_);
        {
            ProfScope ps(PK_CHOL, s_, 0.0);
            bd_cholinv(dlch, nl, s_);
        }
        lrc.launch(s_);
        lhr.launch(PISCES_FP64, s_);
        lgx.launch(dp, s_);
        ProfScope ps(PK_RESID, s_, lres_bytes);
        bd_resid(dlrs, nl, lres_nch, dp, s_);
    }
    PC_HIP(hipMemcpyAsync(hsmall_, dgather_.p, sizeof(double) * 2 * kBdMaxK * n,
                          hipMemcpyDeviceToHost, s_));
    mid_ = true;
}

int BdEngine::step_b()
{
    if (!mid_) return this->live();
    mid_ = false;
    const std::vector<Inst *> &act = act_, &rs = rs_, &lt = lt_;
    const int n = (int)act.size();
    const double tol = std::pow(10.0, -ptol_);
    const bool lo = prec_ != PISCES_FP64;
    const int e32 = lo ? 1 : 0;
    const size_t e = lo ? 4 : 8;
    const int r32 = (lo && precond_ == 3) ? 1 : 0;
    PC_HIP(hipStreamSynchronize(s_));
    for (Inst *in : act) {
        in->have_prev = true;
        in->vec_x = false;
    }
    for (Inst *in : rs) {
        in->init = false;
        in->restart = false;
    }
    for (Inst *in : lt) {
        // FP64-fresh (convergence may be declared) unless H X came from the
        // low precision recurrence
        in->fresh = !lo || in->late_rf;
        in->vec_x = true;
        in->confirm = false;
        in->res.n_restart++;
    }

    // ============================ control
    std::vector<Inst *> corr;
    for (int i = 0; i < n; ++i) {
        Inst *in = act[i];
        const double *h = hsmall_ + (size_t)2 * kBdMaxK * i;
        for (int r = 0; r < k_; ++r) {
            in->theta[r] = h[r];
            in->rn[r] = std::sqrt(std::max(0.0, h[kBdMaxK + r]));
        }
        in->res.n_iter++;
        if (trace_ && i == 0) {
            std::fprintf(stderr, "[bd] it %d mc %d fresh %d |", in->res.n_iter, in->mc, (int)in->fresh);
            for (int r = 0; r < k_; ++r) std::fprintf(stderr, " %.9f:%.1e", in->theta[r], in->rn[r]);
            std::fprintf(stderr, "\n");
        }
        bool wanted = true;
        for (int r = 0; r < in->nroots; ++r)
            if (!(in->rn[r] < tol)) wanted = false;
        const bool was_fresh = in->fresh;
        in->fresh = false;
        in->nnew = 0;
        in->gnew = 0;
        in->nt = 0;
        if (in->late) {
            // the basis is now [X | P]: P's columns of G are formed next step
            in->mc = k_;
            in->gnew = k_;
        }
        if (wanted && was_fresh) {
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
        if (wanted) {
            // H X and the Ritz block are kept by recurrence (and, in the low
            // precision forms, carry the FP32/BF16 products): confirm on a
            // freshly computed FP64 H X before declaring convergence
            in->restart = true;
            in->confirm = true;
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
    if (corr.empty()) return this->live();

    // ============================ phase B: corrections
    // precondition the residual columns into T = C[mc .. mc + nt), then two
    // orthogonalisation rounds against V = [Xc | C[0 .. mc)] and within the
    // block, each one Gram task ([Xc|C|T]^T T: T is read from memory once),
    // one factor kernel and one in-place combination T = T W + V F
    Arena &B = *arena_[1];
    B.used = 0;
    const int nc = (int)corr.size();
    const int t32 = e32;
    const size_t et = e;
    PrecTask *hp = nullptr, *dpc = nullptr;
    FdmTask *hf = nullptr, *dfd = nullptr;
    int nfd = 0, fd_maxn = 0;
    long long maxN = 0;
    if (precond_ == 1) {
        hp = B.take<PrecTask>(nc, &dpc);
        for (int i = 0; i < nc; ++i) {
            Inst *in = corr[i];
            PrecTask t{};
            t.R = in->R.as<double>(); t.ld = in->g.N; t.theta = in->sm() + oTh; t.D = in->g.D;
            t.T = in->corr(in->mc); t.N = in->g.N; t.nt = in->nt; t.t32 = t32;
            std::memcpy(t.src, in->src, sizeof(in->src));
            hp[i] = t;
            maxN = std::max(maxN, in->g.N);
        }
    } else {
        int tot = 0;
        for (Inst *in : corr) tot += in->nt;
        hf = B.take<FdmTask>(tot, &dfd);
        for (Inst *in : corr) {
            maxN = std::max(maxN, in->g.N);
            fd_maxn = std::max({fd_maxn, in->g.n[0], in->g.n[1], in->g.n[2]});
            for (int t = 0; t < in->nt; ++t) {
                FdmTask f{};
                for (int a = 0; a < 3; ++a) f.ax[a] = in->g.ax[a];
                f.src = in->R.as<unsigned char>() + (r32 ? 4 : 8) * in->g.N * in->src[t];
                f.dst = in->corr(in->mc + t);
                f.src32 = r32;
                f.dst32 = t32;
                // BF16 intermediate of the split-BF16 form
                f.work = reinterpret_cast<uint16_t *>(in->T2.p) + in->g.N * t;
                // t = -|T - theta|^{-1} r: (T - theta)^{-1} for bound states
                // (theta below the kinetic spectrum); where theta lies inside
                // it, T - theta is indefinite and its corrections stall, and
                // for a constant potential it is exact, so that its correction
                // would be the Ritz vector itself
                f.shift = in->theta[in->src[t]];
                f.absden = 1;
                f.scale = -1.0;
                f.floor = 1e-3;
                hf[nfd++] = f;
            }
        }
    }
    // the split-BF16 preconditioner's three passes in one XCD-ordered launch
    int *fplan_h = nullptr, *fplan_d = nullptr;
    if (precond_ != 1 && lo && nfd && fdm_fused_enabled()) {
        fplan_h = B.take<int>(fdm_fused_plan_ints(nfd), &fplan_d);
        fdm_fused_plan(hf, nfd, fplan_h);
    }
    GramBatch og[2];
    CombBatch oc[2];
    OrthTask *hot[2], *dot[2];
    // the orth kernel sums the partials of a one-task Gram [Xc|C|T]^T T itself;
    // a pass reduces on its own only where an instance needs two Gram tasks
    std::vector<int> og_idx[2];
    bool og_red[2] = {false, false};
    // FP32 columns with a one-task Gram: both rounds from the first Gram and S
    // (orth2), the second round's launches skip them on the device
    Orth2Task *h2t = nullptr, *d2t = nullptr;
    h2t = B.take<Orth2Task>(nc, &d2t);
    int n2 = 0;
    std::vector<int> o2_inst;
    const int *one = dinit_.as<int>();  // a device 1 (the all-ones keep mask)
    for (int pass = 0; pass < 2; ++pass) {
        og_idx[pass].assign(nc, -1);
        og[pass].h = B.take<GramTask>(2 * nc, &og[pass].d);
        oc[pass].h = B.take<CombTask>(nc, &oc[pass].d);
        hot[pass] = B.take<OrthTask>(nc, &dot[pass]);
        for (int i = 0; i < nc; ++i) {
            Inst *in = corr[i];
            if (pass == 0) in->s_alg = !data_round2_ && t32 && k_ + in->mc + in->nt <= kBdMaxM;
            if (pass == 1 && !data_round2_ && t32 && k_ + in->mc + in->nt <= kBdMaxM) {
                // both rounds from S: nothing of the second round is launched
                hot[1][i] = OrthTask{};
                hot[1][i].noop = one;
                continue;
            }
            const long long N = in->g.N;
            const int m = k_ + in->mc, nt = in->nt;
            double *W = in->sm() + oWc, *F = in->sm() + oF;
            void *T = in->corr(in->mc);
            const double *h, *G;
            int hld, gld;
            if (m + nt <= kBdMaxM) {
                double *hg = in->sm() + oHG;
                og_idx[pass][i] = og[pass].n;
                og[pass].add(in->xc(), m + nt, e32, T, nt, e32, N, hg, m + nt);
                h = hg;
                G = hg + m;
                hld = gld = m + nt;
            } else {
                double *hv = in->sm() + oH, *gt = in->sm() + oGt;
                og_red[pass] = true;
                og[pass].add(in->xc(), m, e32, T, nt, e32, N, hv, m);
                og[pass].add(T, nt, e32, T, nt, e32, N, gt, nt);
                h = hv;
                G = gt;
                hld = m;
                gld = nt;
            }
            // FP32 columns: the second round's combination is left out where it
            // would change T by far less than the columns' rounding
            int *skip = (pass == 1 && t32) ? in->skip2() : nullptr;
            hot[pass][i] = OrthTask{h, G, m, nt, W, F, pass ? in->keepT() : nullptr, in->keepT(),
                                    pass ? in->keep() + k_ + in->mc : nullptr,
                                    pass ? 1e-10 : 1e-13, hld, gld, skip, kSkipTol};
            CombTask ct{T, N, nt, t32, W, nt, in->xc(), N, m, e32, F, m, T, N, t32, nt, 0, N};
            ct.skip = skip;
            oc[pass].add(ct);
            if (t32 && m + nt <= kBdMaxM) {
                if (pass == 0) {
                    hot[0][i].noop = one;
                    h2t[n2++] = Orth2Task{nullptr, 0, 0, m, in->Sv(), kBdMaxM, nt, W, F, in->keepT(),
                                          in->keep() + k_ + in->mc, 1e-13, 1e-10, amp2_max_, in->alg(),
                                          in->skip2(), in->Sv(), in->ndata()};
                    o2_inst.push_back(i);
                } else {
                    og[1].h[og[1].n - 1].skip = in->alg();
                    hot[1][i].noop = in->alg();
                    ++og[1].nexp;
                    ++oc[1].nexp;
                }
            } else if (lo && pass == 0) {
                // the data path: next step's Gram forms S's new columns
                hot[0][i].clear = in->alg();
            }
        }
    }
    long long partB = 1;
    for (int pass = 0; pass < 2; ++pass) {
        og[pass].finish();
        partB = std::max(partB, og[pass].part);
    }
    grow_part(sizeof(double) * (size_t)partB);
    double *dp = dpart_.as<double>();
    for (int pass = 0; pass < 2; ++pass) {
        og[pass].reduce = og_red[pass];
        for (int i = 0; i < nc; ++i) {
            const int q = og_idx[pass][i];
            if (q < 0) continue;
            hot[pass][i].part = dp;
            hot[pass][i].part_off = og[pass].h[q].part_off;
            hot[pass][i].nch = og[pass].h[q].nch;
        }
    }
    for (int q = 0; q < n2; ++q) {
        const GramTask &gt = og[0].h[og_idx[0][o2_inst[q]]];
        h2t[q].part = dp;
        h2t[q].part_off = gt.part_off;
        h2t[q].nch = gt.nch;
    }
    B.flush(s_);
    {
        // R columns read once, corrections written once (+ diag H for the diagonal form)
        double pb = 0;
        for (Inst *in : corr)
            pb += (double)in->g.N * in->nt * ((r32 ? 4 : 8) + et) + (precond_ == 1 ? 8.0 * in->g.N : 0.0);
        ProfScope ps(PK_PREC, s_, pb);
        if (precond_ == 1) bd_precond(dpc, nc, maxN, s_);
        else if (fplan_d) fdm_fused_launch(dfd, fplan_d, fplan_h, fd_maxn, s_);
        else fdm_launch(dfd, nfd, fd_maxn, lo ? PISCES_BF16X2 : PISCES_FP64, s_);
    }
    for (int pass = 0; pass < 2; ++pass) {
        og[pass].launch(dp, s_);
        {
            ProfScope ps(PK_CHOL, s_, 0.0);
            // the classic factor tasks of orth2 instances are no-ops
            if (n2 < nc || (pass == 1 && data_round2_)) bd_orth(dot[pass], nc, s_);
            if (pass == 0) bd_orth2(d2t, n2, s_);
        }
        oc[pass].launch(s_);
    }
    for (Inst *in : corr) {
        in->mc += in->nt;
        in->nnew = in->nt;
        in->gnew += in->nt;
    }
    return this->live();
}

}  // namespace pisces
