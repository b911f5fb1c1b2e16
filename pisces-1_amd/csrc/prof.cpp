// prof.cpp -- HIP-event timing of batched launch families (see prof.hpp).
#include "prof.hpp"

#include "bd.hpp"
#include "context.hpp"

#include <cstring>
#include <vector>

namespace pisces {

namespace {

struct Pair {
    hipEvent_t a = nullptr, b = nullptr;
    int kind = 0;
    double bytes = 0;
};

struct Totals {
    double ms = 0, bytes = 0;
    long long n = 0;
};

bool g_on = false;
std::vector<Pair> g_pending;   // recorded, not yet resolved
std::vector<hipEvent_t> g_free;
Totals g_tot[PK_COUNT];

hipEvent_t get_event()
{
    if (!g_free.empty()) {
        hipEvent_t e = g_free.back();
        g_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    PC_HIP(hipEventCreate(&e));
    return e;
}

void resolve()
{
    for (Pair &p : g_pending) {
        PC_HIP(hipEventSynchronize(p.b));
        float ms = 0.f;
        PC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        Totals &t = g_tot[p.kind];
        t.ms += ms;
        t.bytes += p.bytes;
        t.n += 1;
        g_free.push_back(p.a);
        g_free.push_back(p.b);
    }
    g_pending.clear();
}

const char *kNames[PK_COUNT] = {"hx_fp64", "hx_lo", "gram", "combine", "eig", "ritz_resid", "precond",
                                "factors", "potential", "dipole_scf", "pol_scp"};

}  // namespace

bool prof_enabled() { return g_on; }

ProfScope::ProfScope(int kind, hipStream_t s, double bytes)
{
    if (!g_on) return;
    s_ = s;
    Pair p;
    p.a = get_event();
    p.b = get_event();
    p.kind = kind;
    p.bytes = bytes;
    PC_HIP(hipEventRecord(p.a, s));
    g_pending.push_back(p);
    slot_ = (int)g_pending.size() - 1;
}

ProfScope::~ProfScope()
{
    if (slot_ < 0) return;
    static_cast<void>(hipEventRecord(g_pending[slot_].b, s_));
}

}  // namespace pisces

using namespace pisces;

extern "C" {

int pisces_prof_enable(int on)
{
    return guarded_call([&] {
        g_on = on != 0;
        return 0;
    });
}

int pisces_prof_reset(void)
{
    return guarded_call([&] {
        resolve();
        for (Totals &t : g_tot) t = Totals{};
        return 0;
    });
}

int pisces_prof_read(int kind, double *ms_total, long long *launches, double *bytes_total)
{
    if (kind < 0 || kind >= PK_COUNT) return PISCES_ERR_ARG;
    return guarded_call([&] {
        resolve();
        if (ms_total) *ms_total = g_tot[kind].ms;
        if (launches) *launches = g_tot[kind].n;
        if (bytes_total) *bytes_total = g_tot[kind].bytes;
        return 0;
    });
}

int pisces_tuning_gram(int *variant, int *rows)
{
    if (!variant || !rows) return PISCES_ERR_ARG;
    *variant = gram_config(rows);
    return 0;
}

const char *pisces_prof_name(int kind)
{
    return (kind >= 0 && kind < PK_COUNT) ? kNames[kind] : "";
}

}  // extern "C"
