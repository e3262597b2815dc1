// ose_generic.hip -- launch plans of the runtime-length pass kernels
// (ose_generic.h) for FFT extents without compiled register kernels.
#include <map>
#include <memory>
#include <mutex>

#include "ose_entry.h"
#include "ose_generic.h"

namespace ose {

namespace {

constexpr size_t kLdsMax = 160 * 1024;
constexpr int kThreadsMax = 512;   // the kernels' launch bound (at most 256 VGPRs)

// Stages: the powers of two as radix 16 (then 8, 4, 2), then 3, 5, 7, 11, 13
// one factor each, then every larger prime factor as a direct stage.
bool plan_radices(int n, GenPlan& g)
{
    g.nf = 0;
    auto push = [&g](int r) {
        if (g.nf >= GEN_MAXF) return false;
        g.r[g.nf++] = (short)r;
        return true;
    };
    for (int r : {16, 8, 4, 2, 3, 5, 7, 11, 13})
        while (n % r == 0) {
            if (!push(r)) return false;
            n /= r;
        }
    for (int p = 17; n > 1; p += 2)
        while (n % p == 0) {
            if (!push(p)) return false;
            n /= p;
        }
    return true;
}

// values each thread holds in a stage of radix r: r per butterfly, at most
// GEN_MAXV; a direct stage's groups of GEN_GP output pairs, at most GEN_MAXV
int threads_per_pencil(const GenPlan& g)
{
    int tp = 1;
    for (int f = 0; f < g.nf; ++f) {
        const int r = g.r[f], m = g.N / r;
        if (r > 16) {
            const int groups = m * (((r - 1) / 2 + GEN_GP) / GEN_GP), per = GEN_MAXV / (2 * GEN_GP);
            tp = std::max(tp, (groups + per - 1) / per);
        } else {
            const int nb = GEN_MAXV / r;
            tp = std::max(tp, (m + nb - 1) / nb);
        }
    }
    return tp;
}

struct GenRegistry {
    std::mutex mu;
    std::map<int, std::unique_ptr<FftEntry>> entries;
    std::map<int, std::unique_ptr<GenPlan[]>> plans;   // [0] strided, [1] x pass
};

GenRegistry& gen_registry()
{
    static GenRegistry r;
    return r;
}

}  // namespace

const FftEntry* generic_entry(int N)
{
    if (N < 2 || N > GEN_NMAX) return nullptr;
    GenRegistry& reg = gen_registry();
    std::lock_guard<std::mutex> lk(reg.mu);
    auto it = reg.entries.find(N);
    if (it != reg.entries.end()) return it->second.get();
    GenPlan g;
    g.N = N;
    if (!plan_radices(N, g)) return nullptr;
    g.TP = threads_per_pencil(g);
    if (g.TP > kThreadsMax) return nullptr;
    // LDS: the root table (N entries), then the pencils or lines
    // strided: the widest T (x columns, rows of T + 1 elements) that fits
    int T = 0;
    for (int t : {16, 8, 4, 2, 1})
        if (t * g.TP <= kThreadsMax && sizeof(float2) * (size_t)N * (t + 2) <= kLdsMax) {
            T = t;
            break;
        }
    // x pass: up to 8 lines per workgroup
    int P = 0;
    for (int p : {8, 4, 2, 1})
        if (p * g.TP <= kThreadsMax && sizeof(float2) * (size_t)(p + 1) * N <= kLdsMax) {
            P = p;
            break;
        }
    if (!T || !P) return nullptr;
    std::unique_ptr<GenPlan[]> pl(new GenPlan[2]);
    pl[0] = g;
    pl[0].W = T;
    pl[1] = g;
    pl[1].W = P;
    auto e = std::make_unique<FftEntry>();
    e->N = N;
    e->N1 = e->N2 = 0;
    e->ns = 1;
    StridedKern k;
    k.T = T;
    k.split = 0;
    k.threads = T * g.TP;
    k.lds = sizeof(float2) * (size_t)N * (T + 2);
    k.gp = &pl[0];
    k.fn = (const void*)&k_gen_strided<-1>;
    e->yf[0] = k;
    k.fn = (const void*)&k_gen_strided<1>;
    e->yi[0] = k;
    k.fn = (const void*)&k_gen_zsym;
    e->zs[0] = k;
    e->nx = 1;
    XKern x;
    x.P = P;
    x.split = 0;
    x.twb = 0;
    x.threads = P * g.TP;
    x.lds = sizeof(float2) * (size_t)(P + 1) * N;
    x.gp = &pl[1];
    x.fn[XM_ORDER0] = (const void*)&k_gen_xpass<XM_ORDER0>;
    x.fn[XM_MID] = (const void*)&k_gen_xpass<XM_MID>;
    x.fn[XM_LAST] = (const void*)&k_gen_xpass<XM_LAST>;
    x.fn[XM_FWD] = (const void*)&k_gen_xpass<XM_FWD>;
    x.fn[XM_INV] = (const void*)&k_gen_xpass<XM_INV>;
    x.fn[XM_LASTA] = (const void*)&k_gen_xpass<XM_LASTA>;
    x.fn[XM_SKIP] = (const void*)&k_gen_xpass<XM_SKIP>;
    x.fn[XM_ORDER0N] = (const void*)&k_gen_xpass<XM_ORDER0N>;
    x.fn[XM_MIDF] = (const void*)&k_gen_xpass<XM_MIDF>;
    e->xk[0] = x;
    const FftEntry* out = e.get();
    reg.plans[N] = std::move(pl);
    reg.entries[N] = std::move(e);
    return out;
}

}  // namespace ose
