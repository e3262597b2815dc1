// ose_api.hip -- C-ABI implementation of the HIP backend (include/ose_api.h).
//
// A context owns a main and a side HIP stream and, per model, the velocity
// (as v/R), the taper profiles, the symbol tables and the twiddle tables,
// and per shot batch the five complex fields of the fused step:
//     X, S    x-spectral work fields: passes 1-3 run X -> S -> X -> S and
//             pass 4 reads the spectrum from S (S is X, in place, for
//             batches under 1 GiB per field)
//     W0, W1  physical Chebyshev terms (W0 holds the state u between steps)
//     A       accumulator (physical space)
// Shots of one model are batched into one launch per pass (blockIdx.z /
// blockIdx.y = shot), so a model's step is 4K + 1 launches and one step of
// the whole context adds a counter increment.
#include <hip/hip_runtime.h>

#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <algorithm>
#include <string>
#include <vector>

#include "../../../include/ose_api.h"
#include "ose_bnet.h"
#include "ose_entry.h"
#include "ose_kernels.h"
#include "ose_snap.h"

namespace ose {

// ---------------------------------------------------------------------
// non-template kernels
// ---------------------------------------------------------------------
__global__ void k_inject(InjectArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int shot = blockIdx.y;
    if (k >= a.nx) return;
    const int it = *a.it;
    if (it >= a.nt) return;
    const float s = a.s[(long)shot * a.nt + it];
    const float h = a.h[(long)shot * a.nt + it];
    const int sx = a.src_x[shot];
    const float2 e = a.tw[(int)(((long)k * sx) % a.nx)];
    float2* p = a.X0 + (long)shot * a.shot_stride + a.src_line[shot] + k;
    const float2 o = *p;
    *p = make_float2(o.x + s * e.x - h * e.y, o.y + s * e.y + h * e.x);
    if (k == 0) {
        float2* q = a.W0 + (long)shot * a.shot_stride + a.src_line[shot] + sx;
        const float2 w = *q;
        *q = make_float2(w.x + s, w.y + h);
    }
}

__global__ void k_inject_prof(ProfArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int shot = blockIdx.z;
    const ProfDesc d = a.desc[shot];
    const int iy = blockIdx.y % a.wy_max, iz = blockIdx.y / a.wy_max;
    if (k >= a.nx || iy >= d.wy || iz >= d.wz) return;
    const int itv = *a.it;
    if (itv >= a.nt) return;
    const float* pv = a.pv + (long)shot * 3 * PROF_W;
    const float w = pv[PROF_W + iy] * pv[2 * PROF_W + iz];
    const float ls = a.s[(long)shot * a.nt + itv] * w;
    const float lh = a.h[(long)shot * a.nt + itv] * w;
    const long line = (long)shot * a.shot_stride + ((long)(d.z0 + iz) * a.ny + d.y0 + iy) * a.nx;
    const float2 e = a.PX[(long)shot * a.nx + k];
    float2* p = a.X0 + line + k;
    const float2 o = *p;
    *p = make_float2(o.x + ls * e.x - lh * e.y, o.y + ls * e.y + lh * e.x);
    if (k < d.wx) {
        float2* q = a.W0 + line + d.x0 + k;
        const float2 u = *q;
        *q = make_float2(u.x + ls * pv[k], u.y + lh * pv[k]);
    }
}

__global__ void k_advance(int* it)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *it = *it + 1;
}

// One receiver of a listed geometry: its padded cell, its shot's slot in the
// model's batch and where its samples go (trace buffer offset base + it *
// stride; the samples of one step of one shot are contiguous).
struct RecvDesc {
    long cell;
    long base;
    int slot;
    int stride;
};

struct RecvArgs {
    const float2* u;        // W_0 of the model's batch: the new state
    long shot_stride;
    const RecvDesc* d;
    int n;
    int nt_cap;
    const int* it;
    float* tr;
};

// Receivers at listed cells (the geometry of sg_init_device,
// ose_3d_kernel.cu:202-212, read back as in the cell lookup of
// rec_shotdata_3d, ose_3d_kernel.cu:417-423): Im(u) after the step.
__global__ void k_gather_recv(RecvArgs a)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.n) return;
    const int itv = *a.it;
    if (itv >= a.nt_cap) return;
    const RecvDesc e = a.d[r];
    a.tr[e.base + (long)itv * e.stride] = a.u[(long)e.slot * a.shot_stride + e.cell].y;
}

// plain streaming copy, the bandwidth yardstick of the passes: each thread
// moves four independent 16-byte vectors (64 B), one pass over the array
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void k_copy(const f32x4* __restrict__ src, f32x4* __restrict__ dst, long n)
{
    const long base = ((long)blockIdx.x * blockDim.x) * 4 + threadIdx.x;
    f32x4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long i = base + (long)k * blockDim.x;
        if (i < n) v[k] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long i = base + (long)k * blockDim.x;
        if (i < n) __builtin_nontemporal_store(v[k], dst + i);
    }
}

__global__ void k_scale_vr(float2* f, const float* vr, long n, long shot_stride, float s)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2* p = f + (long)blockIdx.y * shot_stride + i;
    const float m = vr[i] * s;
    *p = make_float2(p->x * m, p->y * m);
}

#define OSE_SIZE(N) void fill_entry_##N(FftEntry&);
#include "ose_sizes.inc"
#undef OSE_SIZE

namespace {
struct Registry {
    std::vector<FftEntry> entries;
    Registry()
    {
#define OSE_SIZE(N) { FftEntry e; fill_entry_##N(e); entries.push_back(e); }
#include "ose_sizes.inc"
#undef OSE_SIZE
    }
};
Registry& registry()
{
    static Registry r;
    return r;
}
}  // namespace

const FftEntry* find_compiled(int N)
{
    for (auto& e : registry().entries)
        if (e.N == N) return &e;
    return nullptr;
}

const FftEntry* find_entry(int N)
{
    const FftEntry* e = find_compiled(N);
    return e ? e : generic_entry(N);
}

// compiled in several launch shapes (the autotuner chooses), not a single one
static bool tuned(const FftEntry& e) { return e.ns > 1 || e.nx > 1; }

// smallest extent >= n compiled in several launch shapes (the extent=smooth rounding)
int next_supported(int n)
{
    int best = 0;
    for (auto& e : registry().entries)
        if (tuned(e) && e.N >= n && (best == 0 || e.N < best)) best = e.N;
    return best;
}

int extent_kind(int n)
{
    const FftEntry* e = find_compiled(n);
    if (e) return tuned(*e) ? 1 : 3;
    return generic_entry(n) ? 2 : 0;
}

}  // namespace ose

using namespace ose;

// ---------------------------------------------------------------------
// context
// ---------------------------------------------------------------------
namespace {

enum PassKind {
    PK_YF = 0, PK_Z = 1, PK_YI = 2, PK_XMID = 3, PK_X0 = 4, PK_XLAST = 5, PK_INJ = 6, PK_XSKIP = 7, PK_XMIDF = 8,
    PK_N = 9
};

// Which orders of a step touch the accumulator (see the deferred
// accumulation of k_xpass).  Pass 0 holds W_0 (an extra read) and W_1;
// pass j >= 1 holds W_{j-1} and W_{j+1}; the last pass, K-1, always touches
// (it forms u).  Every term W_m, m = 0..K, is added exactly once, by one of
// the passes that hold it.  The set of touching passes minimises the step's
// x-pass bytes per cell (dynamic programming over the passes; a touch costs
// the accumulator's read and write, 16 B, or 8 B for the first, which only
// writes), e.g. K = 11 touches passes 1 2 5 6 9 10: 468 B/cell of x passes
// instead of 556.
struct AccPlan {
    bool touch = false;
    bool add_prev = false, add_next = false;   // W_{j-1} (W_0 at pass 0), W_{j+1}
    int ea = -1, eb = 0, ec = 0;               // powers of g (XArgs)
};

double x_pass_bytes(int j, int K, bool touch, bool stored)
{
    if (j == K - 1) return stored ? 44.0 : 36.0;      // S, W_{K-2}, vr (+A) read; u, X_0 written
    if (j == 0) return touch ? 44.0 : 28.0;           // S, vr (+W_0) read; W_1, X_1 (+A) written
    return touch ? (stored ? 52.0 : 44.0) : 36.0;     // S, W_{j-1}, vr (+A) read; W_{j+1}, X_{j+1} (+A) written
}

std::vector<AccPlan> plan_accumulation(int K)
{
    // state after deciding pass j: (touch j-1, touch j, any touch so far)
    struct Cell { double cost = -1; std::vector<char> s; };
    Cell dp[2][2][2];
    for (int s0 = 0; s0 < 2; ++s0) {
        if (K == 1 && !s0) continue;
        dp[0][s0][s0].cost = x_pass_bytes(0, K, s0, false);
        dp[0][s0][s0].s = {(char)s0};
    }
    for (int j = 1; j < K; ++j) {
        Cell nd[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int any = 0; any < 2; ++any) {
                    const Cell& c = dp[a][b][any];
                    if (c.cost < 0) continue;
                    for (int sj = (j == K - 1); sj < 2; ++sj) {
                        // W_{j-1} is held by passes j-2 and j (W_0: passes 0 and 1)
                        const bool covered = (j == 1) ? (b || sj) : (a || sj);
                        if (!covered) continue;
                        const double cost = c.cost + x_pass_bytes(j, K, sj, any);
                        Cell& n = nd[b][sj][any || sj];
                        if (n.cost < 0 || cost < n.cost) {
                            n.cost = cost;
                            n.s = c.s;
                            n.s.push_back((char)sj);
                        }
                    }
                }
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int any = 0; any < 2; ++any) dp[a][b][any] = nd[a][b][any];
    }
    const Cell* best = nullptr;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int any = 0; any < 2; ++any) {
                const Cell& c = dp[a][b][any];
                // W_{K-1} is held only by pass K-2
                if (c.cost < 0 || (K >= 2 && !c.s[K - 2])) continue;
                if (!best || c.cost < best->cost) best = &c;
            }
    std::vector<AccPlan> plan(K);
    if (!best) return plan;   // unreachable for K >= 1
    for (int j = 0; j < K; ++j) plan[j].touch = best->s[j];
    // assign W_m to the pass that writes it if that pass touches, else to
    // the pass that reads it (W_0: pass 0, else pass 1)
    for (int m = 0; m <= K; ++m) {
        const int p = (m == 0) ? (plan[0].touch ? 0 : 1) : (plan[m - 1].touch ? m - 1 : m + 1);
        if (m == 0 && p == 0) plan[0].add_prev = true;
        else if (p == m - 1) plan[p].add_next = true;
        else plan[p].add_prev = true;
    }
    // taper applications of W_m's term by the end of order j in the
    // reference: W_0, W_1 are added at order 0 and tapered at orders 1..j;
    // W_m (m >= 2) is added at order m-1 and tapered at orders m-1..j
    auto napp = [](int m, int j) { return m <= 1 ? j : j - m + 2; };
    int prev = -1;
    for (int j = 0; j < K; ++j) {
        if (!plan[j].touch) continue;
        plan[j].ea = prev < 0 ? -1 : j - prev;
        plan[j].eb = j == 0 ? 0 : napp(j - 1, j);
        plan[j].ec = napp(j + 1, j);
        prev = j;
    }
    // the kernels raise g to powers 0..3 only (gpow)
    for (int j = 0; j < K; ++j)
        if (plan[j].touch && (plan[j].ea > 3 || plan[j].eb > 3 || plan[j].ec > 3 || plan[j].eb < 0)) return {};
    return plan;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

struct Model {
    ose_model_desc d{};
    std::vector<double> coef;
    std::vector<AccPlan> acc_plan;   // per order of a step
    std::vector<float> gx, gy, gz;
    const FftEntry *ex = nullptr, *ey = nullptr, *ez = nullptr;
    // chosen launch shapes (index into the entry's variants)
    int sel_yf = 0, sel_z = 0, sel_yi = 0, sel_x = 0;
    int sel_mode[9] = {};          // x-pass shape per XMode the step runs (LASTA takes LAST's)
    bool tuned = false;
    long ncell = 0;
    DevBuf vr, g, k2, tw;          // g: gx|gy|gz, k2: kx2|ky2|kz2, tw: twx|twy|twz
    float* d_vr = nullptr;
    float *d_gx = nullptr, *d_gy = nullptr, *d_gz = nullptr;
    float *d_kx2 = nullptr, *d_ky2 = nullptr, *d_kz2 = nullptr;
    float2 *d_twx = nullptr, *d_twy = nullptr, *d_twz = nullptr;
    // shot batch
    std::vector<int> shots;        // context shot ids
    int nt_cap = 0;
    DevBuf fields, traces, src, wav;
    // X, S: x-spectral work fields (passes 1-3 run X -> S -> X -> S; S == X
    // for small batches, which run them in place), W[0] the physical state
    // u = W_0 and with W[1] the alternating physical W_j, A the accumulator
    float2 *X = nullptr, *S = nullptr, *W[2] = {nullptr, nullptr}, *A = nullptr;
    float* d_traces = nullptr;
    long* d_src_line = nullptr;
    int* d_src_x = nullptr;
    float *d_s = nullptr, *d_h = nullptr;
    int rec_ny = 0, rec_nx = 0;    // plane receivers (0 x 0 when every shot has a list)
    // listed receivers of the batch (ose_set_receivers)
    DevBuf pbuf;                   // extended sources: descriptors, PX, profiles
    ProfDesc* d_pdesc = nullptr;
    float2* d_PX = nullptr;
    float* d_pv = nullptr;
    int prof_wy = 0, prof_wz = 0;  // 0: point sources only (k_inject)
    DevBuf rdesc, rtr;
    RecvDesc* d_rdesc = nullptr;
    float* d_rtr = nullptr;
    int nrecv = 0;
    // learned thin pad (boundary network)
    bool learned = false;          // a network is set for this model
    bool learned_on = false;       // used by the step (cleared once the guard trips)
    bool replace_taper = false;    // while the network is on, the pad taper is 1
    float guard = 2.0f;
    DevBuf ones;                   // nxb+nyb+nzb: the taper of the learned mode (1, or the
                                   // classical taper once the guard trips)
    BnetGeom bg{};
    DevBuf bnw, bnbuf;
    uint16_t* d_wbf = nullptr;     // [3][27][16][16] bf16
    float* d_bias = nullptr;       // [3][16]
    BnetBufs bb{};
};

struct Shot {
    int model = 0;
    int slot = 0;
    int sx = 0, sy = 0, sz = 0;
    std::vector<float> s, h;
    int nt = 0;
    std::vector<long> rec;         // listed receivers (padded cells); empty: the plane
    long rbase = 0;                // offset of this shot's samples in the model's list traces
    // extended source (ose_set_source_profile); wx == 0: the point (sx, sy, sz)
    ProfDesc prof{0, 0, 0, 0, 0, 0};
    std::vector<float> px, py, pz;
};

struct Pending {
    int kind;
    hipEvent_t e0, e1;
    double bytes;
};

// Snapshot streaming: every `every` steps each shot's state is encoded on
// the compute stream into a free device slot (the enqueuing thread waits for
// one only when the host side has fallen that far behind).  A copier thread
// moves each encoded snapshot into a free pinned host slot on a second
// stream and frees the device slot; a writer thread hands it to the callback
// and frees the host slot.  There are as many device slots as shots (so one
// snapshot of every shot never waits) within half of the free device memory,
// and three host slots.
struct SnapStream {
    static constexpr int kHost = 3;
    int every = 0, lanes = 3, format = SNAP_F32, interior = 0;
    ose_snapshot_fn fn = nullptr;
    void* user = nullptr;
    bool started = false;
    size_t slot_bytes = 0;
    std::vector<std::unique_ptr<DevBuf>> dev;
    std::vector<hipEvent_t> ready;
    std::vector<char> dev_busy;
    void* host[kHost] = {};
    hipEvent_t copied[kHost] = {};
    bool host_busy[kHost] = {};
    hipStream_t copy = nullptr;
    int next = 0;
    std::thread copier, writer;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<int, ose_snapshot_info>> to_copy, to_write;
    bool stop = false, copier_done = false;
    long enqueued = 0, delivered = 0;
    int err = OSE_OK;
    std::string msg;
};

}  // namespace

struct ose_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second lane of the step (enqueue_step) and its fork/join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int nstreams = 2;
    std::vector<std::unique_ptr<Model>> models;
    std::vector<Shot> shots;
    DevBuf it_buf;
    int* d_it = nullptr;
    bool finalized = false;
    bool use_graph = true;
    bool profile = false;
    bool autotune = true;
    bool generic_only = false;   // ose_set_fft_path(ctx, 1): runtime-length kernels on every axis
    // the step's graphs: [0] the whole step, or with two lanes [0] lane 0's
    // head, [1] lane 0's body, [2] lane 1, [3] the tail
    hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    int ngraph = 0;
    std::string err;
    // profiling
    std::vector<hipEvent_t> ev_pool;
    std::vector<Pending> pending;
    double st_ms[PK_N] = {};
    long st_n[PK_N] = {};
    double st_bytes[PK_N] = {};
    // snapshot streaming and the host's count of enqueued steps
    std::unique_ptr<SnapStream> snap;
    int host_it = 0;

    int fail(int code, const char* fmt, ...)
    {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
};

#define HIPCHK(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return (ctx)->fail(OSE_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

// two timing events, destroyed on every return path
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t create()
    {
        const hipError_t e = hipEventCreate(&e0);
        return e == hipSuccess ? hipEventCreate(&e1) : e;
    }
    ~EventPair()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

namespace {

int dev_alloc(ose_ctx* c, DevBuf& b, size_t bytes)
{
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; }
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return c->fail(OSE_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return OSE_OK;
}

int set_device(ose_ctx* c)
{
    HIPCHK(c, hipSetDevice(c->device));
    return OSE_OK;
}

// root table exp(-2 pi i j / n), FP64 then rounded
// Root table of one axis of extent n = N1 N2, three blocks of n entries:
//   [0, n)    w^j, w = exp(-2 pi i / n)          (injection phase factors)
//   [n, 2n)   w^(t k1) at t N1 + k1, t < N2     (twiddles after the N1-point
//             DFTs: each thread's N1 values are contiguous)
//   [2n, 3n)  w^(t n2) at t N2 + n2, t < N1     (twiddles before the inverse
//             N1-point DFTs)
// Angles in FP64, reduced mod n before the cos/sin.
void roots(int n, std::vector<float2>& out)
{
    const FftEntry* e = find_compiled(n);   // the runtime-length path reads block 0 only
    const int n1 = e ? e->N1 : n, n2 = e ? e->N2 : 1;
    out.resize(3 * (size_t)n);
    auto w = [n](long j) {
        const double a = -2.0 * M_PI * (double)(j % n) / (double)n;
        return make_float2((float)std::cos(a), (float)std::sin(a));
    };
    for (int j = 0; j < n; ++j) out[j] = w(j);
    for (int t = 0; t < n2; ++t)
        for (int k1 = 0; k1 < n1; ++k1) out[n + (size_t)t * n1 + k1] = w((long)t * k1);
    for (int t = 0; t < n1; ++t)
        for (int k2 = 0; k2 < n2; ++k2) out[2 * (size_t)n + (size_t)t * n2 + k2] = w((long)t * k2);
}

// squared folded wavenumbers (scale_kxkz_3D, ose_3d_kernel.cu:109-123)
void k2table(int n, double d, std::vector<float>& out)
{
    out.resize(n);
    const double dk = 2.0 * M_PI / (d * (double)n);
    for (int i = 0; i < n; ++i) {
        const double k = dk * (double)(i <= n / 2 ? i : n - i);
        out[i] = (float)(k * k);
    }
}

hipEvent_t take_event(ose_ctx* c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int launch(ose_ctx* c, int kind, double bytes, const void* fn, dim3 grid, dim3 block,
           size_t lds, void* args)
{
    void* params[] = {args};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->profile && kind >= 0) {
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0) (void)hipEventRecord(e0, c->stream);
    }
    hipError_t e = hipLaunchKernel(fn, grid, block, params, lds, c->stream);
    if (e != hipSuccess)
        return c->fail(OSE_EDEVICE, "kernel launch (pass %d) failed: %s", kind, hipGetErrorString(e));
    if (c->profile && e0 && e1) {
        (void)hipEventRecord(e1, c->stream);
        c->pending.push_back({kind, e0, e1, bytes});
    }
    return OSE_OK;
}

int collect(ose_ctx* c)
{
    for (auto& p : c->pending) {
        float ms = 0.f;
        HIPCHK(c, hipEventSynchronize(p.e1));
        HIPCHK(c, hipEventElapsedTime(&ms, p.e0, p.e1));
        c->st_ms[p.kind] += ms;
        c->st_n[p.kind] += 1;
        c->st_bytes[p.kind] += p.bytes;
        c->ev_pool.push_back(p.e0);
        c->ev_pool.push_back(p.e1);
    }
    c->pending.clear();
    return OSE_OK;
}

// The chunked z pass indexes tiles in 32 bits with the exact magic division
// of StridedTile: < 2^24 tiles, extents < 2^16.
bool chunk_usable(const StridedKern& k, int nx, int ny, int nshot)
{
    if (k.chunk <= 0) return true;
    const double gx = (nx + k.T - 1) / k.T;
    return gx * ny * nshot < 16777216.0 && gx < 65536 && ny < 65536;
}

// launch-shape code of a strided kernel: T, negative for the split exchange,
// 100 * chunk + T for the chunked z pass
int scode(const StridedKern& k)
{
    if (k.chunk > 0) return 100 * k.chunk + k.T;
    return k.split ? -k.T : k.T;
}

// strided pass over the y axis (es = nx) or the z axis (es = nx*ny)
int launch_strided(ose_ctx* c, int kind, const StridedKern& k, const float2* in, float2* out, int nx,
                   int ny, int nz, bool zaxis, int nshot, long shot_stride, const float2* tw,
                   const Model* m, double bytes)
{
    if (!k.fn || !in || !out || !tw) return c->fail(OSE_EINVAL, "strided pass launched with a missing buffer");
    // the compiled kernels address a thread's column with a 32-bit byte
    // offset (col + t es) * 8, t < 32 (the runtime-length kernels use 64 bits)
    if (!k.gp && (double)(k.T + 32.0 * (zaxis ? (double)nx * ny : (double)nx)) * sizeof(float2) >= 4294967296.0)
        return c->fail(OSE_EUNSUPPORTED, "strided pass: plane of %d x %d cells too large", nx, ny);
    StridedArgs a{};
    if (k.gp) a.g = *k.gp;
    a.in = in;
    a.out = out;
    a.ncol = nx;
    a.es = zaxis ? (long)nx * ny : (long)nx;
    a.os = zaxis ? (long)nx : (long)nx * ny;
    a.shot_stride = shot_stride;
    a.tw = tw;
    dim3 grid((nx + k.T - 1) / k.T, zaxis ? ny : nz, nshot);
    // XCD-contiguous tile order (StridedTile); its exact division by the
    // grid extents holds for < 2^24 workgroups and extents < 2^16
    a.xcd = (double)grid.x * grid.y * grid.z < 16777216.0 && grid.x < 65536 && grid.y < 65536;
    a.mgx = (1ULL << 40) / grid.x + 1;
    a.mgy = (1ULL << 40) / grid.y + 1;
    if (k.chunk > 0) {
        // chunked z pass: a 1-D grid of chunks of k.chunk tiles; the kernel
        // decodes tiles with the same exact division, so it needs its bounds
        if (!zaxis || !m) return c->fail(OSE_EINVAL, "chunked strided kernels run the z symbol pass only");
        if (!chunk_usable(k, nx, ny, nshot))
            return c->fail(OSE_EUNSUPPORTED, "chunked z pass: %u x %u x %d tiles exceed its tile indexing", grid.x,
                           grid.y, nshot);
        a.tgx = (int)grid.x;
        a.tgy = (int)grid.y;
        a.ntiles = grid.x * grid.y * grid.z;
        a.chunk = k.chunk;
        grid = dim3((a.ntiles + k.chunk - 1) / k.chunk, 1, 1);
    }
    if (m) {
        a.kx2 = m->d_kx2;
        a.ky2 = m->d_ky2;
        a.kz2 = m->d_kz2;
        a.invN = (float)(1.0 / ((double)nx * ny * nz));
    }
    return launch(c, kind, bytes, k.fn, grid, dim3(k.threads), k.lds, &a);
}

int launch_x(ose_ctx* c, int kind, const XKern& k, int mode, XArgs& a, int nlines, int nshot,
             double bytes)
{
    a.nlines = nlines;
    if (!a.twf) a.twf = a.tw;
    if (k.gp) a.g = *k.gp;
    // every pointer the mode dereferences must be set before launching
    bool ok = a.tw != nullptr && k.fn[mode] != nullptr;
    switch (mode) {
    case XM_FWD: ok = ok && a.A && a.Xout; break;
    case XM_INV: ok = ok && a.A && a.Xprev; break;
    case XM_ORDER0: ok = ok && a.S && a.Win && a.Wout && a.Xout && a.A && a.vr; break;
    case XM_ORDER0N: ok = ok && a.S && a.Wout && a.Xout && a.vr; break;
    case XM_SKIP: ok = ok && a.S && a.Win && a.Wout && a.Xout && a.vr; break;
    case XM_MID:
    case XM_MIDF: ok = ok && a.S && a.Win && a.Wout && a.Xout && a.A && a.vr && a.gx && a.gy && a.gz; break;
    case XM_LAST:
        ok = ok && a.S && a.Win && a.Wout && a.Xout && a.A && a.vr && a.gx && a.gy && a.gz && a.traces && a.it;
        break;
    case XM_LASTA:
        ok = ok && a.S && a.Win && a.Wout && a.A && a.vr && a.gx && a.gy && a.gz && a.traces && a.it;
        break;
    default: ok = false;
    }
    if (!ok) return c->fail(OSE_EINVAL, "x pass mode %d launched with a missing buffer", mode);
    dim3 grid((nlines + k.P - 1) / k.P, nshot);
    return launch(c, kind, bytes, k.fn[mode], grid, dim3(k.threads), k.lds, &a);
}

int prepare_kernels(ose_ctx* c, const FftEntry* e)
{
    // allow > 64 KiB of dynamic LDS where a launch shape needs it
    for (int v = 0; v < e->ns; ++v)
        for (const StridedKern* k : {&e->yf[v], &e->yi[v], &e->zs[v]})
            if (k->lds > 65536)
                HIPCHK(c, hipFuncSetAttribute(k->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)k->lds));
    for (int v = 0; v < e->nx; ++v)
        for (const void* f : e->xk[v].fn)
            if (e->xk[v].lds > 65536)
                HIPCHK(c, hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->xk[v].lds));
    return OSE_OK;
}

XArgs base_xargs(const Model& m, int nshot)
{
    XArgs a{};
    a.vr = m.d_vr;
    a.gx = m.d_gx;
    a.gy = m.d_gy;
    a.gz = m.d_gz;
    if (m.learned_on && m.replace_taper && m.ones.p) {
        a.gx = (const float*)m.ones.p;
        a.gy = a.gx + m.d.nxb;
        a.gz = a.gy + m.d.nyb;
    }
    a.tw = m.d_twx;
    a.shot_stride = m.ncell;
    a.ny = m.d.nyb;
    a.inv_nx = (float)(1.0 / (double)m.d.nxb);
    // MID/LAST default to the reference's per-order update A <- g (A + c1 W_{j+1})
    a.ea = 1;
    a.eb = 0;
    a.ec = 1;
    a.cp = 0.f;
    (void)nshot;
    return a;
}

// Injection and the K operator applications of one OSE step (the body of
// wavefield_extro_ori2) for the shots [s0, s0 + nshot) of model m: every
// per-shot buffer is offset to shot s0 and the launches cover nshot shots.
// Only the launches numbered [p0, p1) are enqueued: 0 is the injection,
// 1 + 4j .. 4 + 4j are order j's forward-y, z, inverse-y and x passes.
int enqueue_model_orders(ose_ctx* c, Model& m, int* d_it, int s0, int nshot, int p0 = 0, int p1 = 1 << 30)
{
    if (nshot <= 0) return OSE_OK;
    auto in = [&](int p) { return p >= p0 && p < p1; };
    const int nx = m.d.nxb, ny = m.d.nyb, nz = m.d.nzb;
    const long cells = m.ncell * nshot;
    const long so = m.ncell * s0;                // field offset of shot s0
    float2* const X = m.X + so;
    float2* const S = m.S + so;
    float2* const W[2] = {m.W[0] + so, m.W[1] + so};
    float2* const A = m.A + so;
    int rc;
    // source injection into X0 (x-spectral) and W_0 (physical)
    if (!in(0)) {
    } else if (m.prof_wy > 0) {
        ProfArgs a{};
        a.X0 = X;
        a.W0 = W[0];
        a.shot_stride = m.ncell;
        a.nx = nx;
        a.ny = ny;
        a.desc = m.d_pdesc + s0;
        a.PX = m.d_PX + (long)s0 * nx;
        a.pv = m.d_pv + (long)s0 * 3 * PROF_W;
        a.s = m.d_s + (long)s0 * m.nt_cap;
        a.h = m.d_h + (long)s0 * m.nt_cap;
        a.nt = m.nt_cap;
        a.it = d_it;
        a.wy_max = m.prof_wy;
        if ((rc = launch(c, PK_INJ, 16.0 * nx * m.prof_wy * m.prof_wz * nshot, (const void*)&k_inject_prof,
                         dim3((nx + 255) / 256, m.prof_wy * m.prof_wz, nshot), dim3(256), 0, &a)))
            return rc;
    } else {
        InjectArgs a{};
        a.X0 = X;
        a.W0 = W[0];
        a.shot_stride = m.ncell;
        a.nx = nx;
        a.src_line = m.d_src_line + s0;
        a.src_x = m.d_src_x + s0;
        a.s = m.d_s + (long)s0 * m.nt_cap;
        a.h = m.d_h + (long)s0 * m.nt_cap;
        a.nt = m.nt_cap;
        a.it = d_it;
        a.tw = m.d_twx;
        if ((rc = launch(c, PK_INJ, 16.0 * nx * nshot, (const void*)&k_inject,
                         dim3((nx + 63) / 64, nshot), dim3(64), 0, &a)))
            return rc;
    }
    const int K = m.d.K;
    for (int j = 0; j < K; ++j) {
        // passes 1-3 out of place, alternating between X and S: forward-y
        // X -> S, the z pass S -> X, inverse-y X -> S (in place where S == X)
        if (in(1 + 4 * j) && (rc = launch_strided(c, PK_YF, m.ey->yf[m.sel_yf], X, S, nx, ny, nz, false, nshot,
                                                  m.ncell, m.d_twy, nullptr, 16.0 * cells)))
            return rc;
        if (in(2 + 4 * j) && (rc = launch_strided(c, PK_Z, m.ez->zs[m.sel_z], S, X, nx, ny, nz, true, nshot,
                                                  m.ncell, m.d_twz, &m, 16.0 * cells)))
            return rc;
        if (in(3 + 4 * j) && (rc = launch_strided(c, PK_YI, m.ey->yi[m.sel_yi], X, S, nx, ny, nz, false, nshot,
                                                  m.ncell, m.d_twy, nullptr, 16.0 * cells)))
            return rc;
        if (!in(4 + 4 * j)) continue;
        XArgs a = base_xargs(m, nshot);
        a.S = S;
        a.A = A;
        a.Xout = X;
        const AccPlan& pl = m.acc_plan[j];
        int mode, kind;
        // the shots of a launch share v/R, read once (x_block keeps a line
        // block's shots together on one XCD, so the others find it in L2)
        double bytes = x_pass_bytes(j, K, pl.touch, pl.ea >= 0) * cells - 4.0 * (double)(cells - m.ncell);
        if (j == 0) {
            mode = pl.touch ? XM_ORDER0 : XM_ORDER0N;
            kind = PK_X0;
            a.Win = W[0];
            a.Wout = W[1];
            a.c0 = pl.add_prev ? (float)m.coef[0] : 0.f;
            a.c1 = pl.add_next ? (float)m.coef[1] : 0.f;
        } else {
            // W_{j-1} lives in W[(j-1)%2] and W_{j+1} replaces it
            a.Win = a.Wout = W[(j - 1) % 2];
            a.c1 = pl.add_next ? (float)m.coef[j + 1] : 0.f;
            a.cp = pl.add_prev ? (float)m.coef[j - 1] : 0.f;
            a.ea = pl.ea;
            a.eb = pl.eb;
            a.ec = pl.ec;
            if (j == K - 1) {
                mode = m.learned_on ? XM_LASTA : XM_LAST;
                kind = PK_XLAST;
                a.Wout = W[0];      // the new state u
                a.trace_shot_stride = (long)m.nt_cap * m.rec_ny * m.rec_nx;
                a.traces = m.d_traces ? m.d_traces + s0 * a.trace_shot_stride : nullptr;
                a.rec_z = m.d.bu + m.d.rec_z;
                a.rec_y0 = m.d.bf;
                a.rec_ny = m.rec_ny;
                a.rec_x0 = m.d.bl;
                a.rec_nx = m.rec_nx;
                a.it = d_it;
                a.nt_cap = m.nt_cap;
                if (m.learned_on) bytes -= 8.0 * cells;   // LASTA writes no X_0
            } else if (pl.touch) {
                mode = pl.ea >= 0 ? XM_MID : XM_MIDF;
                kind = pl.ea >= 0 ? PK_XMID : PK_XMIDF;
            } else {
                mode = XM_SKIP;
                kind = PK_XSKIP;
            }
        }
        const XKern& xk = m.ex->xk[m.sel_mode[mode == XM_LASTA ? XM_LAST : mode]];
        if ((rc = launch_x(c, kind, xk, mode, a, ny * nz, nshot, bytes))) return rc;
    }
    return OSE_OK;
}

// The rest of model m's step, over all of its shots: the boundary network
// (learned pad) and the listed-receiver gather.
int enqueue_model_tail(ose_ctx* c, Model& m, int* d_it)
{
    const int nshot = (int)m.shots.size();
    if (nshot == 0) return OSE_OK;
    const int ny = m.d.nyb, nz = m.d.nzb;
    const long cells = m.ncell * nshot;
    int rc;
    if (m.learned_on) {
        // boundary network on the pad of the new state u (in W[0]), then X0 = FFT_x(u)
        if ((rc = bnet_forward(c->stream, m.bg, m.d_wbf, m.d_bias, m.W[0], m.d_vr, m.ncell, nshot, m.bb, m.guard,
                               true)))
            return c->fail(rc, "boundary network launch failed");
        if (m.replace_taper) {
            float* eff = (float*)m.ones.p;
            if ((rc = bnet_taper_select(c->stream, eff, m.d_gx, m.d.nxb, m.d_gy, m.d.nyb, m.d_gz, m.d.nzb,
                                        m.bb.tripped)))
                return c->fail(rc, "boundary network taper launch failed");
        }
        XArgs a = base_xargs(m, nshot);
        a.A = m.W[0];
        a.Xout = m.X;
        if ((rc = launch_x(c, -1, m.ex->xk[m.sel_x], XM_FWD, a, ny * nz, nshot, 16.0 * cells)))
            return rc;
    }
    if (m.nrecv > 0) {
        RecvArgs r{};
        r.u = m.W[0];
        r.shot_stride = m.ncell;
        r.d = m.d_rdesc;
        r.n = m.nrecv;
        r.nt_cap = m.nt_cap;
        r.it = d_it;
        r.tr = m.d_rtr;
        if ((rc = launch(c, -1, 0.0, (const void*)&k_gather_recv, dim3((m.nrecv + 255) / 256), dim3(256), 0, &r)))
            return rc;
    }
    return OSE_OK;
}

// The step's lanes.  With two streams (the default outside per-pass
// profiling) and shots on two or more models, the shots are dealt to two
// lanes -- a model of two or more shots in halves, single-shot models
// whole, each item to the lane with fewer padded cells so far -- that run
// side by side on the main and the side stream, forked and joined by
// events.  The lanes touch disjoint shots, so they need no other ordering.
// The boundary network, the receiver gather and the step counter (the
// tail) follow the join on the main stream.  Measured on MI355X: 16 shots
// of 600^2 x 480 plus four smaller models step 4 % faster in two lanes (the
// small models' launches no longer run alone), while the shots of a single
// model gain nothing (768^3 x 8 shots: 0.6 %, within noise; the 784^3 runs
// lose 2-3 %), so one model's shots stay in one lane.  lane[1] is left
// empty where there is nothing to split.
struct LaneItem {
    Model* m;
    int s0, n;
    double cells;
};

void plan_lanes(ose_ctx* c, std::vector<LaneItem> lane[2])
{
    lane[0].clear();
    lane[1].clear();
    std::vector<LaneItem> items;
    int models = 0;
    for (auto& m : c->models) models += !m->shots.empty();
    const bool split = c->nstreams > 1 && !c->profile && c->side && models >= 2;
    for (auto& m : c->models) {
        const int n = (int)m->shots.size();
        if (n == 0) continue;
        if (split && n >= 2) {
            const int h = n - n / 2;
            items.push_back({m.get(), 0, h, (double)m->ncell * h});
            items.push_back({m.get(), h, n - h, (double)m->ncell * (n - h)});
        } else {
            items.push_back({m.get(), 0, n, (double)m->ncell * n});
        }
    }
    if (!split || items.size() < 2) {
        lane[0] = items;
        return;
    }
    std::stable_sort(items.begin(), items.end(), [](const LaneItem& a, const LaneItem& b) { return a.cells > b.cells; });
    double load[2] = {0.0, 0.0};
    for (const LaneItem& it : items) {
        const int l = load[1] < load[0] ? 1 : 0;
        lane[l].push_back(it);
        load[l] += it.cells;
    }
}

// part 0: the whole lane; 1: the head of its first item (injection,
// forward-y and z of order 0); 2: the rest.  The second lane starts after
// the first lane's head, so the two lanes run two passes apart and one's
// z pass meets the other's x or y passes instead of another z pass.
constexpr int kLaneHead = 3;

int enqueue_lane(ose_ctx* c, const std::vector<LaneItem>& items, int part)
{
    int rc;
    for (size_t i = 0; i < items.size(); ++i) {
        const LaneItem& it = items[i];
        int p0 = 0, p1 = 1 << 30;
        if (part == 1) {
            if (i) break;
            p1 = kLaneHead;
        } else if (part == 2 && i == 0) {
            p0 = kLaneHead;
        }
        if ((rc = enqueue_model_orders(c, *it.m, c->d_it, it.s0, it.n, p0, p1))) return rc;
    }
    return OSE_OK;
}

int enqueue_tail(ose_ctx* c)
{
    int rc;
    for (auto& m : c->models)
        if ((rc = enqueue_model_tail(c, *m, c->d_it))) return rc;
    return launch(c, -1, 0.0, (const void*)&k_advance, dim3(1), dim3(64), 0, &c->d_it);
}

// fork the side stream off the main one / join it back
int fork_side(ose_ctx* c)
{
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    return OSE_OK;
}

int join_side(ose_ctx* c)
{
    HIPCHK(c, hipEventRecord(c->ev_join, c->side));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return OSE_OK;
}

// one step, launched eagerly
int enqueue_step(ose_ctx* c)
{
    std::vector<LaneItem> lane[2];
    plan_lanes(c, lane);
    int rc;
    if (lane[1].empty()) {
        if ((rc = enqueue_lane(c, lane[0], 0))) return rc;
        return enqueue_tail(c);
    }
    if ((rc = enqueue_lane(c, lane[0], 1))) return rc;
    if ((rc = fork_side(c))) return rc;
    rc = enqueue_lane(c, lane[0], 2);
    if (!rc) {
        hipStream_t main = c->stream;
        c->stream = c->side;   // launch() enqueues on c->stream
        rc = enqueue_lane(c, lane[1], 0);
        c->stream = main;
    }
    const int jr = join_side(c);   // joined even after a failed launch
    if (rc) return rc;
    if (jr) return jr;
    return enqueue_tail(c);
}

// Launch-shape autotuner: time every compiled variant of each pass on the
// model's own (still zero) buffers and keep the fastest.  Linear passes on
// zero fields leave them zero, so the state is unchanged.
int time_variant(ose_ctx* c, const std::function<int()>& body, int reps, float* ms)
{
    int rc;
    if ((rc = body())) return rc;   // warm-up
    EventPair ev;
    HIPCHK(c, ev.create());
    HIPCHK(c, hipEventRecord(ev.e0, c->stream));
    for (int r = 0; r < reps && !rc; ++r) rc = body();
    HIPCHK(c, hipEventRecord(ev.e1, c->stream));
    HIPCHK(c, hipEventSynchronize(ev.e1));
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, ev.e0, ev.e1));
    *ms = t / reps;
    return rc;
}

int autotune_model(ose_ctx* c, Model& m)
{
    const int nshot = (int)m.shots.size();
    if (m.tuned || nshot == 0) return OSE_OK;
    const int nx = m.d.nxb, ny = m.d.nyb, nz = m.d.nzb;
    const bool prof = c->profile;
    c->profile = false;
    int rc = OSE_OK;
    float best, t;
    // keep each timed region around a few milliseconds
    const int reps = std::max(1, std::min(5, (int)(2e8 / ((double)m.ncell * nshot))));
    best = 1e30f;
    for (int v = 0; v < m.ey->ns && !rc; ++v) {
        if (!m.ey->yf[v].fn) continue;   // a chunked z-pass slot
        rc = time_variant(c, [&] { return launch_strided(c, -1, m.ey->yf[v], m.X, m.S, nx, ny, nz, false, nshot,
                                                         m.ncell, m.d_twy, nullptr, 0.0); }, reps, &t);
        if (!rc && t < best) best = t, m.sel_yf = v;
    }
    best = 1e30f;
    for (int v = 0; v < m.ez->ns && !rc; ++v) {
        if (!chunk_usable(m.ez->zs[v], nx, ny, nshot)) continue;
        rc = time_variant(c, [&] { return launch_strided(c, -1, m.ez->zs[v], m.S, m.X, nx, ny, nz, true, nshot,
                                                         m.ncell, m.d_twz, &m, 0.0); }, reps, &t);
        if (!rc && t < best) best = t, m.sel_z = v;
    }
    best = 1e30f;
    for (int v = 0; v < m.ey->ns && !rc; ++v) {
        if (!m.ey->yi[v].fn) continue;
        rc = time_variant(c, [&] { return launch_strided(c, -1, m.ey->yi[v], m.X, m.S, nx, ny, nz, false, nshot,
                                                         m.ncell, m.d_twy, nullptr, 0.0); }, reps, &t);
        if (!rc && t < best) best = t, m.sel_yi = v;
    }
    // x pass: every mode the step runs gets its own fastest shape -- MID
    // (middle orders that update the accumulator; also the shape of the
    // transform utilities), SKIP (those that leave it alone), MIDF (the
    // first update), the first order and the last order
    bool any_skip = false, any_midf = false;
    for (int j = 1; j + 1 < m.d.K; ++j) {
        const AccPlan& pl = m.acc_plan[j];
        any_skip = any_skip || !pl.touch;
        any_midf = any_midf || (pl.touch && pl.ea < 0);
    }
    // the last order's entry: with the learned pad on, the step runs
    // XM_LASTA there (XM_LAST only after a guard trip), and a shape that
    // suits one need not suit the other (800^3: 44.9 against 30.7 ms in
    // XM_LAST's shape), so the mode that runs picks the shape of both
    const int last = m.learned_on ? XM_LASTA : XM_LAST;
    const int modes[5] = {XM_MID, XM_SKIP, XM_MIDF, m.acc_plan[0].touch ? XM_ORDER0 : XM_ORDER0N, last};
    const bool used[5] = {true, any_skip, any_midf, true, true};
    float best_m[5] = {1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
    for (int v = 0; v < m.ex->nx && !rc; ++v) {
        // all on zero buffers; the last order records no receivers here
        XArgs a = base_xargs(m, nshot);
        a.S = m.S;
        a.A = m.A;
        a.Win = a.Wout = m.W[1];
        a.Xout = m.X;
        a.c1 = 0.5f;
        a.traces = m.d_traces ? m.d_traces : reinterpret_cast<float*>(m.A);   // never written: rec_ny = 0
        a.it = c->d_it;
        a.rec_ny = 0;
        for (int i = 0; i < 5 && !rc; ++i) {
            if (!used[i]) continue;
            rc = time_variant(c, [&] { return launch_x(c, -1, m.ex->xk[v], modes[i], a, ny * nz, nshot, 0.0); },
                              reps, &t);
            if (!rc && t < best_m[i]) best_m[i] = t, m.sel_mode[modes[i]] = v;
        }
    }
    m.sel_x = m.sel_mode[XM_MID];
    for (int i = 0; i < 5; ++i)
        if (!used[i]) m.sel_mode[modes[i]] = m.sel_x;
    m.sel_mode[XM_LAST] = m.sel_mode[XM_LASTA] = m.sel_mode[last];
    c->profile = prof;
    if (!rc) m.tuned = true;
    return rc;
}

// device buffers of the boundary network for nshot shots (training adds
// the pre-activations)
int alloc_bnet_bufs(ose_ctx* c, Model& m, int nshot, bool training)
{
    if (nshot > 64) return c->fail(OSE_EINVAL, "boundary network: at most 64 shots per model");
    const long n0 = m.bg.sh[0].count, n1 = m.bg.sh[1].count, n2 = m.bg.sh[2].count, n3 = m.bg.sh[3].count;
    size_t off = 0;
    auto take = [&off](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t o_feat = take(sizeof(uint16_t) * 16 * n3 * nshot);
    const size_t o_h1 = take(sizeof(uint16_t) * 16 * n2 * nshot);
    const size_t o_h2 = take(sizeof(uint16_t) * 16 * n1 * nshot);
    const size_t o_y1 = training ? take(sizeof(float) * 16 * n2) : 0;
    const size_t o_y2 = training ? take(sizeof(float) * 16 * n1) : 0;
    const size_t o_y3 = take(sizeof(float) * 2 * n0 * nshot);
    const size_t o_st = take(sizeof(float) * 4 * nshot);
    const size_t o_tr = take(sizeof(int) * 4);
    int rc;
    if ((rc = dev_alloc(c, m.bnbuf, off))) return rc;
    char* b = (char*)m.bnbuf.p;
    m.bb.feat = (uint16_t*)(b + o_feat);
    m.bb.h1 = (uint16_t*)(b + o_h1);
    m.bb.h2 = (uint16_t*)(b + o_h2);
    m.bb.y1 = training ? (float*)(b + o_y1) : nullptr;
    m.bb.y2 = training ? (float*)(b + o_y2) : nullptr;
    m.bb.y3 = (float*)(b + o_y3);
    m.bb.stats = (float*)(b + o_st);
    m.bb.tripped = (int*)(b + o_tr);
    HIPCHK(c, hipMemset(m.bb.tripped, 0, sizeof(int) * 4));
    return OSE_OK;
}

int finalize(ose_ctx* c)
{
    if (c->finalized) return OSE_OK;
    int rc;
    if ((rc = set_device(c))) return rc;
    for (auto& mp : c->models) {
        Model& m = *mp;
        const int nshot = (int)m.shots.size();
        if (nshot == 0) continue;
        m.nt_cap = 0;
        for (int id : m.shots) m.nt_cap = std::max(m.nt_cap, c->shots[id].nt);
        // the plane gather runs unless every shot of the model lists its receivers
        bool plane = false;
        long rsamples = 0;
        m.nrecv = 0;
        for (int id : m.shots) {
            Shot& sh = c->shots[id];
            plane |= sh.rec.empty();
            sh.rbase = rsamples;
            rsamples += (long)m.nt_cap * (long)sh.rec.size();
            m.nrecv += (int)sh.rec.size();
        }
        m.rec_ny = plane ? m.d.ny : 0;
        m.rec_nx = plane ? m.d.nx : 0;
        if (m.nrecv > 0) {
            std::vector<RecvDesc> desc;
            desc.reserve(m.nrecv);
            for (int id : m.shots) {
                const Shot& sh = c->shots[id];
                for (size_t r = 0; r < sh.rec.size(); ++r)
                    desc.push_back({sh.rec[r], sh.rbase + (long)r, sh.slot, (int)sh.rec.size()});
            }
            if ((rc = dev_alloc(c, m.rdesc, sizeof(RecvDesc) * desc.size()))) return rc;
            m.d_rdesc = (RecvDesc*)m.rdesc.p;
            HIPCHK(c, hipMemcpy(m.d_rdesc, desc.data(), sizeof(RecvDesc) * desc.size(), hipMemcpyHostToDevice));
            if ((rc = dev_alloc(c, m.rtr, sizeof(float) * (size_t)rsamples))) return rc;
            m.d_rtr = (float*)m.rtr.p;
            HIPCHK(c, hipMemsetAsync(m.rtr.p, 0, sizeof(float) * (size_t)rsamples, c->stream));
        }
        // Batches of 1 GiB or more per field get the spectrum field S and
        // run passes 1-3 out of place (faster in HBM); smaller ones keep S = X
        // and run them in place, so a pass's working set (one field) can stay
        // in the 256 MB Infinity Cache (a 272^3 field is 161 MB).
        const size_t fbytes = sizeof(float2) * (size_t)m.ncell * nshot;
        const bool oop = fbytes >= ((size_t)1 << 30);
        const size_t nfield = oop ? 5 : 4;
        if ((rc = dev_alloc(c, m.fields, nfield * fbytes))) return rc;
        char* base = (char*)m.fields.p;
        m.X = (float2*)base;
        m.W[0] = (float2*)(base + fbytes);
        m.W[1] = (float2*)(base + 2 * fbytes);
        m.A = (float2*)(base + 3 * fbytes);
        m.S = oop ? (float2*)(base + 4 * fbytes) : m.X;
        HIPCHK(c, hipMemsetAsync(m.fields.p, 0, nfield * fbytes, c->stream));
        const size_t tbytes = sizeof(float) * (size_t)m.nt_cap * m.rec_ny * m.rec_nx * nshot;
        if ((rc = dev_alloc(c, m.traces, tbytes))) return rc;
        m.d_traces = (float*)m.traces.p;
        HIPCHK(c, hipMemsetAsync(m.traces.p, 0, tbytes, c->stream));
        // sources
        std::vector<long> line(nshot);
        std::vector<int> sx(nshot);
        std::vector<float> s((size_t)nshot * m.nt_cap, 0.f), h((size_t)nshot * m.nt_cap, 0.f);
        for (int k = 0; k < nshot; ++k) {
            const Shot& sh = c->shots[m.shots[k]];
            line[k] = ((long)sh.sz * m.d.nyb + sh.sy) * m.d.nxb;
            sx[k] = sh.sx;
            std::copy(sh.s.begin(), sh.s.end(), s.begin() + (size_t)k * m.nt_cap);
            std::copy(sh.h.begin(), sh.h.end(), h.begin() + (size_t)k * m.nt_cap);
        }
        const size_t sb = sizeof(long) * nshot + sizeof(int) * nshot;
        if ((rc = dev_alloc(c, m.src, sb + 64))) return rc;
        m.d_src_line = (long*)m.src.p;
        m.d_src_x = (int*)((char*)m.src.p + sizeof(long) * nshot);
        HIPCHK(c, hipMemcpy(m.d_src_line, line.data(), sizeof(long) * nshot, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(m.d_src_x, sx.data(), sizeof(int) * nshot, hipMemcpyHostToDevice));
        const size_t wb = sizeof(float) * (size_t)nshot * m.nt_cap;
        if ((rc = dev_alloc(c, m.wav, 2 * wb))) return rc;
        m.d_s = (float*)m.wav.p;
        m.d_h = (float*)((char*)m.wav.p + wb);
        if (wb) {
            HIPCHK(c, hipMemcpy(m.d_s, s.data(), wb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(m.d_h, h.data(), wb, hipMemcpyHostToDevice));
        }
        // extended sources: every shot of a model with one goes through the
        // profile kernel (a point becomes the unit profile)
        bool anyprof = false;
        for (int id : m.shots) anyprof |= c->shots[id].prof.wx > 0;
        m.prof_wy = m.prof_wz = 0;
        if (anyprof) {
            const int nxb = m.d.nxb;
            std::vector<ProfDesc> desc(nshot);
            std::vector<float2> PX((size_t)nshot * nxb);
            std::vector<float> pv((size_t)nshot * 3 * PROF_W, 0.f);
            for (int k = 0; k < nshot; ++k) {
                const Shot& sh = c->shots[m.shots[k]];
                ProfDesc d = sh.prof;
                float* v = pv.data() + (size_t)k * 3 * PROF_W;
                if (d.wx == 0) {
                    d = ProfDesc{sh.sx, sh.sy, sh.sz, 1, 1, 1};
                    v[0] = v[PROF_W] = v[2 * PROF_W] = 1.f;
                } else {
                    std::copy(sh.px.begin(), sh.px.end(), v);
                    std::copy(sh.py.begin(), sh.py.end(), v + PROF_W);
                    std::copy(sh.pz.begin(), sh.pz.end(), v + 2 * PROF_W);
                }
                desc[k] = d;
                m.prof_wy = std::max(m.prof_wy, d.wy);
                m.prof_wz = std::max(m.prof_wz, d.wz);
                // PX[kx] = sum_j px[j] exp(-2 pi i kx (x0 + j) / nx), in FP64
                for (int kx = 0; kx < nxb; ++kx) {
                    double re = 0.0, im = 0.0;
                    for (int j = 0; j < d.wx; ++j) {
                        const long ph = ((long)kx * (d.x0 + j)) % nxb;
                        const double a = -2.0 * M_PI * (double)ph / (double)nxb;
                        re += (double)v[j] * std::cos(a);
                        im += (double)v[j] * std::sin(a);
                    }
                    PX[(size_t)k * nxb + kx] = make_float2((float)re, (float)im);
                }
            }
            const size_t db = sizeof(ProfDesc) * nshot, xb = sizeof(float2) * PX.size(),
                         vb = sizeof(float) * pv.size();
            if ((rc = dev_alloc(c, m.pbuf, db + xb + vb + 256))) return rc;
            char* pb = (char*)m.pbuf.p;
            m.d_PX = (float2*)pb;
            m.d_pv = (float*)(pb + xb);
            m.d_pdesc = (ProfDesc*)(pb + xb + vb);
            HIPCHK(c, hipMemcpy(m.d_PX, PX.data(), xb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(m.d_pv, pv.data(), vb, hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(m.d_pdesc, desc.data(), db, hipMemcpyHostToDevice));
        }
        if (m.learned && (rc = alloc_bnet_bufs(c, m, nshot, false))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->autotune)
        for (auto& mp : c->models)
            if ((rc = autotune_model(c, *mp))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->finalized = true;
    return OSE_OK;
}

void drop_graph(ose_ctx* c);

// capture what body enqueues on the main stream into an executable graph
int capture_graph(ose_ctx* c, const std::function<int()>& body, hipGraph_t* g, hipGraphExec_t* x)
{
    *g = nullptr;
    *x = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = body();
    hipGraph_t gg = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &gg);
    if (rc) {
        if (gg) (void)hipGraphDestroy(gg);
        return rc;
    }
    if (e != hipSuccess) return c->fail(OSE_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(x, gg, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(gg);
        *x = nullptr;
        return c->fail(OSE_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *g = gg;
    return OSE_OK;
}

// The step as graphs: one graph, or with two lanes the first lane's head
// and body, the second lane and the tail.  A graph's kernels run one after
// another on the stream it is launched on, so the lanes are separate graphs
// launched on two streams.
int build_graph(ose_ctx* c)
{
    if (c->ngraph) return OSE_OK;
    std::vector<LaneItem> lane[2];
    plan_lanes(c, lane);
    int rc;
    if (lane[1].empty()) {
        rc = capture_graph(c, [&] { int r = enqueue_lane(c, lane[0], 0); return r ? r : enqueue_tail(c); },
                           &c->graph[0], &c->gexec[0]);
        if (!rc) c->ngraph = 1;
    } else {
        rc = capture_graph(c, [&] { return enqueue_lane(c, lane[0], 1); }, &c->graph[0], &c->gexec[0]);
        if (!rc) rc = capture_graph(c, [&] { return enqueue_lane(c, lane[0], 2); }, &c->graph[1], &c->gexec[1]);
        if (!rc) rc = capture_graph(c, [&] { return enqueue_lane(c, lane[1], 0); }, &c->graph[2], &c->gexec[2]);
        if (!rc) rc = capture_graph(c, [&] { return enqueue_tail(c); }, &c->graph[3], &c->gexec[3]);
        if (!rc) c->ngraph = 4;
    }
    if (rc) drop_graph(c);
    return rc;
}

// one step from the built graphs
int launch_graph_step(ose_ctx* c)
{
    if (c->ngraph == 1) {
        HIPCHK(c, hipGraphLaunch(c->gexec[0], c->stream));
        return OSE_OK;
    }
    int rc;
    HIPCHK(c, hipGraphLaunch(c->gexec[0], c->stream));
    if ((rc = fork_side(c))) return rc;
    HIPCHK(c, hipGraphLaunch(c->gexec[1], c->stream));
    const hipError_t e = hipGraphLaunch(c->gexec[2], c->side);
    if ((rc = join_side(c))) return rc;
    if (e != hipSuccess) return c->fail(OSE_EDEVICE, "hipGraphLaunch: %s", hipGetErrorString(e));
    HIPCHK(c, hipGraphLaunch(c->gexec[3], c->stream));
    return OSE_OK;
}

void drop_graph(ose_ctx* c)
{
    for (int i = 0; i < 4; ++i) {
        if (c->gexec[i]) (void)hipGraphExecDestroy(c->gexec[i]);
        if (c->graph[i]) (void)hipGraphDestroy(c->graph[i]);
        c->gexec[i] = nullptr;
        c->graph[i] = nullptr;
    }
    c->ngraph = 0;
}

Shot* get_shot(ose_ctx* c, int id)
{
    if (id < 0 || id >= (int)c->shots.size()) return nullptr;
    return &c->shots[id];
}

// ---- snapshot streaming ---------------------------------------------------

void snap_box(const SnapStream& s, const Model& m, int* o, int* n)
{
    if (s.interior) {
        o[0] = m.d.bl, o[1] = m.d.bf, o[2] = m.d.bu;
        n[0] = m.d.nx, n[1] = m.d.ny, n[2] = m.d.nz;
    } else {
        o[0] = o[1] = o[2] = 0;
        n[0] = m.d.nxb, n[1] = m.d.nyb, n[2] = m.d.nzb;
    }
}

int snap_lanes(int lanes) { return (lanes & 1) + ((lanes >> 1) & 1); }

void snap_fail(SnapStream* s, const char* what, hipError_t e)
{
    if (s->err == OSE_OK) {
        s->err = OSE_EDEVICE;
        s->msg = std::string(what) + ": " + hipGetErrorString(e);
    }
}

void snap_copier(int device, SnapStream* s)
{
    (void)hipSetDevice(device);
    for (;;) {
        std::pair<int, ose_snapshot_info> job;
        int h = -1;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [s] { return s->stop || !s->to_copy.empty(); });
            if (s->to_copy.empty()) break;
            job = s->to_copy.front();
            s->to_copy.pop_front();
            s->cv.wait(lk, [s, &h] {
                for (int k = 0; k < SnapStream::kHost; ++k)
                    if (!s->host_busy[k]) return (h = k), true;
                return false;
            });
            s->host_busy[h] = true;
        }
        const int d = job.first;
        hipError_t e = hipStreamWaitEvent(s->copy, s->ready[d], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->host[h], s->dev[d]->p, job.second.bytes, hipMemcpyDeviceToHost, s->copy);
        if (e == hipSuccess) e = hipEventRecord(s->copied[h], s->copy);
        if (e == hipSuccess) e = hipEventSynchronize(s->copied[h]);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->dev_busy[d] = 0;
            if (e == hipSuccess) {
                s->to_write.emplace_back(h, job.second);
            } else {
                snap_fail(s, "snapshot copy failed", e);
                s->host_busy[h] = false;
                ++s->delivered;
            }
        }
        s->cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->copier_done = true;
    }
    s->cv.notify_all();
}

void snap_writer(SnapStream* s)
{
    for (;;) {
        std::pair<int, ose_snapshot_info> job;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [s] { return s->copier_done || !s->to_write.empty(); });
            if (s->to_write.empty()) return;
            job = s->to_write.front();
            s->to_write.pop_front();
        }
        s->fn(s->user, &job.second, s->host[job.first]);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->host_busy[job.first] = false;
            ++s->delivered;
        }
        s->cv.notify_all();
    }
}

// slots sized for the largest model's snapshot, the copy stream and the threads
int snap_start(ose_ctx* c)
{
    SnapStream& s = *c->snap;
    if (s.started) return OSE_OK;
    size_t need = 256;
    for (auto& mp : c->models) {
        int o[3], n[3];
        snap_box(s, *mp, o, n);
        need = std::max(need, snap_bytes(s.format, snap_lanes(s.lanes), n[0], n[1], n[2]));
    }
    s.slot_bytes = need;
    size_t free_b = 0, total_b = 0;
    HIPCHK(c, hipMemGetInfo(&free_b, &total_b));
    const size_t fit = free_b / 2 / need;
    const int nd = (int)std::max<size_t>(1, std::min<size_t>(c->shots.size(), fit));
    int rc;
    for (int i = 0; i < nd; ++i) {
        s.dev.emplace_back(new DevBuf());
        if ((rc = dev_alloc(c, *s.dev.back(), need))) return rc;
        hipEvent_t e = nullptr;
        HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s.ready.push_back(e);
        s.dev_busy.push_back(0);
    }
    for (int i = 0; i < SnapStream::kHost; ++i) {
        HIPCHK(c, hipHostMalloc(&s.host[i], need, hipHostMallocDefault));
        HIPCHK(c, hipEventCreateWithFlags(&s.copied[i], hipEventDisableTiming));
    }
    HIPCHK(c, hipStreamCreateWithFlags(&s.copy, hipStreamNonBlocking));
    s.copier = std::thread(snap_copier, c->device, &s);
    s.writer = std::thread(snap_writer, &s);
    s.started = true;
    return OSE_OK;
}

int snap_drain(ose_ctx* c)
{
    if (!c->snap || !c->snap->started) return OSE_OK;
    SnapStream& s = *c->snap;
    std::unique_lock<std::mutex> lk(s.mu);
    s.cv.wait(lk, [&s] { return s.delivered == s.enqueued; });
    if (s.err != OSE_OK) return c->fail(s.err, "%s", s.msg.c_str());
    return OSE_OK;
}

// drain, stop the threads and release the slots
int snap_stop(ose_ctx* c)
{
    if (!c->snap) return OSE_OK;
    SnapStream& s = *c->snap;
    int rc = OSE_OK;
    if (s.started) rc = snap_drain(c);
    {
        std::lock_guard<std::mutex> lk(s.mu);
        s.stop = true;
    }
    s.cv.notify_all();
    if (s.copier.joinable()) s.copier.join();
    if (s.writer.joinable()) s.writer.join();
    for (auto e : s.ready)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < SnapStream::kHost; ++i) {
        if (s.host[i]) (void)hipHostFree(s.host[i]);
        if (s.copied[i]) (void)hipEventDestroy(s.copied[i]);
    }
    if (s.copy) (void)hipStreamDestroy(s.copy);
    c->snap.reset();
    return rc;
}

// encode and stream the state of every shot after `step` steps
int snap_enqueue(ose_ctx* c, int step)
{
    SnapStream& s = *c->snap;
    for (auto& mp : c->models) {
        Model& m = *mp;
        for (int id : m.shots) {
            const Shot& sh = c->shots[id];
            const int i = s.next;
            s.next = (i + 1) % (int)s.dev.size();
            {
                std::unique_lock<std::mutex> lk(s.mu);
                s.cv.wait(lk, [&s, i] { return !s.dev_busy[i] || s.err != OSE_OK; });
                if (s.err != OSE_OK) return c->fail(s.err, "%s", s.msg.c_str());
                s.dev_busy[i] = 1;
            }
            int o[3], n[3];
            snap_box(s, m, o, n);
            SnapArgs a{};
            a.u = m.W[0] + (long)sh.slot * m.ncell;
            a.nxb = m.d.nxb;
            a.nyb = m.d.nyb;
            a.x0 = o[0], a.y0 = o[1], a.z0 = o[2];
            a.nx = n[0], a.ny = n[1], a.nz = n[2];
            a.lanes = s.lanes;
            a.format = s.format;
            a.out = s.dev[i]->p;
            const void* fn = s.format == SNAP_F32    ? (const void*)&k_snap<SNAP_F32>
                             : s.format == SNAP_BF16 ? (const void*)&k_snap<SNAP_BF16>
                             : s.format == SNAP_Q16  ? (const void*)&k_snap<SNAP_Q16>
                                                     : (const void*)&k_snap<SNAP_Q8>;
            ose_snapshot_info info{};
            info.shot_id = id;
            info.step = step;
            info.lanes = s.lanes;
            info.format = s.format;
            info.x0 = o[0], info.y0 = o[1], info.z0 = o[2];
            info.nx = n[0], info.ny = n[1], info.nz = n[2];
            info.bytes = snap_bytes(s.format, snap_lanes(s.lanes), n[0], n[1], n[2]);
            int rc = launch(c, -1, 0.0, fn, dim3(n[1], n[2]), dim3(kSnapThreads), 0, &a);
            hipError_t e = hipSuccess;
            if (!rc) e = hipEventRecord(s.ready[i], c->stream);
            if (rc || e != hipSuccess) {
                std::lock_guard<std::mutex> lk(s.mu);
                s.dev_busy[i] = 0;
                return rc ? rc : c->fail(OSE_EDEVICE, "snapshot enqueue failed: %s", hipGetErrorString(e));
            }
            {
                std::lock_guard<std::mutex> lk(s.mu);
                s.to_copy.emplace_back(i, info);
                ++s.enqueued;
            }
            s.cv.notify_all();
        }
    }
    return OSE_OK;
}

}  // namespace

// ---------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------
extern "C" {

int ose_create(ose_ctx** out, int device)
{
    if (!out) return OSE_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return OSE_EDEVICE;
    if (device < 0 || device >= n) return OSE_EINVAL;
    auto* c = new (std::nothrow) ose_ctx();
    if (!c) return OSE_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return OSE_EDEVICE;
    }
    if (dev_alloc(c, c->it_buf, sizeof(int) * 64)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return OSE_ENOMEM;
    }
    c->d_it = (int*)c->it_buf.p;
    if (hipMemset(c->d_it, 0, sizeof(int)) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
        if (c->side) (void)hipStreamDestroy(c->side);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return OSE_EDEVICE;
    }
    *out = c;
    return OSE_OK;
}

int ose_destroy(ose_ctx* c)
{
    if (!c) return OSE_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    (void)snap_stop(c);
    drop_graph(c);
    for (auto& p : c->pending) {
        c->ev_pool.push_back(p.e0);
        c->ev_pool.push_back(p.e1);
    }
    for (auto e : c->ev_pool)
        if (e) (void)hipEventDestroy(e);
    c->models.clear();
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OSE_OK;
}

int ose_set_streams(ose_ctx* c, int n)
{
    if (!c) return OSE_EINVAL;
    if (n != 1 && n != 2) return c->fail(OSE_EINVAL, "ose_set_streams: n must be 1 or 2");
    if (n != c->nstreams) {
        int rc;
        if ((rc = set_device(c))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->nstreams = n;
        drop_graph(c);
    }
    return OSE_OK;
}

const char* ose_last_error(const ose_ctx* c)
{
    if (!c) return "null context";
    return c->err.c_str();
}

int ose_supported_extent(int n) { return next_supported(n); }

int ose_accumulation_plan(int K, int* plan)
{
    if (!plan || K < 2 || K > 64) return OSE_EINVAL;
    const std::vector<AccPlan> p = plan_accumulation(K);
    if ((int)p.size() != K) return OSE_EINVAL;
    for (int j = 0; j < K; ++j) {
        int* o = plan + 6 * j;
        o[0] = p[j].touch;
        o[1] = p[j].add_prev;
        o[2] = p[j].add_next;
        o[3] = p[j].ea;
        o[4] = p[j].eb;
        o[5] = p[j].ec;
    }
    return OSE_OK;
}
int ose_extent_is_supported(int n) { return find_entry(n) != nullptr; }
int ose_extent_kind(int n) { return extent_kind(n); }

int ose_set_fft_path(ose_ctx* c, int path)
{
    if (!c) return OSE_EINVAL;
    if (path != 0 && path != 1) return c->fail(OSE_EINVAL, "ose_set_fft_path: path must be 0 or 1");
    if (!c->models.empty()) return c->fail(OSE_ESTATE, "ose_set_fft_path: models already added");
    c->generic_only = path == 1;
    return OSE_OK;
}

int ose_add_model(ose_ctx* c, const ose_model_desc* d, const float* v, int* model_id)
{
    if (!c) return OSE_EINVAL;
    if (!d || !v || !d->coef || !d->taper_x || !d->taper_y || !d->taper_z)
        return c->fail(OSE_EINVAL, "ose_add_model: null argument");
    if (c->finalized) return c->fail(OSE_ESTATE, "ose_add_model: context already running");
    if (d->nxb <= 0 || d->nyb <= 0 || d->nzb <= 0 || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 ||
        d->bl < 0 || d->bf < 0 || d->bu < 0 || d->bl + d->nx > d->nxb || d->bf + d->ny > d->nyb ||
        d->bu + d->nz > d->nzb)
        return c->fail(OSE_EINVAL, "ose_add_model: inconsistent extents");
    if (d->K < 2 || d->K > 64) return c->fail(OSE_EINVAL, "ose_add_model: K must be in [2, 64]");
    if (!(d->R > 0) || !(d->dx > 0) || !(d->dy > 0) || !(d->dz > 0) || !(d->dt > 0))
        return c->fail(OSE_EINVAL, "ose_add_model: R, dx, dy, dz, dt must be positive");
    if (d->rec_z < 0 || d->rec_z >= d->nz) return c->fail(OSE_EINVAL, "ose_add_model: rec_z outside the interior");
    auto m = std::make_unique<Model>();
    auto entry = [c](int n) { return c->generic_only ? generic_entry(n) : find_entry(n); };
    m->ex = entry(d->nxb);
    m->ey = entry(d->nyb);
    m->ez = entry(d->nzb);
    if (!m->ex || !m->ey || !m->ez)
        return c->fail(OSE_EUNSUPPORTED, "ose_add_model: extent %dx%dx%d outside [2, %d]", d->nxb, d->nyb, d->nzb,
                       GEN_NMAX);
    m->d = *d;
    m->coef.assign(d->coef, d->coef + d->K + 1);
    m->acc_plan = plan_accumulation(d->K);
    if ((int)m->acc_plan.size() != d->K)
        return c->fail(OSE_EINVAL, "ose_add_model: no accumulation plan for K = %d", d->K);
    for (double x : m->coef)
        if (!std::isfinite(x)) return c->fail(OSE_ENUMERIC, "ose_add_model: non-finite coefficient");
    m->gx.assign(d->taper_x, d->taper_x + d->nxb);
    m->gy.assign(d->taper_y, d->taper_y + d->nyb);
    m->gz.assign(d->taper_z, d->taper_z + d->nzb);
    m->d.coef = nullptr;
    m->d.taper_x = m->d.taper_y = m->d.taper_z = nullptr;
    m->ncell = (long)d->nxb * d->nyb * d->nzb;
    int rc;
    if ((rc = set_device(c))) return rc;
    for (const FftEntry* e : {m->ex, m->ey, m->ez})
        if ((rc = prepare_kernels(c, e))) return rc;
    // v / R
    std::vector<float> vr((size_t)m->ncell);
    for (long i = 0; i < m->ncell; ++i) {
        if (!std::isfinite(v[i]) || v[i] <= 0.f)
            return c->fail(OSE_ENUMERIC, "ose_add_model: velocity not finite and positive at cell %ld", i);
        vr[i] = (float)((double)v[i] / d->R);
    }
    if ((rc = dev_alloc(c, m->vr, sizeof(float) * vr.size()))) return rc;
    m->d_vr = (float*)m->vr.p;
    HIPCHK(c, hipMemcpy(m->d_vr, vr.data(), sizeof(float) * vr.size(), hipMemcpyHostToDevice));
    const int nx = d->nxb, ny = d->nyb, nz = d->nzb;
    if ((rc = dev_alloc(c, m->g, sizeof(float) * (nx + ny + nz)))) return rc;
    m->d_gx = (float*)m->g.p;
    m->d_gy = m->d_gx + nx;
    m->d_gz = m->d_gy + ny;
    HIPCHK(c, hipMemcpy(m->d_gx, m->gx.data(), sizeof(float) * nx, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_gy, m->gy.data(), sizeof(float) * ny, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_gz, m->gz.data(), sizeof(float) * nz, hipMemcpyHostToDevice));
    std::vector<float> kx, ky, kz;
    k2table(nx, d->dx, kx);
    k2table(ny, d->dy, ky);
    k2table(nz, d->dz, kz);
    if ((rc = dev_alloc(c, m->k2, sizeof(float) * (nx + ny + nz)))) return rc;
    m->d_kx2 = (float*)m->k2.p;
    m->d_ky2 = m->d_kx2 + nx;
    m->d_kz2 = m->d_ky2 + ny;
    HIPCHK(c, hipMemcpy(m->d_kx2, kx.data(), sizeof(float) * nx, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_ky2, ky.data(), sizeof(float) * ny, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_kz2, kz.data(), sizeof(float) * nz, hipMemcpyHostToDevice));
    std::vector<float2> tx, ty, tz;
    roots(nx, tx);
    roots(ny, ty);
    roots(nz, tz);
    if ((rc = dev_alloc(c, m->tw, sizeof(float2) * 3 * (nx + ny + nz)))) return rc;
    m->d_twx = (float2*)m->tw.p;
    m->d_twy = m->d_twx + 3 * nx;
    m->d_twz = m->d_twy + 3 * ny;
    HIPCHK(c, hipMemcpy(m->d_twx, tx.data(), sizeof(float2) * tx.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_twy, ty.data(), sizeof(float2) * ty.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(m->d_twz, tz.data(), sizeof(float2) * tz.size(), hipMemcpyHostToDevice));
    c->models.push_back(std::move(m));
    if (model_id) *model_id = (int)c->models.size() - 1;
    return OSE_OK;
}

int ose_add_shot(ose_ctx* c, int model_id, int sx, int sy, int sz, const float* s, const float* h,
                 int nt, int* shot_id)
{
    if (!c) return OSE_EINVAL;
    if (c->finalized) return c->fail(OSE_ESTATE, "ose_add_shot: context already running");
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_add_shot: bad model id");
    Model& m = *c->models[model_id];
    if (sx < 0 || sx >= m.d.nxb || sy < 0 || sy >= m.d.nyb || sz < 0 || sz >= m.d.nzb)
        return c->fail(OSE_EINVAL, "ose_add_shot: source outside the grid");
    if (nt < 0 || (nt > 0 && (!s || !h))) return c->fail(OSE_EINVAL, "ose_add_shot: bad wavelet");
    Shot sh;
    sh.model = model_id;
    sh.slot = (int)m.shots.size();
    sh.sx = sx;
    sh.sy = sy;
    sh.sz = sz;
    sh.nt = nt;
    if (nt > 0) {
        sh.s.assign(s, s + nt);
        sh.h.assign(h, h + nt);
        for (int i = 0; i < nt; ++i)
            if (!std::isfinite(s[i]) || !std::isfinite(h[i]))
                return c->fail(OSE_ENUMERIC, "ose_add_shot: non-finite wavelet sample %d", i);
    }
    c->shots.push_back(sh);
    m.shots.push_back((int)c->shots.size() - 1);
    if (shot_id) *shot_id = (int)c->shots.size() - 1;
    return OSE_OK;
}

int ose_set_receivers(ose_ctx* c, int shot_id, const int* xyz, int n)
{
    if (!c) return OSE_EINVAL;
    if (c->finalized) return c->fail(OSE_ESTATE, "ose_set_receivers: context already running");
    Shot* sh = get_shot(c, shot_id);
    if (!sh || n < 0 || (n > 0 && !xyz)) return c->fail(OSE_EINVAL, "ose_set_receivers: bad shot or receiver list");
    const Model& m = *c->models[sh->model];
    std::vector<long> rec((size_t)n);
    for (int r = 0; r < n; ++r) {
        const int x = xyz[3 * r], y = xyz[3 * r + 1], z = xyz[3 * r + 2];
        if (x < 0 || x >= m.d.nxb || y < 0 || y >= m.d.nyb || z < 0 || z >= m.d.nzb)
            return c->fail(OSE_EINVAL, "ose_set_receivers: receiver %d at (%d, %d, %d) is outside the %dx%dx%d grid", r,
                           x, y, z, m.d.nxb, m.d.nyb, m.d.nzb);
        rec[(size_t)r] = ((long)z * m.d.nyb + y) * m.d.nxb + x;
    }
    sh->rec.swap(rec);
    return OSE_OK;
}

int ose_set_source_profile(ose_ctx* c, int shot_id, const int* box, const float* px, const float* py,
                           const float* pz)
{
    if (!c) return OSE_EINVAL;
    if (c->finalized) return c->fail(OSE_ESTATE, "ose_set_source_profile: context already running");
    Shot* sh = get_shot(c, shot_id);
    if (!sh) return c->fail(OSE_EINVAL, "ose_set_source_profile: bad shot");
    if (!box) {
        sh->prof = ProfDesc{0, 0, 0, 0, 0, 0};
        sh->px.clear(), sh->py.clear(), sh->pz.clear();
        return OSE_OK;
    }
    const Model& m = *c->models[sh->model];
    const int x0 = box[0], y0 = box[1], z0 = box[2], wx = box[3], wy = box[4], wz = box[5];
    const int n[3] = {m.d.nxb, m.d.nyb, m.d.nzb}, o[3] = {x0, y0, z0}, w[3] = {wx, wy, wz};
    for (int a = 0; a < 3; ++a)
        if (w[a] < 1 || w[a] > PROF_W || o[a] < 0 || o[a] + w[a] > n[a])
            return c->fail(OSE_EINVAL, "ose_set_source_profile: support %d+%d on axis %d outside [0, %d) or wider than %d",
                           o[a], w[a], a, n[a], PROF_W);
    if (!px || !py || !pz) return c->fail(OSE_EINVAL, "ose_set_source_profile: null profile");
    for (int a = 0; a < 3; ++a) {
        const float* p = a == 0 ? px : a == 1 ? py : pz;
        for (int j = 0; j < w[a]; ++j)
            if (!std::isfinite(p[j])) return c->fail(OSE_ENUMERIC, "ose_set_source_profile: non-finite value");
    }
    sh->prof = ProfDesc{x0, y0, z0, wx, wy, wz};
    sh->px.assign(px, px + wx);
    sh->py.assign(py, py + wy);
    sh->pz.assign(pz, pz + wz);
    return OSE_OK;
}

int ose_receiver_count(ose_ctx* c, int shot_id, int* nrec)
{
    if (!c) return OSE_EINVAL;
    const Shot* sh = get_shot(c, shot_id);
    if (!sh || !nrec) return c->fail(OSE_EINVAL, "ose_receiver_count: bad shot or null output");
    const Model& m = *c->models[sh->model];
    *nrec = sh->rec.empty() ? m.d.nx * m.d.ny : (int)sh->rec.size();
    return OSE_OK;
}

int ose_set_options(ose_ctx* c, int use_graph, int profile)
{
    if (!c) return OSE_EINVAL;
    c->use_graph = use_graph != 0;
    c->profile = profile != 0;
    return OSE_OK;
}

int ose_run(ose_ctx* c, int nsteps)
{
    if (!c) return OSE_EINVAL;
    if (nsteps < 0) return c->fail(OSE_EINVAL, "ose_run: negative step count");
    if (c->shots.empty()) return c->fail(OSE_ESTATE, "ose_run: no shots");
    int rc;
    if ((rc = set_device(c))) return rc;
    if ((rc = finalize(c))) return rc;
    // a tripped boundary-network guard switches the model back to the plain taper
    for (auto& mp : c->models) {
        if (!mp->learned_on) continue;
        int tr = 0;
        HIPCHK(c, hipMemcpy(&tr, mp->bb.tripped, sizeof(int), hipMemcpyDeviceToHost));
        if (tr) {
            mp->learned_on = false;
            drop_graph(c);
        }
    }
    const bool graph = c->use_graph && !c->profile;
    if (graph && (rc = build_graph(c))) {
        // capture is not available: run eagerly
        (void)hipGetLastError();
        c->use_graph = false;
        drop_graph(c);
    }
    const bool snaps = c->snap && c->snap->every > 0;
    if (snaps && (rc = snap_start(c))) return rc;
    for (int s = 0; s < nsteps; ++s) {
        if (c->use_graph && !c->profile && c->ngraph) {
            if ((rc = launch_graph_step(c))) return rc;
        } else if ((rc = enqueue_step(c))) {
            return rc;
        }
        ++c->host_it;
        if (snaps && c->host_it % c->snap->every == 0 && (rc = snap_enqueue(c, c->host_it))) return rc;
    }
    return OSE_OK;
}

int ose_sync(ose_ctx* c)
{
    if (!c) return OSE_EINVAL;
    int rc;
    if ((rc = set_device(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = snap_drain(c))) return rc;
    return collect(c);
}

int ose_set_snapshots(ose_ctx* c, int every, int lanes, int format, int interior, ose_snapshot_fn fn, void* user)
{
    if (!c) return OSE_EINVAL;
    if (every < 0 || (every > 0 && (!fn || lanes < 1 || lanes > 3 || format < OSE_SNAP_F32 || format > OSE_SNAP_Q8)))
        return c->fail(OSE_EINVAL, "ose_set_snapshots: bad cadence %d, lanes %d, format %d or callback", every, lanes,
                       format);
    int rc;
    if ((rc = set_device(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = snap_stop(c))) return rc;
    if (every == 0) return OSE_OK;
    c->snap.reset(new (std::nothrow) SnapStream());
    if (!c->snap) return c->fail(OSE_ENOMEM, "ose_set_snapshots: out of memory");
    SnapStream& s = *c->snap;
    s.every = every;
    s.lanes = lanes;
    s.format = format;
    s.interior = interior != 0;
    s.fn = fn;
    s.user = user;
    return OSE_OK;
}

int ose_snapshot_drain(ose_ctx* c)
{
    if (!c) return OSE_EINVAL;
    return snap_drain(c);
}

int ose_timed_run(ose_ctx* c, int nsteps, double* ms)
{
    if (!c) return OSE_EINVAL;
    int rc;
    if ((rc = set_device(c))) return rc;
    if ((rc = finalize(c))) return rc;
    if (c->use_graph && !c->profile && (rc = build_graph(c))) {
        (void)hipGetLastError();
        c->use_graph = false;
        drop_graph(c);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    EventPair ev;
    HIPCHK(c, ev.create());
    HIPCHK(c, hipEventRecord(ev.e0, c->stream));
    rc = ose_run(c, nsteps);
    if (rc == OSE_OK) {
        HIPCHK(c, hipEventRecord(ev.e1, c->stream));
        HIPCHK(c, hipEventSynchronize(ev.e1));
        float f = 0.f;
        HIPCHK(c, hipEventElapsedTime(&f, ev.e0, ev.e1));
        if (ms) *ms = f;
    }
    if (rc) return rc;
    return collect(c);
}

int ose_pattern_probe(ose_ctx* c, int model_id, int axis, int reps, double* gbs)
{
    if (!c || !gbs || reps <= 0) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_pattern_probe: bad model");
    if (axis != 1 && axis != 2) return c->fail(OSE_EINVAL, "ose_pattern_probe: axis must be 1 (y) or 2 (z)");
    int rc;
    if ((rc = set_device(c)) || (rc = finalize(c)) || (rc = ose_sync(c))) return rc;
    Model& m = *c->models[model_id];
    const int nshot = (int)m.shots.size();
    if (nshot == 0) return c->fail(OSE_ESTATE, "ose_pattern_probe: model has no shots");
    const bool zaxis = axis == 2;
    const FftEntry* e = zaxis ? m.ez : m.ey;
    int sel = zaxis ? m.sel_z : m.sel_yf;
    if (zaxis && m.ez->zs[sel].chunk > 0) {
        // a chunked z pass: its one-tile counterpart of the same width (split exchange)
        for (int u = 0; u < m.ez->ns; ++u)
            if (m.ez->zs[u].chunk == 0 && m.ez->zs[u].T == m.ez->zs[sel].T && m.ez->zs[u].split) sel = u;
        if (m.ez->zs[sel].chunk > 0) sel = 0;
    }
    const StridedKern& k = zaxis ? m.ez->zs[sel] : m.ey->yf[sel];
    if (e->N1 <= 0 || e->N1 > 32 || e->N2 > 32 || (k.T != 8 && k.T != 16 && k.T != 32))
        return c->fail(OSE_EUNSUPPORTED, "ose_pattern_probe: extent %d has no compiled split", e->N);
    const int nx = m.d.nxb, ny = m.d.nyb, nz = m.d.nzb;
    // the pass's own grid, tile order and occupancy (its LDS request); field
    // X (the next step's x spectrum, read) to S (scratch between steps,
    // written), so the state survives between steps
    StridedArgs a{};
    a.in = m.X;
    a.out = m.S;
    a.ncol = nx;
    a.es = zaxis ? (long)nx * ny : (long)nx;
    a.os = zaxis ? (long)nx : (long)nx * ny;
    a.shot_stride = m.ncell;
    const dim3 grid((nx + k.T - 1) / k.T, zaxis ? ny : nz, nshot);
    a.xcd = (double)grid.x * grid.y * grid.z < 16777216.0 && grid.x < 65536 && grid.y < 65536;
    a.mgx = (1ULL << 40) / grid.x + 1;
    a.mgy = (1ULL << 40) / grid.y + 1;
    int n1 = e->N1, n2 = e->N2;
    const void* fn = k.T == 8 ? (const void*)&k_strided_copy<8>
                   : k.T == 16 ? (const void*)&k_strided_copy<16> : (const void*)&k_strided_copy<32>;
    const dim3 block(k.T * std::max(n1, n2));
    if (k.lds > 65536) HIPCHK(c, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)k.lds));
    void* args[] = {&a, &n1, &n2};
    HIPCHK(c, hipLaunchKernel(fn, grid, block, args, k.lds, c->stream));   // warm-up
    EventPair ev;
    HIPCHK(c, ev.create());
    HIPCHK(c, hipEventRecord(ev.e0, c->stream));
    for (int r = 0; r < reps; ++r) HIPCHK(c, hipLaunchKernel(fn, grid, block, args, k.lds, c->stream));
    HIPCHK(c, hipEventRecord(ev.e1, c->stream));
    HIPCHK(c, hipEventSynchronize(ev.e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, ev.e0, ev.e1));
    *gbs = 16.0 * (double)m.ncell * nshot * reps / (ms * 1e-3) / 1e9;
    return OSE_OK;
}

int ose_copy_probe(ose_ctx* c, size_t bytes, int reps, double* gbs)
{
    if (!c || !gbs || reps <= 0 || bytes < 4096) return OSE_EINVAL;
    int rc;
    if ((rc = set_device(c)) || (rc = ose_sync(c))) return rc;
    DevBuf a, b;
    bytes &= ~(size_t)15;
    if ((rc = dev_alloc(c, a, bytes)) || (rc = dev_alloc(c, b, bytes))) return rc;
    HIPCHK(c, hipMemsetAsync(a.p, 0, bytes, c->stream));
    const long n = (long)(bytes / 16);
    const dim3 grid((unsigned)((n + 1023) / 1024)), block(256);
    const f32x4* src = (const f32x4*)a.p;
    f32x4* dst = (f32x4*)b.p;
    void* args[] = {(void*)&src, (void*)&dst, (void*)&n};
    HIPCHK(c, hipLaunchKernel((const void*)&k_copy, grid, block, args, 0, c->stream));
    EventPair ev;
    HIPCHK(c, ev.create());
    HIPCHK(c, hipEventRecord(ev.e0, c->stream));
    for (int r = 0; r < reps; ++r) HIPCHK(c, hipLaunchKernel((const void*)&k_copy, grid, block, args, 0, c->stream));
    HIPCHK(c, hipEventRecord(ev.e1, c->stream));
    HIPCHK(c, hipEventSynchronize(ev.e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, ev.e0, ev.e1));
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return OSE_OK;
}

int ose_pass_probe(ose_ctx* c, int model_id, int which, int reps, double* ms)
{
    if (!c || !ms || reps <= 0) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_pass_probe: bad model");
    int rc;
    if ((rc = set_device(c)) || (rc = finalize(c)) || (rc = ose_sync(c))) return rc;
    Model& m = *c->models[model_id];
    const int nshot = (int)m.shots.size();
    if (nshot == 0) return c->fail(OSE_ESTATE, "ose_pass_probe: model has no shots");
    const int nx = m.d.nxb, ny = m.d.nyb, nz = m.d.nzb;
    // the probe works on W[1] and A only, which hold nothing between steps
    // (ORDER0 writes both before reading them), so a state being stepped is
    // not disturbed; the x probe keeps its spectrum and W in W[1] (same
    // traffic, different lines of the buffer are never shared across threads)
    std::function<int()> body;
    XArgs a = base_xargs(m, nshot);
    a.S = m.W[1];
    a.Xout = m.W[1];
    a.A = m.A;
    a.Win = a.Wout = m.W[1];
    a.c1 = 0.5f;
    float2* F = m.W[1];
    switch (which) {
    case 0: body = [&] { return launch_strided(c, -1, m.ey->yf[m.sel_yf], F, F, nx, ny, nz, false, nshot, m.ncell, m.d_twy, nullptr, 0.0); }; break;
    case 1: {
        // forward-z alone in the z symbol's shape (a chunked z symbol has no
        // forward-z kernel: the T = 16 split one stands in)
        int v = m.sel_z;
        for (int u = 0; !m.ez->yf[v].fn && u < m.ez->ns; ++u)
            if (m.ez->yf[u].fn && m.ez->yf[u].T == m.ez->zs[m.sel_z].T) v = u;
        if (!m.ez->yf[v].fn) v = 0;
        body = [&, v] { return launch_strided(c, -1, m.ez->yf[v], F, F, nx, ny, nz, true, nshot, m.ncell, m.d_twz, nullptr, 0.0); };
        break;
    }
    case 2: body = [&] { return launch_strided(c, -1, m.ez->zs[m.sel_z], F, F, nx, ny, nz, true, nshot, m.ncell, m.d_twz, &m, 0.0); }; break;
    case 3: body = [&] { return launch_strided(c, -1, m.ey->yi[m.sel_yi], F, F, nx, ny, nz, false, nshot, m.ncell, m.d_twy, nullptr, 0.0); }; break;
    case 4: body = [&] { return launch_x(c, -1, m.ex->xk[m.sel_x], XM_MID, a, ny * nz, nshot, 0.0); }; break;
    case 5: body = [&] { return launch_x(c, -1, m.ex->xk[m.sel_x], XM_FWD, a, ny * nz, nshot, 0.0); }; break;
    default: return c->fail(OSE_EINVAL, "ose_pass_probe: unknown pass %d", which);
    }
    float t = 0.f;
    const bool prof = c->profile;
    c->profile = false;
    rc = time_variant(c, body, reps, &t);
    c->profile = prof;
    if (rc) return rc;
    *ms = t;
    return OSE_OK;
}

int ose_set_autotune(ose_ctx* c, int on)
{
    if (!c) return OSE_EINVAL;
    if (c->finalized) return c->fail(OSE_ESTATE, "ose_set_autotune: context already running");
    c->autotune = on != 0;
    return OSE_OK;
}

// x-pass launch-shape code: P + 1000 twb, negative for the split exchange
static int xcode(const XKern& k)
{
    const int v = k.P + 1000 * k.twb;
    return k.split ? -v : v;
}

int ose_get_tuning(ose_ctx* c, int model_id, int* shape)
{
    if (!c || !shape) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_get_tuning: bad model");
    const Model& m = *c->models[model_id];
    shape[0] = scode(m.ey->yf[m.sel_yf]);
    shape[1] = scode(m.ez->zs[m.sel_z]);
    shape[2] = scode(m.ey->yi[m.sel_yi]);
    shape[3] = xcode(m.ex->xk[m.sel_x]);
    return OSE_OK;
}

int ose_set_tuning(ose_ctx* c, int model_id, const int* shape)
{
    if (!c || !shape) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_set_tuning: bad model");
    Model& m = *c->models[model_id];
    auto find_strided = [&](const FftEntry* e, int which, int want) {
        for (int v = 0; v < e->ns; ++v) {
            const StridedKern& k = which == 0 ? e->yf[v] : which == 1 ? e->zs[v] : e->yi[v];
            if (k.fn && scode(k) == want) return v;
        }
        return -1;
    };
    const int yf = find_strided(m.ey, 0, shape[0]);
    const int z = find_strided(m.ez, 1, shape[1]);
    const int yi = find_strided(m.ey, 2, shape[2]);
    int x = -1;
    for (int v = 0; v < m.ex->nx; ++v)
        if (xcode(m.ex->xk[v]) == shape[3]) x = v;
    if (yf < 0 || z < 0 || yi < 0 || x < 0)
        return c->fail(OSE_EINVAL, "ose_set_tuning: launch shape {%d, %d, %d, %d} is not compiled for %dx%dx%d",
                       shape[0], shape[1], shape[2], shape[3], m.d.nxb, m.d.nyb, m.d.nzb);
    m.sel_yf = yf;
    m.sel_z = z;
    m.sel_yi = yi;
    m.sel_x = x;
    for (int& k : m.sel_mode) k = x;
    m.tuned = true;
    drop_graph(c);
    return OSE_OK;
}

// x-pass modes in the order of ose_get_tuning_modes' entries 3..7 (the first
// order's entry covers ORDER0 and ORDER0N, the last order's LAST and LASTA)
static const int kTuneModes[5] = {XM_MID, XM_SKIP, XM_MIDF, XM_ORDER0N, XM_LAST};

int ose_get_tuning_modes(ose_ctx* c, int model_id, int* shape)
{
    if (!c || !shape) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_get_tuning_modes: bad model");
    const Model& m = *c->models[model_id];
    int rc;
    if ((rc = ose_get_tuning(c, model_id, shape))) return rc;
    for (int i = 0; i < 5; ++i) shape[3 + i] = xcode(m.ex->xk[m.sel_mode[kTuneModes[i]]]);
    // ORDER0 and ORDER0N share the first order's entry; the plan uses one of them
    if (!m.acc_plan.empty() && m.acc_plan[0].touch) shape[6] = xcode(m.ex->xk[m.sel_mode[XM_ORDER0]]);
    return OSE_OK;
}

int ose_set_tuning_modes(ose_ctx* c, int model_id, const int* shape)
{
    if (!c || !shape) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_set_tuning_modes: bad model");
    Model& m = *c->models[model_id];
    int xs[5];
    for (int i = 0; i < 5; ++i) {
        xs[i] = -1;
        for (int v = 0; v < m.ex->nx; ++v)
            if (xcode(m.ex->xk[v]) == shape[3 + i]) xs[i] = v;
        if (xs[i] < 0)
            return c->fail(OSE_EINVAL, "ose_set_tuning_modes: x-pass shape %d is not compiled for extent %d",
                           shape[3 + i], m.d.nxb);
    }
    const int base[4] = {shape[0], shape[1], shape[2], shape[3]};
    int rc;
    if ((rc = ose_set_tuning(c, model_id, base))) return rc;
    for (int i = 0; i < 5; ++i) m.sel_mode[kTuneModes[i]] = xs[i];
    m.sel_mode[XM_ORDER0] = xs[3];
    m.sel_mode[XM_LASTA] = xs[4];
    m.sel_x = xs[0];
    return OSE_OK;
}

int ose_launch_shapes(ose_ctx* c, int model_id, int* strided, int* nstrided, int* xshapes, int* nxs)
{
    if (!c || !nstrided || !nxs) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_launch_shapes: bad model");
    const Model& m = *c->models[model_id];
    // strided shapes common to the y and z extents
    int n = 0;
    for (int v = 0; v < m.ey->ns; ++v) {
        if (!m.ey->yf[v].fn) continue;   // a chunked z-pass slot
        const int code = scode(m.ey->yf[v]);
        bool inz = false;
        for (int u = 0; u < m.ez->ns; ++u) inz |= scode(m.ez->zs[u]) == code;
        if (inz) {
            if (strided && n < *nstrided) strided[n] = code;
            ++n;
        }
    }
    *nstrided = n;
    n = 0;
    for (int v = 0; v < m.ex->nx; ++v) {
        if (xshapes && n < *nxs) xshapes[n] = xcode(m.ex->xk[v]);
        ++n;
    }
    *nxs = n;
    return OSE_OK;
}

int ose_pass_stats(ose_ctx* c, int pass, double* avg_ms, long* launches, double* bytes)
{
    if (!c || pass < 0 || pass >= PK_N) return OSE_EINVAL;
    if (avg_ms) *avg_ms = c->st_n[pass] ? c->st_ms[pass] / c->st_n[pass] : 0.0;
    if (launches) *launches = c->st_n[pass];
    if (bytes) *bytes = c->st_n[pass] ? c->st_bytes[pass] / c->st_n[pass] : 0.0;
    return OSE_OK;
}

int ose_reset_stats(ose_ctx* c)
{
    if (!c) return OSE_EINVAL;
    int rc = ose_sync(c);
    for (int i = 0; i < PK_N; ++i) c->st_ms[i] = 0, c->st_n[i] = 0, c->st_bytes[i] = 0;
    return rc;
}

int ose_get_step(ose_ctx* c, int* it)
{
    if (!c || !it) return OSE_EINVAL;
    int rc;
    if ((rc = ose_sync(c))) return rc;
    HIPCHK(c, hipMemcpy(it, c->d_it, sizeof(int), hipMemcpyDeviceToHost));
    return OSE_OK;
}

int ose_set_step(ose_ctx* c, int it)
{
    if (!c || it < 0) return OSE_EINVAL;
    int rc;
    if ((rc = ose_sync(c))) return rc;
    HIPCHK(c, hipMemcpy(c->d_it, &it, sizeof(int), hipMemcpyHostToDevice));
    c->host_it = it;
    return OSE_OK;
}

int ose_get_traces(ose_ctx* c, int shot_id, int it0, int nsteps, float* out)
{
    if (!c) return OSE_EINVAL;
    Shot* sh = get_shot(c, shot_id);
    if (!sh || !out) return c->fail(OSE_EINVAL, "ose_get_traces: bad shot or null output");
    int rc;
    if ((rc = finalize(c))) return rc;
    Model& m = *c->models[sh->model];
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > m.nt_cap)
        return c->fail(OSE_EINVAL, "ose_get_traces: steps [%d,%d) outside the trace buffer (%d)", it0,
                       it0 + nsteps, m.nt_cap);
    if ((rc = ose_sync(c))) return rc;
    const long nrec = sh->rec.empty() ? (long)m.rec_ny * m.rec_nx : (long)sh->rec.size();
    std::vector<float> buf((size_t)nrec * nsteps);
    const float* src = sh->rec.empty() ? m.d_traces + (long)sh->slot * m.nt_cap * nrec + (long)it0 * nrec
                                       : m.d_rtr + sh->rbase + (long)it0 * nrec;
    if (nsteps) HIPCHK(c, hipMemcpy(buf.data(), src, sizeof(float) * buf.size(), hipMemcpyDeviceToHost));
    for (int it = 0; it < nsteps; ++it)
        for (long r = 0; r < nrec; ++r) out[r * nsteps + it] = buf[(size_t)it * nrec + r];
    return OSE_OK;
}

int ose_set_traces(ose_ctx* c, int shot_id, int it0, int nsteps, const float* in)
{
    if (!c) return OSE_EINVAL;
    Shot* sh = get_shot(c, shot_id);
    if (!sh || !in) return c->fail(OSE_EINVAL, "ose_set_traces: bad shot or null input");
    int rc;
    if ((rc = set_device(c)) || (rc = finalize(c))) return rc;
    Model& m = *c->models[sh->model];
    if (it0 < 0 || nsteps < 0 || it0 + nsteps > m.nt_cap)
        return c->fail(OSE_EINVAL, "ose_set_traces: steps [%d,%d) outside the trace buffer (%d)", it0,
                       it0 + nsteps, m.nt_cap);
    if ((rc = ose_sync(c))) return rc;
    const long nrec = sh->rec.empty() ? (long)m.rec_ny * m.rec_nx : (long)sh->rec.size();
    std::vector<float> buf((size_t)nrec * nsteps);
    for (int it = 0; it < nsteps; ++it)
        for (long r = 0; r < nrec; ++r) {
            const float v = in[r * nsteps + it];
            if (!std::isfinite(v)) return c->fail(OSE_ENUMERIC, "ose_set_traces: non-finite sample");
            buf[(size_t)it * nrec + r] = v;
        }
    float* dst = sh->rec.empty() ? m.d_traces + (long)sh->slot * m.nt_cap * nrec + (long)it0 * nrec
                                 : m.d_rtr + sh->rbase + (long)it0 * nrec;
    if (nsteps) HIPCHK(c, hipMemcpy(dst, buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
    return OSE_OK;
}

int ose_get_wavefield(ose_ctx* c, int shot_id, float* re, float* im)
{
    if (!c) return OSE_EINVAL;
    Shot* sh = get_shot(c, shot_id);
    if (!sh || !re || !im) return c->fail(OSE_EINVAL, "ose_get_wavefield: bad shot or null output");
    int rc;
    if ((rc = set_device(c)) || (rc = finalize(c)) || (rc = ose_sync(c))) return rc;
    Model& m = *c->models[sh->model];
    // the state u is kept in physical form in W[0]
    std::vector<float2> buf((size_t)m.ncell);
    HIPCHK(c, hipMemcpy(buf.data(), m.W[0] + (long)sh->slot * m.ncell, sizeof(float2) * buf.size(),
                        hipMemcpyDeviceToHost));
    for (long i = 0; i < m.ncell; ++i) {
        re[i] = buf[i].x;
        im[i] = buf[i].y;
    }
    return OSE_OK;
}

int ose_set_wavefield(ose_ctx* c, int shot_id, const float* re, const float* im)
{
    if (!c) return OSE_EINVAL;
    Shot* sh = get_shot(c, shot_id);
    if (!sh || !re || !im) return c->fail(OSE_EINVAL, "ose_set_wavefield: bad shot or null input");
    int rc;
    if ((rc = set_device(c)) || (rc = finalize(c)) || (rc = ose_sync(c))) return rc;
    Model& m = *c->models[sh->model];
    std::vector<float2> buf((size_t)m.ncell);
    for (long i = 0; i < m.ncell; ++i) {
        if (!std::isfinite(re[i]) || !std::isfinite(im[i]))
            return c->fail(OSE_ENUMERIC, "ose_set_wavefield: non-finite value at cell %ld", i);
        buf[i] = make_float2(re[i], im[i]);
    }
    XArgs a = base_xargs(m, 1);
    a.A = m.W[0] + (long)sh->slot * m.ncell;   // u, and X_0 = FFT_x(u)
    a.Xout = m.X + (long)sh->slot * m.ncell;
    HIPCHK(c, hipMemcpy(a.A, buf.data(), sizeof(float2) * buf.size(), hipMemcpyHostToDevice));
    const bool prof = c->profile;
    c->profile = false;
    rc = launch_x(c, 0, m.ex->xk[m.sel_x], XM_FWD, a, m.d.nyb * m.d.nzb, 1, 0.0);
    c->profile = prof;
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OSE_OK;
}

}  // extern "C"

namespace {

// The operator application on two caller-provided device work fields of
// m.ncell complex cells each (A: the field and the result, X: its spectrum).
int apply_operator_on(ose_ctx* c, Model& m, const float* in, float* out, float2* A, float2* X)
{
    int rc;
    const int nx = m.d.nxb, ny = m.d.nyb, nz = m.d.nzb;
    const size_t fb = sizeof(float2) * (size_t)m.ncell;
    HIPCHK(c, hipMemcpy(A, in, fb, hipMemcpyHostToDevice));
    const bool prof = c->profile;
    c->profile = false;
    XArgs a = base_xargs(m, 1);
    a.A = A;
    a.Xout = X;
    rc = launch_x(c, 0, m.ex->xk[m.sel_x], XM_FWD, a, ny * nz, 1, 0.0);
    if (!rc) rc = launch_strided(c, 0, m.ey->yf[m.sel_yf], X, X, nx, ny, nz, false, 1, m.ncell, m.d_twy, nullptr, 0);
    if (!rc) rc = launch_strided(c, 0, m.ez->zs[m.sel_z], X, X, nx, ny, nz, true, 1, m.ncell, m.d_twz, &m, 0);
    if (!rc) rc = launch_strided(c, 0, m.ey->yi[m.sel_yi], X, X, nx, ny, nz, false, 1, m.ncell, m.d_twy, nullptr, 0);
    if (!rc) {
        XArgs b = base_xargs(m, 1);
        b.Xprev = X;
        b.A = A;
        rc = launch_x(c, 0, m.ex->xk[m.sel_x], XM_INV, b, ny * nz, 1, 0.0);
    }
    if (!rc) {
        void* args[] = {&A, &m.d_vr, &m.ncell, &m.ncell, nullptr};
        float s = (float)nx;   // XM_INV divided by nx; the operator's 1/N sits in the symbol
        args[4] = &s;
        hipError_t e = hipLaunchKernel((const void*)&k_scale_vr, dim3((unsigned)((m.ncell + 255) / 256), 1),
                                       dim3(256), args, 0, c->stream);
        if (e != hipSuccess) rc = c->fail(OSE_EDEVICE, "k_scale_vr launch: %s", hipGetErrorString(e));
    }
    c->profile = prof;
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, A, fb, hipMemcpyDeviceToHost));
    return OSE_OK;
}

// Scratch of ose_apply_operator_noalloc: its two work fields, each of at most
// kNoAllocCells complex cells (64 MiB in all).  Every GPU the library runs on
// holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel or a copy (on the host it is only a stand-in for
// the symbol): it uses the device address hipGetSymbolAddress gives on each
// device, taken once per device by noalloc_scratch.
constexpr long kNoAllocCells = 1L << 22;
__device__ __align__(256) float2 scratch[2 * kNoAllocCells];

std::mutex g_scratch_mu;
float2* g_scratch_ptr[64] = {};

// the device address of this device's scratch (looked up once per device)
int noalloc_scratch(ose_ctx* c, float2** ptr)
{
    if (c->device < 0 || c->device >= 64) return c->fail(OSE_EDEVICE, "no scratch for device %d", c->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return c->fail(OSE_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[c->device] = (float2*)p;
    }
    *ptr = g_scratch_ptr[c->device];
    return OSE_OK;
}

}  // namespace

extern "C" {

int ose_apply_operator(ose_ctx* c, int model_id, const float* in, float* out)
{
    if (!c) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size() || !in || !out)
        return c->fail(OSE_EINVAL, "ose_apply_operator: bad model or null pointer");
    int rc;
    if ((rc = set_device(c)) || (rc = ose_sync(c))) return rc;
    Model& m = *c->models[model_id];
    DevBuf tmp;
    const size_t fb = sizeof(float2) * (size_t)m.ncell;
    if ((rc = dev_alloc(c, tmp, 2 * fb))) return rc;
    return apply_operator_on(c, m, in, out, (float2*)tmp.p, (float2*)((char*)tmp.p + fb));
}

int ose_apply_operator_noalloc(ose_ctx* c, int model_id, const float* in, float* out)
{
    if (!c) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size() || !in || !out)
        return c->fail(OSE_EINVAL, "ose_apply_operator_noalloc: bad model or null pointer");
    Model& m = *c->models[model_id];
    if (m.ncell > kNoAllocCells)
        return c->fail(OSE_EINVAL, "ose_apply_operator_noalloc: %ld cells exceed the scratch limit of %ld cells",
                       m.ncell, kNoAllocCells);
    int rc;
    if ((rc = set_device(c)) || (rc = ose_sync(c))) return rc;
    float2* s = nullptr;
    if ((rc = noalloc_scratch(c, &s))) return rc;
    return apply_operator_on(c, m, in, out, s, s + kNoAllocCells);
}

int ose_fft3d(ose_ctx* c, int nx, int ny, int nz, const float* in, float* out, int sign)
{
    if (!c) return OSE_EINVAL;
    if (!in || !out || (sign != 1 && sign != -1)) return c->fail(OSE_EINVAL, "ose_fft3d: bad argument");
    auto entry = [c](int n) { return c->generic_only ? generic_entry(n) : find_entry(n); };
    const FftEntry *ex = entry(nx), *ey = entry(ny), *ez = entry(nz);
    if (!ex || !ey || !ez) return c->fail(OSE_EUNSUPPORTED, "ose_fft3d: extent outside [2, %d]", GEN_NMAX);
    int rc;
    if ((rc = set_device(c)) || (rc = ose_sync(c))) return rc;
    for (const FftEntry* e : {ex, ey, ez})
        if ((rc = prepare_kernels(c, e))) return rc;
    const long n = (long)nx * ny * nz;
    const size_t fb = sizeof(float2) * (size_t)n;
    std::vector<float2> tx, ty, tz;
    roots(nx, tx);
    roots(ny, ty);
    roots(nz, tz);
    DevBuf tmp;
    if ((rc = dev_alloc(c, tmp, 2 * fb + sizeof(float2) * 3 * (nx + ny + nz) + sizeof(float) * (nx + ny + nz))))
        return rc;
    float2* A = (float2*)tmp.p;
    float2* X = A + n;
    float2* twx = X + n;
    float2* twy = twx + 3 * nx;
    float2* twz = twy + 3 * ny;
    float* ones = (float*)(twz + 3 * nz);
    std::vector<float> one(nx + ny + nz, 1.f);
    HIPCHK(c, hipMemcpy(twx, tx.data(), sizeof(float2) * tx.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(twy, ty.data(), sizeof(float2) * ty.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(twz, tz.data(), sizeof(float2) * tz.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(ones, one.data(), sizeof(float) * one.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(A, in, fb, hipMemcpyHostToDevice));
    const bool prof = c->profile;
    c->profile = false;
    XArgs a{};
    a.tw = twx;
    a.shot_stride = n;
    a.ny = ny;
    a.inv_nx = 1.0f;
    a.gx = a.gy = a.gz = ones;
    if (sign < 0) {
        a.A = A;
        a.Xout = X;
        rc = launch_x(c, 0, ex->xk[0], XM_FWD, a, ny * nz, 1, 0.0);
        if (!rc) rc = launch_strided(c, 0, ey->yf[0], X, X, nx, ny, nz, false, 1, n, twy, nullptr, 0);
        if (!rc) rc = launch_strided(c, 0, ez->yf[0], X, X, nx, ny, nz, true, 1, n, twz, nullptr, 0);
    } else {
        rc = launch_strided(c, 0, ez->yi[0], A, A, nx, ny, nz, true, 1, n, twz, nullptr, 0);
        if (!rc) rc = launch_strided(c, 0, ey->yi[0], A, A, nx, ny, nz, false, 1, n, twy, nullptr, 0);
        a.Xprev = A;
        a.A = X;   // XM_INV writes IFFT_x * inv_nx (= 1 here) into A
        if (!rc) rc = launch_x(c, 0, ex->xk[0], XM_INV, a, ny * nz, 1, 0.0);
    }
    c->profile = prof;
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, X, fb, hipMemcpyDeviceToHost));
    return OSE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// learned thin pad: set on a model, status
// ---------------------------------------------------------------------
static BnetGeom make_geom(int nxb, int nyb, int nzb, int bl, int bf, int bu, int nx, int ny, int nz, double vmax,
                          double R)
{
    BnetGeom g{};
    g.nxb = nxb;
    g.nyb = nyb;
    g.nzb = nzb;
    g.bl = bl;
    g.bf = bf;
    g.bu = bu;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.vmax = (float)vmax;
    g.R = (float)R;
    for (int d = 0; d < 4; ++d) g.sh[d] = make_shell(nxb, nyb, nzb, bl, bf, bu, nx, ny, nz, d);
    return g;
}

static int upload_params(ose_ctx* c, DevBuf& buf, uint16_t** wbf, uint16_t** wtbf, float** bias, const float* p,
                         float** master)
{
    // layout: fp32 master [NPAR] | bf16 W [3*WSZ] | bf16 WT [3*WSZ] | fp32 bias [48]
    const size_t o_m = 0, o_w = (sizeof(float) * BN_NPAR + 255) & ~(size_t)255;
    const size_t o_wt = o_w + ((sizeof(uint16_t) * 3 * BN_WSZ + 255) & ~(size_t)255);
    const size_t o_b = o_wt + ((sizeof(uint16_t) * 3 * BN_WSZ + 255) & ~(size_t)255);
    int rc;
    if (!buf.p && (rc = dev_alloc(c, buf, o_b + sizeof(float) * 48))) return rc;
    char* b = (char*)buf.p;
    float* m = (float*)(b + o_m);
    *wbf = (uint16_t*)(b + o_w);
    if (wtbf) *wtbf = (uint16_t*)(b + o_wt);
    *bias = (float*)(b + o_b);
    if (master) *master = m;
    if (p) HIPCHK(c, hipMemcpy(m, p, sizeof(float) * BN_NPAR, hipMemcpyHostToDevice));
    if ((rc = bnet_convert(c->stream, m, *wbf, (uint16_t*)(b + o_wt), *bias)))
        return c->fail(rc, "boundary network weight conversion failed");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return OSE_OK;
}

extern "C" {

int ose_boundary_net_size(void) { return BN_NPAR; }

int ose_set_boundary_net(ose_ctx* c, int model_id, const float* params, size_t n, double guard, int replace_taper)
{
    if (!c) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_set_boundary_net: bad model");
    Model& m = *c->models[model_id];
    int rc;
    if ((rc = set_device(c))) return rc;
    if (!params) {   // remove: plain taper
        m.learned = m.learned_on = false;
        drop_graph(c);
        return OSE_OK;
    }
    if (n != (size_t)BN_NPAR) return c->fail(OSE_EINVAL, "ose_set_boundary_net: expected %d parameters", BN_NPAR);
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i]))
            return c->fail(OSE_ENUMERIC, "ose_set_boundary_net: non-finite parameter %zu (keeping the taper)", i);
    if (!(guard > 1.0)) return c->fail(OSE_EINVAL, "ose_set_boundary_net: guard must exceed 1");
    const ose_model_desc& d = m.d;
    if (d.nx < 8 || d.ny < 8 || d.nz < 8)
        return c->fail(OSE_EINVAL, "ose_set_boundary_net: interior too small for the network's receptive field");
    float vmax = 0.f;
    {
        // vmax of the model from v/R
        std::vector<float> vr((size_t)m.ncell);
        HIPCHK(c, hipMemcpy(vr.data(), m.d_vr, sizeof(float) * vr.size(), hipMemcpyDeviceToHost));
        for (float x : vr) vmax = std::max(vmax, x);
        vmax = (float)(vmax * d.R);
    }
    m.bg = make_geom(d.nxb, d.nyb, d.nzb, d.bl, d.bf, d.bu, d.nx, d.ny, d.nz, vmax, d.R);
    if ((rc = upload_params(c, m.bnw, &m.d_wbf, nullptr, &m.d_bias, params, nullptr))) return rc;
    m.guard = (float)guard;
    m.replace_taper = replace_taper != 0;
    if (m.replace_taper) {
        const int nt = d.nxb + d.nyb + d.nzb;
        if (!m.ones.p && (rc = dev_alloc(c, m.ones, sizeof(float) * nt))) return rc;
        std::vector<float> one((size_t)nt, 1.f);
        HIPCHK(c, hipMemcpy(m.ones.p, one.data(), sizeof(float) * nt, hipMemcpyHostToDevice));
    }
    m.learned = m.learned_on = true;
    if (c->finalized && (rc = alloc_bnet_bufs(c, m, (int)m.shots.size(), false))) return rc;
    drop_graph(c);
    return OSE_OK;
}

int ose_trip_boundary_net(ose_ctx* c, int model_id)
{
    if (!c) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_trip_boundary_net: bad model");
    Model& m = *c->models[model_id];
    if (!m.learned) return c->fail(OSE_ESTATE, "ose_trip_boundary_net: the model has no boundary network");
    int rc;
    if ((rc = set_device(c)) || (rc = finalize(c)) || (rc = ose_sync(c))) return rc;
    if (!m.bb.tripped) return c->fail(OSE_ESTATE, "ose_trip_boundary_net: the model has no shots");
    const int one = 1;
    HIPCHK(c, hipMemcpy(m.bb.tripped, &one, sizeof(int), hipMemcpyHostToDevice));
    m.learned_on = false;
    drop_graph(c);
    return OSE_OK;
}

int ose_boundary_status(ose_ctx* c, int model_id, int* active, int* tripped)
{
    if (!c) return OSE_EINVAL;
    if (model_id < 0 || model_id >= (int)c->models.size()) return c->fail(OSE_EINVAL, "ose_boundary_status: bad model");
    Model& m = *c->models[model_id];
    int rc;
    if ((rc = ose_sync(c))) return rc;
    int tr = 0;
    if (m.learned && m.bb.tripped) HIPCHK(c, hipMemcpy(&tr, m.bb.tripped, sizeof(int), hipMemcpyDeviceToHost));
    if (active) *active = (m.learned && !tr) ? 1 : 0;
    if (tripped) *tripped = tr;
    return OSE_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// training object
// ---------------------------------------------------------------------
struct ose_bnet {
    ose_ctx c;                 // device, stream and error text
    Model m;                   // geometry, v/R and network buffers
    long ncell = 0;
    DevBuf params, opt, scratch, fields, lossbuf;
    float* d_master = nullptr;
    uint16_t *d_wt = nullptr;
    float *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t *dy_a = nullptr, *dy_b = nullptr;
    float2 *d_u = nullptr, *d_t = nullptr;
    std::vector<std::unique_ptr<DevBuf>> samples;   // each: u | t
    int step = 0;
};

extern "C" {

int ose_bnet_create(ose_bnet** out, int device, int nxb, int nyb, int nzb, int bl, int bf, int bu, int nx, int ny,
                    int nz, const float* v_padded)
{
    if (!out || !v_padded) return OSE_EINVAL;
    *out = nullptr;
    if (nxb <= 0 || nyb <= 0 || nzb <= 0 || bl < 0 || bf < 0 || bu < 0 || nx < 8 || ny < 8 || nz < 8 ||
        bl + nx > nxb || bf + ny > nyb || bu + nz > nzb)
        return OSE_EINVAL;
    auto* b = new (std::nothrow) ose_bnet();
    if (!b) return OSE_ENOMEM;
    ose_ctx* c = &b->c;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return OSE_EDEVICE;
    }
    const long n = (long)nxb * nyb * nzb;
    b->ncell = n;
    double vmax = 0.0;
    for (long i = 0; i < n; ++i) {
        if (!std::isfinite(v_padded[i]) || v_padded[i] <= 0.f) {
            (void)hipStreamDestroy(c->stream);
            delete b;
            return OSE_ENUMERIC;
        }
        vmax = std::max(vmax, (double)v_padded[i]);
    }
    Model& m = b->m;
    m.ncell = n;
    m.d.nxb = nxb; m.d.nyb = nyb; m.d.nzb = nzb;
    m.d.bl = bl; m.d.bf = bf; m.d.bu = bu; m.d.nx = nx; m.d.ny = ny; m.d.nz = nz;
    m.d.R = vmax;   // features use v/vmax; with R = vmax, v/R is that directly
    m.bg = make_geom(nxb, nyb, nzb, bl, bf, bu, nx, ny, nz, vmax, vmax);
    std::vector<float> vr((size_t)n);
    for (long i = 0; i < n; ++i) vr[i] = (float)(v_padded[i] / vmax);
    int rc = OSE_OK;
    if ((rc = dev_alloc(c, m.vr, sizeof(float) * n))) goto fail;
    m.d_vr = (float*)m.vr.p;
    if (hipMemcpy(m.d_vr, vr.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess) { rc = OSE_EDEVICE; goto fail; }
    if ((rc = alloc_bnet_bufs(c, m, 1, true))) goto fail;
    {
        // grads, Adam moments
        if ((rc = dev_alloc(c, b->opt, sizeof(float) * 3 * BN_NPAR))) goto fail;
        b->d_grad = (float*)b->opt.p;
        b->d_m = b->d_grad + BN_NPAR;
        b->d_v = b->d_m + BN_NPAR;
        if (hipMemset(b->opt.p, 0, sizeof(float) * 3 * BN_NPAR) != hipSuccess) { rc = OSE_EDEVICE; goto fail; }
        const long n1 = m.bg.sh[1].count, n2 = m.bg.sh[2].count;
        if ((rc = dev_alloc(c, b->scratch, sizeof(uint16_t) * 16 * (n2 + n1) + 512))) goto fail;
        b->dy_a = (uint16_t*)b->scratch.p;
        b->dy_b = b->dy_a + 16 * n2;
        if ((rc = dev_alloc(c, b->fields, sizeof(float2) * 2 * n))) goto fail;
        b->d_u = (float2*)b->fields.p;
        b->d_t = b->d_u + n;
        if ((rc = dev_alloc(c, b->lossbuf, sizeof(float) * (1 << 20)))) goto fail;
        // initial parameters: He-scaled layers 1-2, zero last layer (starts as the plain taper)
        std::vector<float> p(BN_NPAR, 0.f);
        std::mt19937_64 rng(12345);
        std::normal_distribution<float> nd(0.f, 1.f);
        for (int l = 0; l < 2; ++l) {
            const float sd = std::sqrt(2.0f / (27.0f * (l == 0 ? 9.0f : 16.0f)));
            for (int i = 0; i < BN_WSZ; ++i) p[l * (BN_WSZ + 16) + i] = sd * nd(rng);
        }
        if ((rc = upload_params(c, b->params, &m.d_wbf, &b->d_wt, &m.d_bias, p.data(), &b->d_master))) goto fail;
    }
    *out = b;
    return OSE_OK;
fail:
    (void)hipStreamDestroy(c->stream);
    delete b;
    return rc;
}

int ose_bnet_destroy(ose_bnet* b)
{
    if (!b) return OSE_EINVAL;
    (void)hipSetDevice(b->c.device);
    (void)hipStreamSynchronize(b->c.stream);
    hipStream_t st = b->c.stream;
    delete b;
    (void)hipStreamDestroy(st);
    return OSE_OK;
}

const char* ose_bnet_error(const ose_bnet* b) { return b ? b->c.err.c_str() : "null network"; }

int ose_bnet_set_params(ose_bnet* b, const float* p, size_t n)
{
    if (!b || !p || n != (size_t)BN_NPAR) return OSE_EINVAL;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return b->c.fail(OSE_ENUMERIC, "non-finite parameter %zu", i);
    (void)hipSetDevice(b->c.device);
    return upload_params(&b->c, b->params, &b->m.d_wbf, &b->d_wt, &b->m.d_bias, p, &b->d_master);
}

int ose_bnet_get_params(ose_bnet* b, float* p, size_t n)
{
    if (!b || !p || n != (size_t)BN_NPAR) return OSE_EINVAL;
    ose_ctx* c = &b->c;
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(p, b->d_master, sizeof(float) * n, hipMemcpyDeviceToHost));
    return OSE_OK;
}

int ose_bnet_forward(ose_bnet* b, const float* u, float* delta)
{
    if (!b || !u || !delta) return OSE_EINVAL;
    ose_ctx* c = &b->c;
    (void)hipSetDevice(c->device);
    Model& m = b->m;
    HIPCHK(c, hipMemcpy(b->d_u, u, sizeof(float2) * b->ncell, hipMemcpyHostToDevice));
    int rc;
    if ((rc = bnet_forward(c->stream, m.bg, m.d_wbf, m.d_bias, b->d_u, m.d_vr, b->ncell, 1, m.bb, 0.f, false)))
        return c->fail(rc, "bnet_forward launch failed");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const Shell& s0 = m.bg.sh[0];
    std::vector<float> y3((size_t)s0.count * 2);
    float st[4];
    HIPCHK(c, hipMemcpy(y3.data(), m.bb.y3, sizeof(float) * y3.size(), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(st, m.bb.stats, sizeof(st), hipMemcpyDeviceToHost));
    (void)st;
    std::memset(delta, 0, sizeof(float) * 2 * b->ncell);
    for (long i = 0; i < s0.count; ++i) {
        int x, y, z;
        shell_coords(s0, i, x, y, z);
        const long cell = ((long)z * s0.ny + y) * s0.nx + x;
        const float ur = u[2 * cell], ui = u[2 * cell + 1];
        float qx = 1.f + y3[2 * i], qy = y3[2 * i + 1];   // passive gain G = q / max(1, |q|)
        const float m2 = qx * qx + qy * qy;
        if (m2 > 1.f) {
            const float r = 1.f / std::sqrt(m2);
            qx *= r;
            qy *= r;
        }
        delta[2 * cell] = ur * (qx - 1.f) - ui * qy;     // u (G - 1) on the pad
        delta[2 * cell + 1] = ui * (qx - 1.f) + ur * qy;
    }
    return OSE_OK;
}

int ose_bnet_gradient(ose_bnet* b, const float* u, const float* t, float* loss, float* grad)
{
    if (!b || !u || !t || !loss || !grad) return OSE_EINVAL;
    ose_ctx* c = &b->c;
    (void)hipSetDevice(c->device);
    Model& m = b->m;
    HIPCHK(c, hipMemcpy(b->d_u, u, sizeof(float2) * b->ncell, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(b->d_t, t, sizeof(float2) * b->ncell, hipMemcpyHostToDevice));
    float* dl = (float*)b->lossbuf.p;
    int rc;
    if ((rc = bnet_backward(c->stream, m.bg, m.d_wbf, b->d_wt, m.d_bias, b->d_u, b->d_t, m.d_vr, b->ncell, m.bb,
                            b->dy_a, b->dy_b, b->d_grad, dl)))
        return c->fail(rc, "bnet_backward launch failed");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(grad, b->d_grad, sizeof(float) * BN_NPAR, hipMemcpyDeviceToHost));
    return OSE_OK;
}

int ose_bnet_add_sample(ose_bnet* b, const float* u, const float* t, int* count)
{
    if (!b || !u || !t) return OSE_EINVAL;
    ose_ctx* c = &b->c;
    (void)hipSetDevice(c->device);
    auto s = std::make_unique<DevBuf>();
    int rc;
    if ((rc = dev_alloc(c, *s, sizeof(float2) * 2 * b->ncell))) return rc;
    HIPCHK(c, hipMemcpy(s->p, u, sizeof(float2) * b->ncell, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy((float2*)s->p + b->ncell, t, sizeof(float2) * b->ncell, hipMemcpyHostToDevice));
    b->samples.push_back(std::move(s));
    if (count) *count = (int)b->samples.size();
    return OSE_OK;
}

int ose_bnet_train(ose_bnet* b, int iters, float lr, unsigned long long seed, float* losses)
{
    if (!b || iters <= 0 || iters > (1 << 20) || !(lr > 0.f)) return OSE_EINVAL;
    ose_ctx* c = &b->c;
    if (b->samples.empty()) return c->fail(OSE_ESTATE, "ose_bnet_train: no samples");
    (void)hipSetDevice(c->device);
    Model& m = b->m;
    std::mt19937_64 rng(seed);
    float* dl = (float*)b->lossbuf.p;
    const uint16_t* wt_dummy = nullptr;
    (void)wt_dummy;
    for (int it = 0; it < iters; ++it) {
        const DevBuf& smp = *b->samples[rng() % b->samples.size()];
        const float2* u = (const float2*)smp.p;
        const float2* t = u + b->ncell;
        int rc;
        if ((rc = bnet_backward(c->stream, m.bg, m.d_wbf, b->d_wt, m.d_bias, u, t, m.d_vr, b->ncell, m.bb, b->dy_a,
                                b->dy_b, b->d_grad, dl + it)))
            return c->fail(rc, "bnet_backward launch failed");
        ++b->step;
        if ((rc = bnet_adam(c->stream, b->d_master, b->d_grad, b->d_m, b->d_v, b->step, lr, 0.9f, 0.999f, 1e-8f)))
            return c->fail(rc, "adam launch failed");
        if ((rc = bnet_convert(c->stream, b->d_master, m.d_wbf, b->d_wt, m.d_bias)))
            return c->fail(rc, "weight conversion failed");
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (losses) HIPCHK(c, hipMemcpy(losses, dl, sizeof(float) * iters, hipMemcpyDeviceToHost));
    return OSE_OK;
}

}  // extern "C"
