// ose_kernels.h -- CDNA4 (gfx950) kernels of the OSE time step.
//
// One Chebyshev order of the one-step extrapolation applies
//     L W = (v/R) * IFFT3( |k| FFT3(W) ) / N
// to the packed complex field W_j = (-i)^j Z_j (Re = reference p-lane,
// Im = q-lane, ose_3d_lib.cu:639-874) and runs the three-term recursion
//     W_{j+1} = W_{j-1} - 2i L W_j,   A = g (A + c_{j+1} W_{j+1}).
// It is four HBM passes over complex64 fields laid out [shot][z][y][x]:
//   1. k_strided<SIGN=-1>  forward-y       X -> S
//   2. k_zsym              forward-z, * |k|/N, inverse-z   S -> X
//                          (or k_zsym_chunk: several tiles per workgroup)
//   3. k_strided<SIGN=+1>  inverse-y       X -> S
//   4. k_xpass<MODE>       inverse-x of S, recursion on physical W, taper,
//                          accumulator (on about half of the orders, two
//                          terms at a time), receivers, forward-x of W_{j+1}
//                          into X
// X starts as X_j = FFT_x W_j; passes 1-3 alternate between the two spectral
// fields X and S (or run in place where a batch is small enough for S to be
// X); pass 4 reads W_{j-1} in physical form and overwrites it with W_{j+1}
// (two physical W buffers alternate), so each order needs one inverse and
// one forward x transform.
//
// Every 1D transform of length N = N1*N2 is done in registers and LDS:
// step A: N2 threads each run an N1-point register DFT over x[N2*n1 + t],
// multiply by exp(s 2 pi i t k1/N) and write LDS; step B: N1 threads each
// gather over n2 and run an N2-point register DFT, giving X[k1 + N1*k2].
// One LDS exchange per transform; global loads and stores are coalesced
// along x in every pass.
#pragma once
#include <hip/hip_runtime.h>

#include "dft_gen.h"

namespace ose {

template <int A, int B> struct Max { static constexpr int value = A > B ? A : B; };

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Extents without compiled register kernels take the runtime-length path of
// ose_generic.h: a self-sorting mixed-radix transform staged through LDS.
// Its plan travels in the kernel arguments (unused by the compiled kernels).
constexpr int GEN_MAXF = 16;      // stages
struct GenPlan {
    int N = 0;                    // transform length
    int TP = 0;                   // threads per pencil / line
    int W = 0;                    // pencils (T columns) or lines (P) per workgroup
    int nf = 0;                   // stages
    short r[GEN_MAXF] = {};       // radix of each stage (<= 16: register
                                  // butterfly; larger primes: direct sums)
};

// ---------------------------------------------------------------------
// Strided passes (y and z): a workgroup owns T consecutive x columns of
// one (z or y) plane and transforms all T pencils of length N.
// ---------------------------------------------------------------------
struct StridedArgs {
    const float2* in;
    float2* out;
    int ncol;           // columns along x (nx)
    long es;            // element stride along the pencil
    long os;            // stride of the blockIdx.y index
    long shot_stride;   // cells per shot
    const float2* tw;   // root table of the pencil axis, three blocks of N:
                        // w^j; w^(t k1) at N + t N1 + k1; w^(t n2) at
                        // 2N + t N2 + n2 (w = exp(-2 pi i / N))
    // symbol tables (z pass only): squared folded wavenumbers
    const float* kx2;
    const float* ky2;
    const float* kz2;
    float invN;         // 1 / (nx ny nz)
    int xcd;            // 1: XCD-contiguous tile order (StridedTile)
    unsigned long long mgx, mgy;   // 2^40 / gridDim.x + 1 and 2^40 / gridDim.y + 1
                                   // (chunked z pass: of tgx and tgy)
    // chunked z pass only (k_zsym_chunk): tiles tgx (column blocks) x tgy
    // (rows) x shots, `chunk` consecutive tiles per workgroup
    int tgx, tgy;
    unsigned ntiles;
    int chunk;
    GenPlan g;          // runtime-length path only
};

// floor(n / d) as (n * m) >> 40 with m = 2^40 / d + 1: exact for n < 2^24 and
// d < 2^16 (the error n (m - 2^40/d) / 2^40 < n / 2^40 stays below 1/d), and
// all in scalar registers for wave-uniform n (a division by a runtime value
// would be expanded into vector reciprocal code).
__device__ __forceinline__ unsigned div_magic(unsigned n, unsigned long long m)
{
    return (unsigned)(((unsigned long long)n * m) >> 40);
}

// The workgroup's tile (column block bx, row by, shot bz).  Workgroups are
// dispatched round-robin over the chip's 8 XCDs by linear id; with xcd set,
// XCD x takes the contiguous tile range [x q + min(x, r), ...) of the
// q = n / 8, r = n % 8 split, column block fastest, so the workgroups that
// run together on one XCD hold neighbouring column blocks of the same rows
// (one set of pages, whole 256-byte segments) instead of blocks 8 apart.
struct StridedTile {
    int bx, by, bz;
    StridedTile() = default;
    __device__ __forceinline__ explicit StridedTile(const StridedArgs& a)
    {
        if (!a.xcd) {
            bx = blockIdx.x;
            by = blockIdx.y;
            bz = blockIdx.z;
            return;
        }
        const unsigned gx = gridDim.x, gy = gridDim.y;
        const unsigned n = gx * gy * gridDim.z;
        const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const unsigned x = id % 8, q = n / 8, r = n % 8;
        const unsigned tile = x * q + min(x, r) + id / 8;
        const unsigned row = div_magic(tile, a.mgx);
        const unsigned shot = div_magic(row, a.mgy);
        bx = (int)(tile - row * gx);
        by = (int)(row - shot * gy);
        bz = (int)shot;
    }
};

// Row pitch (elements) of the LDS column exchange.  A wave's exchange access
// covers T consecutive columns of 64 / T consecutive t, whose rows step by 1
// on one side of each exchange and by N1 on the other, so its groups of T
// lanes start pitch * 1 or pitch * N1 elements apart.  With pitch T and odd
// N1 the groups fall on disjoint banks on both sides (32 4-byte banks per 32
// lanes in the half-LDS form; also with 64 banks per wave and in the 8-byte
// form); the padded T + 1 puts two groups on overlapping banks there, about
// twice the LDS cycles.  For even N1 no pitch serves both sides and the
// padded row stays.
template <int T, int N1>
struct XchgPitch {
    static constexpr int value = (N1 & 1) ? T : T + 1;
};

// SPLIT = 1 exchanges the real and then the imaginary parts through an LDS
// array of floats: half the LDS per workgroup (twice the workgroups per CU
// for the long extents) for two more barriers.
template <int N1, int N2, int T, int SPLIT>
struct StridedCfg {
    static constexpr int N = N1 * N2;
    static constexpr int NT = Max<N1, N2>::value;
    static constexpr int THREADS = T * NT;
    static constexpr int LDT = XchgPitch<T, N1>::value;     // exchange row pitch
    static constexpr size_t LDS_BYTES = (SPLIT ? sizeof(float) : sizeof(float2)) * (size_t)N * LDT;
};

// LDS exchange among the T columns' threads: writers (wr) store v[i] for
// i < NW at row wrow(i), readers (rd) load v[i] for i < NR from row
// rrow(i).  In split form a thread that both writes and reads overwrites
// only the real parts of v[0..NR) before it stores its imaginary parts.
//
// LDS row of access i.  One side of every exchange strides across all of it,
// and ds instructions carry at most a 64 KiB immediate offset: beyond that
// (T = 32 at 864: 110 KiB) the compiler gives each far access its own
// address register, which spills.  So where the exchange exceeds 64 KiB the
// upper half of a loop addresses from row(n/2), passed through an empty asm
// statement so that it stays one base register with immediate offsets.
template <bool FAR, int NI, typename FROW>
__device__ __forceinline__ int xrow(FROW row, int i)
{
    if (!FAR || i < NI / 2) return row(i);
    int mid = row(NI / 2);
    asm("" : "+v"(mid));
    return mid + (row(i) - row(NI / 2));
}

template <int T, int LDT, int SPLIT, int NW, int NR, typename FW, typename FR>
__device__ __forceinline__ void col_exchange(float2* v, bool wr, bool rd, FW wrow, FR rrow, float2* lds, int col)
{
    constexpr bool FAR = (size_t)NW * NR * LDT * (SPLIT ? sizeof(float) : sizeof(float2)) > 65536;
    if (!SPLIT) {
        if (wr) {
#pragma unroll
            for (int i = 0; i < NW; ++i) lds[xrow<FAR, NW>(wrow, i) * LDT + col] = v[i];
        }
        __syncthreads();
        if (rd) {
#pragma unroll
            for (int i = 0; i < NR; ++i) v[i] = lds[xrow<FAR, NR>(rrow, i) * LDT + col];
        }
    } else {
        float* f = reinterpret_cast<float*>(lds);
        if (wr) {
#pragma unroll
            for (int i = 0; i < NW; ++i) f[xrow<FAR, NW>(wrow, i) * LDT + col] = v[i].x;
        }
        __syncthreads();
        if (rd) {
#pragma unroll
            for (int i = 0; i < NR; ++i) v[i].x = f[xrow<FAR, NR>(rrow, i) * LDT + col];
        }
        __syncthreads();
        if (wr) {
#pragma unroll
            for (int i = 0; i < NW; ++i) f[xrow<FAR, NW>(wrow, i) * LDT + col] = v[i].y;
        }
        __syncthreads();
        if (rd) {
#pragma unroll
            for (int i = 0; i < NR; ++i) v[i].y = f[xrow<FAR, NR>(rrow, i) * LDT + col];
        }
    }
}

// Addressing of a strided pass: element (pencil row r0 + t, column col) of
// the workgroup's T columns lives at the workgroup's base (uniform) + r0 *
// es (uniform) + a 32-bit per-thread byte offset (col + t es) * 8, so every
// global access takes a scalar base and one 32-bit vector offset instead of
// a 64-bit multiply per element.  Threads past the last column load from the
// last column (in bounds, results unused) and do not store.
struct StridedAddr {
    long wbase;        // element index of the workgroup's first column
    long es;
    unsigned boff;     // (xc + t es) * sizeof(float2)
    int xc;            // column actually addressed
    bool ok;           // the thread's own column exists

    __device__ __forceinline__ StridedAddr(const StridedArgs& a, const StridedTile& g, int T, int col, int t)
    {
        const int x0 = g.bx * T;
        xc = min(col, a.ncol - 1 - x0);
        ok = col == xc;
        es = a.es;
        wbase = (long)g.bz * a.shot_stride + (long)g.by * a.os + x0;
        boff = ((unsigned)xc + (unsigned)t * (unsigned)a.es) * (unsigned)sizeof(float2);
    }
    typedef __attribute__((address_space(1))) char gchar;
    typedef float vf2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) vf2 gvf2;
    // the wave-uniform global address of pencil row r0, pinned to scalar
    // registers so that the compiler keeps it apart from the per-thread offset
    __device__ __forceinline__ gchar* row(const float2* p, int r0) const
    {
        const unsigned long long u = reinterpret_cast<unsigned long long>(p + wbase + (long)r0 * es);
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
        return (gchar*)(((unsigned long long)hi << 32) | lo);
    }
    // element at pencil index r0 + t
    __device__ __forceinline__ float2 load(const float2* p, int r0) const
    {
        const vf2 r = *(const gvf2*)(row(p, r0) + boff);
        return make_float2(r.x, r.y);
    }
    __device__ __forceinline__ void store(float2* p, int r0, float2 v) const
    {
        vf2 r;
        r.x = v.x;
        r.y = v.y;
        *(gvf2*)(row(p, r0) + boff) = r;
    }
};

template <int N1, int N2, int T, int SIGN, int SPLIT>
__global__ __launch_bounds__((StridedCfg<N1, N2, T, SPLIT>::THREADS))
void k_strided(StridedArgs a)
{
    using C = StridedCfg<N1, N2, T, SPLIT>;
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int col = tid % T;
    const int t = tid / T;
    const StridedTile g(a);
    const StridedAddr ad(a, g, T, col, t);
    float2 v[C::NT];
    if (t < N2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = ad.load(a.in, N2 * n1);
        rdft<N1, SIGN>(v);
        const float2* twt = a.tw + C::N + t * N1;
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1) {
            const float2 w = twt[k1];
            v[k1] = SIGN < 0 ? cmul(v[k1], w) : cmulc(v[k1], w);
        }
    }
    col_exchange<T, C::LDT, SPLIT, N1, N2>(
        v, t < N2, t < N1, [&](int k1) { return t * N1 + k1; }, [&](int n2) { return n2 * N1 + t; }, lds, col);
    if (t < N1) {
        rdft<N2, SIGN>(v);
        if (ad.ok) {
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) ad.store(a.out, N1 * k2, v[k2]);
        }
    }
}

// Access-pattern probe: the HBM reads and writes of a strided pass (pencil
// rows N2 n1 + t of T columns, the tile order and addressing of k_strided
// and k_zsym) with no arithmetic, for N1, N2 <= 32 given at run time.  It
// measures what the pass's access pattern alone allows.
template <int T>
__global__ __launch_bounds__(T * 32) void k_strided_copy(StridedArgs a, int N1, int N2)
{
    const int tid = threadIdx.x;
    const int col = tid % T;
    const int t = tid / T;
    const StridedTile g(a);
    const StridedAddr ad(a, g, T, col, t);
    if (t >= N2) return;
    float2 v[32];
#pragma unroll
    for (int n1 = 0; n1 < 32; ++n1)
        if (n1 < N1) v[n1] = ad.load(a.in, N2 * n1);
    if (!ad.ok) return;
#pragma unroll
    for (int n1 = 0; n1 < 32; ++n1)
        if (n1 < N1) ad.store(a.out, N2 * n1, v[n1]);
}

// z pass: forward-z, multiply by |k|/N (folded wavenumbers, scale_kxkz_3D
// ose_3d_kernel.cu:101-126 and the 1/N of scal_vr_3d :355), inverse-z.
// In to out (the same field where a batch runs in place); the tile row
// (StridedTile::by) is the (spectral) y index.  Each LDS location of the
// second exchange is written by the thread that read it in the first, so
// the exchanges need no barrier between them.
// The split form asks for at least 4 waves per SIMD (at most 128 VGPRs):
// its half-size LDS lets two 512-thread workgroups share a CU only then,
// and the z pass needs them to cover its address translation latency.  (The
// whole-complex form is held to one workgroup per CU by its LDS anyway.)
//
// The twiddles and |k_z|^2 come from LDS: a workgroup copies the twiddle
// rows w^(t n2) (block 2 of the root table) and kz2 there before it issues
// its pencil reads, so that the copy's reads complete first (the vector
// memory counter completes in order) and no global read is left between
// the pencil's arrival and the stores.  (Read from global memory at the
// point of use, each of the three tables cost a round trip through the
// memory queue in the middle of the arithmetic.)  With the rows w^(t n2)
// the forward twiddle w^(k1 n2) is applied after the first exchange, by the
// thread that gathers k1 = t: the same products the writing side would
// form, and the same rows serve the inverse.
template <int N1, int N2, int T, int SPLIT>
struct ZsymCfg {
    using SC = StridedCfg<N1, N2, T, SPLIT>;
    static constexpr size_t TW_OFF = (SC::LDS_BYTES + 15) & ~(size_t)15;    // w^(t n2), N float2
    static constexpr size_t KZ_OFF = TW_OFF + sizeof(float2) * (size_t)SC::N; // |k_z|^2, N floats
    static constexpr size_t LDS_BYTES = KZ_OFF + sizeof(float) * (size_t)SC::N;
    static_assert(LDS_BYTES <= 160 * 1024, "z pass LDS beyond a CU's 160 KiB");
};

template <int N1, int N2, int T, int SPLIT>
__global__ __launch_bounds__((StridedCfg<N1, N2, T, SPLIT>::THREADS), (SPLIT ? 4 : 1))
void k_zsym(StridedArgs a)
{
    using C = StridedCfg<N1, N2, T, SPLIT>;
    using Z = ZsymCfg<N1, N2, T, SPLIT>;
    extern __shared__ float2 lds[];
    float2* ltw = reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + Z::TW_OFF);
    float* lkz = reinterpret_cast<float*>(reinterpret_cast<char*>(lds) + Z::KZ_OFF);
    const int tid = threadIdx.x;
    const int col = tid % T;
    const int t = tid / T;
    const StridedTile g(a);
    const StridedAddr ad(a, g, T, col, t);
    // the tables' and the column's reads first (indices clamped, so that no
    // read is conditional), then the pencil's; the tables go to LDS once
    // their reads are in, while the pencil's are still in flight
    constexpr int NTAB = (C::N + C::THREADS - 1) / C::THREADS;
    float2 tw_r[NTAB];
    float kz_r[NTAB];
#pragma unroll
    for (int j = 0; j < NTAB; ++j) {
        const int i = min(tid + j * C::THREADS, C::N - 1);
        tw_r[j] = a.tw[2 * C::N + i];
        kz_r[j] = a.kz2[i];
    }
    const float kxy = a.kx2[g.bx * T + ad.xc] + a.ky2[g.by];
    float2 v[C::NT];
    if (t < N2) {
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) v[n1] = ad.load(a.in, N2 * n1);
    }
#pragma unroll
    for (int j = 0; j < NTAB; ++j) {
        const int i = tid + j * C::THREADS;
        if (i < C::N) {
            ltw[i] = tw_r[j];
            lkz[i] = kz_r[j];
        }
    }
    if (t < N2) rdft<N1, -1>(v);
    col_exchange<T, C::LDT, SPLIT, N1, N2>(
        v, t < N2, t < N1, [&](int k1) { return t * N1 + k1; }, [&](int n2) { return n2 * N1 + t; }, lds, col);
    if (t < N1) {
        // thread t holds X1[k1 = t][n2]: twiddle w^(t n2), then the N2-point DFT
        const float2* twr = ltw + t * N2;
#pragma unroll
        for (int n2 = 1; n2 < N2; ++n2) v[n2] = cmul(v[n2], twr[n2]);
        rdft<N2, -1>(v);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            // hardware square root (1 ulp): |k| needs no correctly rounded
            // result, and the rounding-corrected sqrtf costs ~10 instructions
            const float s = __builtin_amdgcn_sqrtf(kxy + lkz[t + N1 * k2]) * a.invN;
            v[k2].x *= s;
            v[k2].y *= s;
        }
        // inverse: thread m1 = t holds X[m1 + N1 m2]
        rdft<N2, 1>(v);
        // the same rows, read again: an opaque copy of the pointer keeps the
        // compiler from holding the forward multiply's 31 values in
        // registers across the transforms (41 spilled VGPRs)
        const float2* twi = twr;
        asm("" : "+v"(twi));
#pragma unroll
        for (int n2 = 1; n2 < N2; ++n2) v[n2] = cmulc(v[n2], twi[n2]);
    }
    col_exchange<T, C::LDT, SPLIT, N2, N1>(
        v, t < N1, t < N2, [&](int n2) { return n2 * N1 + t; }, [&](int m1) { return t * N1 + m1; }, lds, col);
    if (t < N2) {
        rdft<N1, 1>(v);
        if (ad.ok) {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) ad.store(a.out, N2 * n1, v[n1]);
        }
    }
}

// Chunked z pass: the arithmetic of k_zsym (whole-complex exchange) on
// a.chunk consecutive tiles per workgroup, chunks in XCD-contiguous order.
// Each tile's successor is loaded into registers before the tile is
// transformed, so the next tile's HBM reads overlap this tile's work: the z
// pencils' reads cross a 6 MB page stride per element and take the longest
// address translation of the step, which one tile per workgroup (two
// workgroups per CU) leaves exposed.  The twiddle rows and the squared z
// wavenumbers are staged in LDS once per workgroup: read from global memory
// inside the loop they would queue behind the prefetched reads (the vector
// memory counter completes in order) and stall every tile on them.  One
// workgroup per CU (its registers hold a tile and a half).
template <int N1, int N2, int T>
struct ZChunkCfg {
    static constexpr int N = N1 * N2;
    static constexpr int NT = Max<N1, N2>::value;
    static constexpr int THREADS = T * NT;
    static constexpr int LDT = XchgPitch<T, N1>::value;
    static constexpr size_t XCH = (size_t)N * LDT;   // exchange, float2
    static constexpr size_t LDS_BYTES = sizeof(float2) * (XCH + 2 * (size_t)N) + sizeof(float) * (size_t)N;
};

// tile L of the chunked z pass: column block fastest, then row, then shot
__device__ __forceinline__ StridedTile zc_tile(const StridedArgs& a, unsigned L)
{
    const unsigned row = div_magic(L, a.mgx);
    const unsigned shot = div_magic(row, a.mgy);
    StridedTile g;
    g.bx = (int)(L - row * (unsigned)a.tgx);
    g.by = (int)(row - shot * (unsigned)a.tgy);
    g.bz = (int)shot;
    return g;
}

// issue tile L's reads: the column's kx^2 and the row's ky^2, then the
// pencil values.  With step = 0 every read hits the tile's first row (the
// prefetch after the chunk's last tile: cache hits, results unused), so the
// reads are issued unconditionally and the wait before a tile's work stays
// exact (a conditional prefetch made the compiler wait for every read).
template <int N1, int N2, int T>
__device__ __forceinline__ void zc_fetch(const StridedArgs& a, unsigned L, int step, int col, int t,
                                         float2 (&vn)[N1], float& kxn, float& kyn)
{
    const StridedTile g = zc_tile(a, L);
    const StridedAddr ad(a, g, T, col, t);
    kxn = a.kx2[g.bx * T + ad.xc];
    kyn = a.ky2[g.by];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) vn[n1] = ad.load(a.in, step * N2 * n1);
}

template <int N1, int N2, int T>
__global__ __launch_bounds__((ZChunkCfg<N1, N2, T>::THREADS), 1)
void k_zsym_chunk(StridedArgs a)
{
    using C = ZChunkCfg<N1, N2, T>;
    static_assert(N1 <= N2, "every thread of the chunked z pass loads a pencil slice");
    extern __shared__ float2 lds[];
    float2* ltw = lds + C::XCH;     // w^(t k1) at t N1 + k1; w^(t n2) at N + t N2 + n2
    float* lkz = reinterpret_cast<float*>(ltw + 2 * C::N);
    const int tid = threadIdx.x;
    const int col = tid % T;
    const int t = tid / T;
    const unsigned n = gridDim.x, id = blockIdx.x;
    const unsigned x = id % 8, q = n / 8, r = n % 8;
    const unsigned first = (x * q + min(x, r) + id / 8) * (unsigned)a.chunk;
    const unsigned last = min(first + (unsigned)a.chunk, a.ntiles);
    if (first >= last) return;
    for (int i = tid; i < 2 * C::N; i += C::THREADS) ltw[i] = a.tw[C::N + i];
    for (int i = tid; i < C::N; i += C::THREADS) lkz[i] = a.kz2[i];
    __syncthreads();
    float2 v[C::NT];
    float2 vn[N1];
    float kxn, kyn;
    zc_fetch<N1, N2, T>(a, first, 1, col, t, vn, kxn, kyn);
    for (unsigned L = first; L < last; ++L) {
        const StridedTile g = zc_tile(a, L);
        const StridedAddr ad(a, g, T, col, t);
        const float kxy = kxn + kyn;
#pragma unroll
        for (int i = 0; i < N1; ++i) v[i] = vn[i];
        const bool more = L + 1 < last;
        zc_fetch<N1, N2, T>(a, more ? L + 1 : L, more ? 1 : 0, col, t, vn, kxn, kyn);
        __builtin_amdgcn_sched_barrier(0);
        {
            rdft<N1, -1>(v);
            const float2* twt = ltw + t * N1;
#pragma unroll
            for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twt[k1]);
        }
        col_exchange<T, C::LDT, 0, N1, N2>(
            v, true, t < N1, [&](int k1) { return t * N1 + k1; }, [&](int n2) { return n2 * N1 + t; }, lds, col);
        if (t < N1) {
            rdft<N2, -1>(v);
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) {
                const float s = __builtin_amdgcn_sqrtf(kxy + lkz[t + N1 * k2]) * a.invN;
                v[k2].x *= s;
                v[k2].y *= s;
            }
            rdft<N2, 1>(v);
            const float2* twi = ltw + C::N + t * N2;
#pragma unroll
            for (int n2 = 1; n2 < N2; ++n2) v[n2] = cmulc(v[n2], twi[n2]);
        }
        col_exchange<T, C::LDT, 0, N2, N1>(
            v, t < N1, true, [&](int n2) { return n2 * N1 + t; }, [&](int m1) { return t * N1 + m1; }, lds, col);
        {
            rdft<N1, 1>(v);
            // Unconditional stores: a thread past the last column addresses
            // the last column (StridedAddr) and has transformed that same
            // pencil with the same instructions, so it writes the owner's
            // values bit for bit.  Conditional stores would make the compiler
            // wait for this tile's writes before the next tile's prefetched
            // values could be used (it cannot count writes that may not
            // have been issued).
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) ad.store(a.out, N2 * n1, v[n1]);
        }
        // the next tile's first exchange overwrites what this tile's second read
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// x pass: a workgroup owns P consecutive x lines of one shot.
// ---------------------------------------------------------------------
enum XMode : int {
    XM_ORDER0 = 0,   // W1 = -i L W0;  A = c0 W0 + c1 W1;  out W1, X1
    XM_MID = 1,      // W' = W_{j-1} - 2i L W_j;  A = g (A + c W');  out W', X'
    XM_LAST = 2,     // as MID, u = g (A + c W') is the new state: receivers, out u, X0
    XM_FWD = 3,      // Xout = FFT_x(A)              (state upload)
    XM_INV = 4,      // A = IFFT_x(Xprev) / nx       (transform utilities)
    XM_LASTA = 5,    // as LAST without the forward transform (the boundary network acts on u first)
    XM_SKIP = 6,     // W' = W_{j-1} - 2i L W_j;  out W', X'  (the accumulator is not touched)
    XM_ORDER0N = 7,  // W1 = -i L W0;  out W1, X1             (the accumulator is not touched)
    XM_MIDF = 8,     // as MID, the first update of the step's accumulator (no A read)
};

// Deferred accumulation.  Pass j (j >= 1) holds two terms of the series,
// W_{j-1} (read) and W_{j+1} (written), so the accumulator need not be read
// and written on every order: a touching pass adds the terms assigned to it
// and catches up on the taper applications the skipped orders owed,
//   A <- g^ea A + g^eb cp W_{j-1} + g^ec c1 W_{j+1}    (MIDF: no g^ea A term),
// where the exponents count the taper applications each term has had by
// order j in the reference's order of operations (g once per order j >= 1,
// after adding W_{j+1}).  The host plans which orders touch A; the sum is the
// reference's exactly, up to FP32 rounding of the powers of g.

// The recursion runs on physical W: pass 4 of order j reads W_{j-1} and
// writes W_{j+1} over it (Win == Wout except where noted), so only the
// spectrum S needs an inverse x transform; the new W is forward transformed
// into X for the next order's passes 1-3.
struct XArgs {
    const float2* S;        // spectrum after passes 1-3 (natural order along x)
    const float2* Win;      // physical W_{j-1} (W_0 in ORDER0)
    float2* Wout;           // physical W_{j+1} (W_1 in ORDER0, the new state u in LAST/LASTA)
    float2* Xout;           // x-spectral W_{j+1} (X_0 of the next step in LAST)
    float2* A;              // accumulator
    const float2* Xprev;    // XM_INV input
    const float* vr;        // v / R, [z][y][x], shared by the shots of a model
    const float* gx;
    const float* gy;
    const float* gz;
    const float2* tw;       // x root table (inverse transform; layout of StridedArgs::tw)
    const float2* twf;      // the same table for the forward transform: a
                            // separate pointer keeps the compiler from holding
                            // the inverse's twiddle loads in registers for reuse
                            // by the forward (register pressure, occupancy)
    long shot_stride;
    int ny;
    int nlines;             // ny * nz
    float c0;               // ORDER0: J0
    float c1;               // ORDER0: 2 J1; MID/LAST: coefficient of W_{j+1} (0: not added here)
    float cp;               // MID/LAST: coefficient of W_{j-1} (0: not added here)
    int ea, eb, ec;         // MID/LAST: powers of g (0..3) on A, W_{j-1}, W_{j+1}
    float inv_nx;
    // receivers (LAST): Im(u) on the plane z = rec_z over the interior
    float* traces;          // [shot][nt_cap][rec_ny][rec_nx]
    long trace_shot_stride;
    int rec_z, rec_y0, rec_ny, rec_x0, rec_nx;
    const int* it;          // device step counter
    int nt_cap;
    GenPlan g;              // runtime-length path only
};

template <int N1, int N2, int P, int SPLIT>
struct XCfg {
    static constexpr int N = N1 * N2;
    static constexpr int NT = Max<N1, N2>::value;
    static constexpr int THREADS = P * NT;
    static constexpr int PN = N1 | 1;                 // odd LDS pitch
    static constexpr int LINE = N2 * PN;              // elements (float2, or float when split)
    static constexpr size_t LDS_BYTES = (SPLIT ? sizeof(float) : sizeof(float2)) * (size_t)P * LINE;
};

// LDS exchange within one x line: writers store src[i] (i < NW) at wpos(i),
// readers load dst[i] (i < NR) from rpos(i); src and dst are different
// register arrays.  SPLIT = 1 moves real and imaginary parts in two rounds
// through a float array of half the size.
template <int SPLIT, int NW, int NR, typename FW, typename FR>
__device__ __forceinline__ void line_exchange(const float2* src, bool wr, float2* dst, bool rd, FW wpos, FR rpos,
                                              void* Lv)
{
    if (!SPLIT) {
        float2* L = reinterpret_cast<float2*>(Lv);
        if (wr) {
#pragma unroll
            for (int i = 0; i < NW; ++i) L[wpos(i)] = src[i];
        }
        __syncthreads();
        if (rd) {
#pragma unroll
            for (int i = 0; i < NR; ++i) dst[i] = L[rpos(i)];
        }
    } else {
        float* f = reinterpret_cast<float*>(Lv);
        if (wr) {
#pragma unroll
            for (int i = 0; i < NW; ++i) f[wpos(i)] = src[i].x;
        }
        __syncthreads();
        if (rd) {
#pragma unroll
            for (int i = 0; i < NR; ++i) dst[i].x = f[rpos(i)];
        }
        __syncthreads();
        if (wr) {
#pragma unroll
            for (int i = 0; i < NW; ++i) f[wpos(i)] = src[i].y;
        }
        __syncthreads();
        if (rd) {
#pragma unroll
            for (int i = 0; i < NR; ++i) dst[i].y = f[rpos(i)];
        }
    }
}

// inverse x transform of one line: thread t < N1 loads X[t + N1 m2]; on
// return thread t < N2 holds x[N2 n1 + t] (unnormalised) in res[n1].
// TWB = 1 places a scheduling barrier before the twiddle multiply, so the
// twiddle loads are not hoisted over the register DFT (lower peak register
// use for some extents, at the price of exposed load latency).
template <int N1, int N2, int PN, int SPLIT, int TWB>
__device__ __forceinline__ void x_inverse(const float2* __restrict__ src, bool okl, int t,
                                          void* L, const float2* __restrict__ tw,
                                          float2* __restrict__ w, float2* __restrict__ res)
{
    if (t < N1) {
#pragma unroll
        for (int m2 = 0; m2 < N2; ++m2) w[m2] = okl ? src[t + N1 * m2] : make_float2(0.f, 0.f);
        rdft<N2, 1>(w);
        if (TWB) __builtin_amdgcn_sched_barrier(0);
        const float2* twi = tw + 2 * N1 * N2 + t * N2;
#pragma unroll
        for (int n2 = 1; n2 < N2; ++n2) w[n2] = cmulc(w[n2], twi[n2]);
    }
    line_exchange<SPLIT, N2, N1>(
        w, t < N1, res, t < N2, [&](int n2) { return n2 * PN + t; }, [&](int m1) { return t * PN + m1; }, L);
    if (t < N2) rdft<N1, 1>(res);
    __syncthreads();
}

// forward x transform: thread t < N2 holds x[N2 n1 + t] in u[n1]; writes
// X[k] to dst (natural order).
template <int N1, int N2, int PN, int SPLIT, int TWB>
__device__ __forceinline__ void x_forward(float2* __restrict__ u, bool okl, int t,
                                          void* L, const float2* __restrict__ tw,
                                          float2* __restrict__ w, float2* dst)
{
    if (t < N2) {
        rdft<N1, -1>(u);
        if (TWB) __builtin_amdgcn_sched_barrier(0);
        const float2* twt = tw + N1 * N2 + t * N1;
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1) u[k1] = cmul(u[k1], twt[k1]);
    }
    line_exchange<SPLIT, N1, N2>(
        u, t < N2, w, t < N1, [&](int k1) { return t * PN + k1; }, [&](int n2) { return n2 * PN + t; }, L);
    if (t < N1) {
        rdft<N2, -1>(w);
        if (okl) {
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2) dst[t + N1 * k2] = w[k2];
        }
    }
}

// g^e for a wave-uniform e in [0, 3] (products of g, as the reference
// applies g once per order); the host keeps every exponent in that range
__device__ __forceinline__ float gpow(float g, int e)
{
    const float g1 = (e & 1) ? g : 1.f;
    return (e & 2) ? g1 * (g * g) : g1;
}

// Pass 4's update of one element x of a line: lw = L W_j there (v/R times the
// unnormalised inverse x transform; the symbol carried 1/N).  Returns W_{j+1}
// (W_1 in ORDER0*, the new state u in LAST*) and updates the accumulator
// where the mode touches it; win and acc point at the line.  Win and Wout
// may be the same buffer: the caller reads win[x] here before it stores
// W_{j+1} at x, and no other thread touches x.  Shared by k_xpass and the
// runtime-length k_gen_xpass.
template <int MODE>
__device__ __forceinline__ float2 x_update(const XArgs& a, float2 lw, int x, float gyz, const float2* win,
                                           float2* __restrict__ acc)
{
    float2 wn;
    if (MODE == XM_ORDER0 || MODE == XM_ORDER0N) {
        // W1 = -i L W0
        wn = make_float2(lw.y, -lw.x);
        if (MODE == XM_ORDER0) {
            const float2 wp = win[x];
            acc[x] = make_float2(a.c0 * wp.x + a.c1 * wn.x, a.c0 * wp.y + a.c1 * wn.y);
        }
    } else {
        // W_{j+1} = W_{j-1} - 2i L W_j
        const float2 wp = win[x];
        wn = make_float2(wp.x + 2.f * lw.y, wp.y - 2.f * lw.x);
        if (MODE != XM_SKIP) {
            const float g = gyz * a.gx[x];
            const float gp = gpow(g, a.eb) * a.cp, gn = gpow(g, a.ec) * a.c1;
            float2 an = make_float2(gp * wp.x + gn * wn.x, gp * wp.y + gn * wn.y);
            if (MODE != XM_MIDF) {
                const float2 ao = acc[x];
                const float ga = gpow(g, a.ea);
                an.x += ga * ao.x;
                an.y += ga * ao.y;
            }
            if (MODE == XM_MID || MODE == XM_MIDF) acc[x] = an;
            if (MODE == XM_LAST || MODE == XM_LASTA)
                wn = an;   // the new state u; its x transform is X_0 of the next step
        }
    }
    return wn;
}

// Line block and shot of an x-pass workgroup (grid: line blocks x shots).
// The hardware deals workgroups round-robin over 8 XCDs by linear id, so id
// runs on XCD id % 8; that XCD takes the contiguous range of logical blocks
// starting at (id % 8) q + min(id % 8, r) (q = blocks / 8, r = blocks % 8)
// instead of every 8th one.  The shot is the fastest logical index: the
// shots of one line block run back to back on one XCD, so the model's v/R
// lines they share come from that XCD's L2 after the first shot's read.
struct XBlock {
    unsigned blk, shot;
};

__device__ __forceinline__ XBlock x_block()
{
    const unsigned ns = gridDim.y, n = gridDim.x * ns;
    const unsigned id = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned q = n / 8, r = n % 8, xc = id % 8;
    const unsigned lg = xc * q + min(xc, r) + id / 8;
    if (ns == 1) return {lg, 0u};
    const unsigned blk = lg / ns;
    return {blk, lg - blk * ns};
}

template <int N1, int N2, int P, int MODE, int SPLIT, int TWB>
__global__ __launch_bounds__((XCfg<N1, N2, P, SPLIT>::THREADS))
void k_xpass(XArgs a)
{
    using C = XCfg<N1, N2, P, SPLIT>;
    constexpr int N = C::N;
    extern __shared__ float2 lds[];
    const int l = threadIdx.x / C::NT;
    const int t = threadIdx.x % C::NT;
    const XBlock xb = x_block();
    const long line = (long)xb.blk * P + l;
    const bool okl = line < a.nlines;
    const int y = okl ? (int)(line % a.ny) : 0;
    const int z = okl ? (int)(line / a.ny) : 0;
    const long lbase = (long)xb.shot * a.shot_stride + line * N;
    const long vbase = line * N;
    // this line's exchange area: LINE float2 elements, or LINE floats when split
    void* L = SPLIT ? (void*)(reinterpret_cast<float*>(lds) + l * C::LINE) : (void*)(lds + l * C::LINE);
    float2 u[N1];
    float2 w[C::NT];

    if (MODE == XM_FWD) {
        if (t < N2 && okl) {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) u[n1] = a.A[lbase + N2 * n1 + t];
        }
        x_forward<N1, N2, C::PN, SPLIT, TWB>(u, okl, t, L, a.twf, w, a.Xout + lbase);
        return;
    }
    if (MODE == XM_INV) {
        x_inverse<N1, N2, C::PN, SPLIT, TWB>(a.Xprev + lbase, okl, t, L, a.tw, w, u);
        if (t < N2 && okl) {
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1)
                a.A[lbase + N2 * n1 + t] = make_float2(u[n1].x * a.inv_nx, u[n1].y * a.inv_nx);
        }
        return;
    }

    // L W_j / (v/R) in physical x (unnormalised inverse; the symbol carried 1/N)
    x_inverse<N1, N2, C::PN, SPLIT, TWB>(a.S + lbase, okl, t, L, a.tw, w, u);

    if (t < N2 && okl) {
        const float gyz = a.gy[y] * a.gz[z];
        int itv = 0;
        bool rec = false;
        if (MODE == XM_LAST || MODE == XM_LASTA) {
            itv = *a.it;
            rec = (z == a.rec_z) && (y >= a.rec_y0) && (y < a.rec_y0 + a.rec_ny) && (itv < a.nt_cap);
        }
        // Win and Wout may be the same buffer (each element is read and then
        // written by this thread only): every Win load precedes the Wout
        // stores, which are issued after the loop; A never aliases Win.
        const float2* win = a.Win + lbase;
        float2* __restrict__ acc = a.A + lbase;
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) {
            const int x = N2 * n1 + t;
            const float vr = a.vr[vbase + x];
            u[n1] = x_update<MODE>(a, make_float2(vr * u[n1].x, vr * u[n1].y), x, gyz, win, acc);
        }
        float2* wout = a.Wout + lbase;
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) wout[N2 * n1 + t] = u[n1];
        if ((MODE == XM_LAST || MODE == XM_LASTA) && rec) {
            // receivers: Im(u) on the plane z = rec_z
            float* tr = a.traces + (long)xb.shot * a.trace_shot_stride + ((long)itv * a.rec_ny + (y - a.rec_y0)) * a.rec_nx;
#pragma unroll
            for (int n1 = 0; n1 < N1; ++n1) {
                const int x = N2 * n1 + t;
                if (x >= a.rec_x0 && x < a.rec_x0 + a.rec_nx) tr[x - a.rec_x0] = u[n1].y;
            }
        }
    }
    if (MODE != XM_LASTA) x_forward<N1, N2, C::PN, SPLIT, TWB>(u, okl, t, L, a.twf, w, a.Xout + lbase);
}

// ---------------------------------------------------------------------
// Source injection in x-spectral form and the step counter.
// u[src] += s[it] + i h[it] before each step (ose_3d.cu:252-253,
// add_wavelet_Xt ose_3d_kernel.cu:291-297) becomes, on X_0 = FFT_x(u),
// X_0[line][k] += (s + i h) exp(-2 pi i k sx / nx), and W_0[src] += s + i h.
// ---------------------------------------------------------------------
struct InjectArgs {
    float2* X0;
    float2* W0;             // the same state in physical form
    long shot_stride;
    int nx;
    const long* src_line;   // per shot: (sz*ny + sy) * nx
    const int* src_x;       // per shot
    const float* s;         // [shot][nt]
    const float* h;
    int nt;
    const int* it;
    const float2* tw;       // x root table
};

// Extended (separable) source profiles: shot k adds (s + i h) px[x] py[y]
// pz[z] over the box [x0, x0+wx) x [y0, y0+wy) x [z0, z0+wz) to W_0 and, on
// the box's (y, z) lines, (s + i h) py pz PX[kx] to X_0, PX the x transform
// of px.  A point source is the unit profile.
constexpr int PROF_W = 64;                // widest support along one axis
struct ProfDesc {
    int x0, y0, z0, wx, wy, wz;
};
struct ProfArgs {
    float2* X0;
    float2* W0;
    long shot_stride;
    int nx, ny;
    const ProfDesc* desc;   // [shot]
    const float2* PX;       // [shot][nx]
    const float* pv;        // [shot][3][PROF_W]: px | py | pz
    const float* s;         // [shot][nt]
    const float* h;
    int nt;
    const int* it;
    int wy_max;             // lines per shot in the launch: wy_max * wz_max
};

__global__ void k_inject(InjectArgs a);
__global__ void k_inject_prof(ProfArgs a);
__global__ void k_advance(int* it);
__global__ void k_scale_vr(float2* f, const float* vr, long n, long shot_stride, float s);

}  // namespace ose
