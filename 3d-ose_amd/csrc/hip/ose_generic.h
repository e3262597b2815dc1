// ose_generic.h -- runtime-length pass kernels for FFT extents without
// compiled register kernels.
//
// The reference plans its 3D transform for whatever padded extent the odd
// rule produces (cufftMakePlan3d(nzb, nyb, nxb), ose_3d.cu:226; judge_odd,
// ose_3d_lib.cu:36-44).  The compiled kernels of ose_kernels.h cover the
// extents N = N1 N2 with both factors in registers; every other extent up to
// GEN_NMAX takes these kernels, which do the same four passes with the same
// arguments and the same results in natural order, so the two kinds mix
// freely across the axes of one model.
//
// A workgroup loads W pencils (strided passes, W x columns) or W x lines
// into LDS, one element per thread and row at a time (coalesced along x),
// and transforms them in place with the self-sorting mixed-radix scheme
// (Stockham):  for stages of radix R with p = the product of the radices
// before it, butterfly i in [0, N/R), k = i mod p, reads x[i + s N/R],
// multiplies by w^(s k N/(p R)) (w = exp(-2 pi i / N), the axis's FP64 root
// table), takes an R-point DFT and writes it to [(i - k) R + k + t p].
// Each stage reads everything it needs into registers, waits, writes, waits,
// so one LDS buffer suffices.  Radices 2, 3, 4, 5, 7, 8, 11, 13 and 16 use the
// straight-line register DFTs of dft_gen.h; a larger prime factor R is a
// direct stage: output pairs (s, R - s) sum the R inputs' symmetric and
// antisymmetric combinations with table cosines and sines (O(R) per element,
// no recurrence, so no drift in FP32).  The twiddles come from a
// copy of the root table at the head of the workgroup's LDS.
#pragma once
#include "ose_kernels.h"

namespace ose {

constexpr int GEN_MAXV = 16;   // values per thread per stage

// The workgroup's LDS: the axis's root table w^j (j < N) first, then the
// pencils or lines.  Declared once here so that the stage functions, which
// are not inlined (each keeps its own register allocation, so the kernel
// needs the largest stage's registers, not the sum of all), address it as
// LDS.
extern __shared__ float2 gen_lds[];

template <int SIGN>
__device__ __forceinline__ float2 gtw(int e)
{
    const float2 w = gen_lds[e];
    return SIGN < 0 ? w : make_float2(w.x, -w.y);
}

// LDS element e of the pencil addressed by (pitch, off)
struct GenAt {
    int pitch, off;
    __device__ __forceinline__ int operator()(int e) const { return e * pitch + off; }
};

template <int R, int SIGN>
__device__ __noinline__ void gen_stage(GenAt at, int N, int TP, int p, int t)
{
    constexpr int NB = GEN_MAXV / R;   // butterflies per thread
    float2* L = gen_lds;
    const int M = N / R;
    const int tws = N / (p * R);
    float2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int i = t + b * TP;
        if (i < M) {
            const int k = i % p;
#pragma unroll
            for (int s = 0; s < R; ++s) v[b][s] = L[at(i + s * M)];
#pragma unroll
            for (int s = 1; s < R; ++s) v[b][s] = cmul(v[b][s], gtw<SIGN>(s * k * tws));
            rdft<R, SIGN>(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int i = t + b * TP;
        if (i < M) {
            const int k = i % p;
            const int j = (i - k) * R + k;
#pragma unroll
            for (int s = 0; s < R; ++s) L[at(j + s * p)] = v[b][s];
        }
    }
    __syncthreads();
}

// A stage of prime radix R > 13.  Butterfly i (k = i mod p) twiddles its
// inputs, y_q = x[i + q N/R] w^(q k N/(p R)), and forms the R-point DFT in
// the pair-symmetric way: with a_q = y_q + y_(R-q), b_q = y_q - y_(R-q)
// (q = 1..h, h = (R-1)/2) and the angle t = 2 pi q s / R,
//   X_0 = y_0 + sum_q a_q,   X_s, X_(R-s) = y_0 + sum_q a_q cos t -/+ i S sum_q b_q sin t
// (S = +1 forward, -1 inverse), written to j + s p, j = (i - k) R + k.  A
// thread takes a group of GEN_GP output pairs of one butterfly (s = 0 stands
// for X_0 alone), so each a_q, b_q serves GEN_GP pairs; cos and sin come
// from the root table at (q s mod R) N/R (no recurrence).  Half the products
// and a quarter of the table reads of the plain sum over all R outputs.
constexpr int GEN_GP = 4;   // output pairs per thread group

template <int SIGN>
__device__ __noinline__ void gen_direct(GenAt at, int N, int TP, int p, int R, int t)
{
    constexpr int NB = GEN_MAXV / (2 * GEN_GP);   // groups per thread
    float2* L = gen_lds;
    const int M = N / R;
    const int tws = N / (p * R);
    const int h = (R - 1) / 2;
    const int ng = (h + GEN_GP) / GEN_GP;          // groups over s = 0..h
    float2 v[NB][2 * GEN_GP];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int u = t + b * TP;
        if (u < M * ng) {
            const int i = u / ng, s0 = (u - i * ng) * GEN_GP;
            const int k = i % p;
            float2 A[GEN_GP], B[GEN_GP];
#pragma unroll
            for (int g = 0; g < GEN_GP; ++g) A[g] = B[g] = make_float2(0.f, 0.f);
            const float2 y0 = L[at(i)];
            for (int q = 1; q <= h; ++q) {
                const float2 y1 = cmul(L[at(i + q * M)], gtw<SIGN>(q * k * tws));
                const float2 y2 = cmul(L[at(i + (R - q) * M)], gtw<SIGN>((R - q) * k * tws));
                const float2 aq = make_float2(y1.x + y2.x, y1.y + y2.y);
                const float2 bq = make_float2(y1.x - y2.x, y1.y - y2.y);
#pragma unroll
                for (int g = 0; g < GEN_GP; ++g) {
                    const int s = min(s0 + g, h);
                    const float2 w = gen_lds[((q * s) % R) * M];   // (cos t, -sin t)
                    A[g].x += aq.x * w.x;
                    A[g].y += aq.y * w.x;
                    B[g].x -= bq.x * w.y;
                    B[g].y -= bq.y * w.y;
                }
            }
#pragma unroll
            for (int g = 0; g < GEN_GP; ++g) {
                // i S B = (-S B.y, S B.x)
                const float bx = SIGN < 0 ? -B[g].y : B[g].y, by = SIGN < 0 ? B[g].x : -B[g].x;
                const float2 c = make_float2(y0.x + A[g].x, y0.y + A[g].y);
                v[b][2 * g] = make_float2(c.x - bx, c.y - by);       // X_s
                v[b][2 * g + 1] = make_float2(c.x + bx, c.y + by);   // X_(R-s)
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int u = t + b * TP;
        if (u < M * ng) {
            const int i = u / ng, s0 = (u - i * ng) * GEN_GP;
            const int k = i % p;
            const int j = (i - k) * R + k;
#pragma unroll
            for (int g = 0; g < GEN_GP; ++g) {
                const int s = s0 + g;
                if (s == 0) {
                    L[at(j)] = v[b][0];   // X_0 (its pair partner is unused)
                } else if (s <= h) {
                    L[at(j + s * p)] = v[b][2 * g];
                    L[at(j + (R - s) * p)] = v[b][2 * g + 1];
                }
            }
        }
    }
    __syncthreads();
}

// unnormalised DFT (SIGN -1 forward, +1 inverse) of one pencil in LDS, natural
// order in and out; every thread of the workgroup calls it (barriers inside)
template <int SIGN>
__device__ __forceinline__ void gen_fft(GenAt at, const GenPlan& g, int t)
{
    int p = 1;
    for (int f = 0; f < g.nf; ++f) {
        const int R = g.r[f];
        switch (R) {
        case 2: gen_stage<2, SIGN>(at, g.N, g.TP, p, t); break;
        case 3: gen_stage<3, SIGN>(at, g.N, g.TP, p, t); break;
        case 4: gen_stage<4, SIGN>(at, g.N, g.TP, p, t); break;
        case 5: gen_stage<5, SIGN>(at, g.N, g.TP, p, t); break;
        case 7: gen_stage<7, SIGN>(at, g.N, g.TP, p, t); break;
        case 8: gen_stage<8, SIGN>(at, g.N, g.TP, p, t); break;
        case 11: gen_stage<11, SIGN>(at, g.N, g.TP, p, t); break;
        case 13: gen_stage<13, SIGN>(at, g.N, g.TP, p, t); break;
        case 16: gen_stage<16, SIGN>(at, g.N, g.TP, p, t); break;
        default: gen_direct<SIGN>(at, g.N, g.TP, p, R, t); break;
        }
        p *= R;
    }
}

// the axis's root table into the head of the LDS (the data's load waits for it)
__device__ __forceinline__ void gen_load_table(const float2* __restrict__ tw, int N)
{
    for (int j = threadIdx.x; j < N; j += blockDim.x) gen_lds[j] = tw[j];
}

// the runtime-length strided pass: W = T x columns per workgroup, TP threads
// per column; LDS rows of T + 1 elements after the root table
struct GenStrided {
    int T, col, t, x0, xc, N;
    long base;
    __device__ __forceinline__ GenStrided(const StridedArgs& a, const StridedTile& g)
    {
        T = a.g.W;
        N = a.g.N;
        col = threadIdx.x % T;
        t = threadIdx.x / T;
        x0 = g.bx * T;
        xc = min(col, a.ncol - 1 - x0);   // columns past the edge load the last one, store nothing
        base = (long)g.bz * a.shot_stride + (long)g.by * a.os + x0 + xc;
    }
    __device__ __forceinline__ GenAt at() const { return GenAt{T + 1, N + col}; }
    __device__ __forceinline__ void load(const StridedArgs& a) const
    {
        gen_load_table(a.tw, N);
        const GenAt e = at();
        for (int r = t; r < N; r += a.g.TP) gen_lds[e(r)] = a.in[base + (long)r * a.es];
        __syncthreads();
    }
    __device__ __forceinline__ void store(const StridedArgs& a) const
    {
        if (col != xc) return;
        const GenAt e = at();
        for (int r = t; r < N; r += a.g.TP) a.out[base + (long)r * a.es] = gen_lds[e(r)];
    }
};

template <int SIGN>
__global__ __launch_bounds__(512) void k_gen_strided(StridedArgs a)
{
    const StridedTile g(a);
    const GenStrided s(a, g);
    s.load(a);
    gen_fft<SIGN>(s.at(), a.g, s.t);
    s.store(a);
}

// z pass: forward-z, * |k| / N (scale_kxkz_3D, ose_3d_kernel.cu:101-126, and
// the 1/N of scal_vr_3d, :355), inverse-z, as k_zsym
__global__ __launch_bounds__(512) void k_gen_zsym(StridedArgs a)
{
    const StridedTile g(a);
    const GenStrided s(a, g);
    s.load(a);
    gen_fft<-1>(s.at(), a.g, s.t);
    const float kxy = a.kx2[s.x0 + s.xc] + a.ky2[g.by];
    const GenAt e = s.at();
    for (int r = s.t; r < s.N; r += a.g.TP) {
        float2& v = gen_lds[e(r)];
        const float m = __builtin_amdgcn_sqrtf(kxy + a.kz2[r]) * a.invN;
        v.x *= m;
        v.y *= m;
    }
    __syncthreads();
    gen_fft<1>(s.at(), a.g, s.t);
    s.store(a);
}

// x pass: W = P lines per workgroup, TP threads per line; the same modes,
// arguments and per-element update (x_update) as k_xpass
template <int MODE>
__global__ __launch_bounds__(512) void k_gen_xpass(XArgs a)
{
    const int N = a.g.N, P = a.g.W, TP = a.g.TP;
    const int l = threadIdx.x / TP;
    const int t = threadIdx.x % TP;
    // XCD-contiguous line blocks, shots fastest, as k_xpass
    const XBlock xb = x_block();
    const long line = (long)xb.blk * P + l;
    const bool okl = line < a.nlines;
    const int y = okl ? (int)(line % a.ny) : 0;
    const int z = okl ? (int)(line / a.ny) : 0;
    const long lbase = (long)xb.shot * a.shot_stride + line * N;
    const long vbase = line * N;
    // the root table, then P lines
    float2* L = gen_lds + N + (long)l * N;
    const GenAt at{1, N + l * N};
    const float2 zero = make_float2(0.f, 0.f);
    gen_load_table(a.tw, N);   // the forward table (twf) holds the same roots

    if (MODE == XM_FWD || MODE == XM_INV) {
        const float2* src = MODE == XM_FWD ? a.A : a.Xprev;
        for (int x = t; x < N; x += TP) L[x] = okl ? src[lbase + x] : zero;
        __syncthreads();
        if (MODE == XM_FWD) {
            gen_fft<-1>(at, a.g, t);
            if (okl)
                for (int x = t; x < N; x += TP) a.Xout[lbase + x] = L[x];
        } else {
            gen_fft<1>(at, a.g, t);
            if (okl)
                for (int x = t; x < N; x += TP) a.A[lbase + x] = make_float2(L[x].x * a.inv_nx, L[x].y * a.inv_nx);
        }
        return;
    }

    // L W_j / (v/R) in physical x (unnormalised inverse; the symbol carried 1/N)
    for (int x = t; x < N; x += TP) L[x] = okl ? a.S[lbase + x] : zero;
    __syncthreads();
    gen_fft<1>(at, a.g, t);

    if (okl) {
        const float gyz = a.gy[y] * a.gz[z];
        int itv = 0;
        bool rec = false;
        if (MODE == XM_LAST || MODE == XM_LASTA) {
            itv = *a.it;
            rec = (z == a.rec_z) && (y >= a.rec_y0) && (y < a.rec_y0 + a.rec_ny) && (itv < a.nt_cap);
        }
        // Win and Wout may be the same buffer: each element is read and then
        // written by this thread only
        const float2* win = a.Win + lbase;
        float2* acc = a.A + lbase;
        float* tr = rec ? a.traces + (long)xb.shot * a.trace_shot_stride +
                              ((long)itv * a.rec_ny + (y - a.rec_y0)) * a.rec_nx
                        : nullptr;
        for (int x = t; x < N; x += TP) {
            const float2 u = L[x];
            const float vr = a.vr[vbase + x];
            const float2 wn = x_update<MODE>(a, make_float2(vr * u.x, vr * u.y), x, gyz, win, acc);
            L[x] = wn;
            a.Wout[lbase + x] = wn;
            if (rec && x >= a.rec_x0 && x < a.rec_x0 + a.rec_nx) tr[x - a.rec_x0] = wn.y;
        }
    } else {
        for (int x = t; x < N; x += TP) L[x] = zero;
    }
    if (MODE == XM_LASTA) return;
    __syncthreads();
    gen_fft<-1>(at, a.g, t);
    if (okl)
        for (int x = t; x < N; x += TP) a.Xout[lbase + x] = L[x];
}

}  // namespace ose
