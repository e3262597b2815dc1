// ose_bnet.hip -- kernels of the learned thin-pad boundary network and the
// training object behind ose_bnet_* (include/ose_api.h).  The reference has
// no learned boundary (its absorbing layer is generate_att/apply_ABC_Xt,
// ose_3d_kernel.cu:242-288); the network semantics are the ones the FP64
// checker (oracle/bnet_ref.py) restates.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../../include/ose_api.h"
#include "ose_bnet.h"

namespace ose {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float x)
{
    const float k = 0.7978845608028654f;
    return 0.5f * x * (1.f + tanhf(k * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x)
{
    const float k = 0.7978845608028654f;
    const float t = tanhf(k * (x + 0.044715f * x * x * x));
    return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * k * (1.f + 3.f * 0.044715f * x * x);
}

template <int N>
__device__ __forceinline__ void block_sum(float (&v)[N], float* red)
{
    // red: shared [N][256/64]
#pragma unroll
    for (int k = 0; k < N; ++k) {
        float x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        v[k] = x;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) red[k * 4 + w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = red[k * 4] + red[k * 4 + 1] + red[k * 4 + 2] + red[k * 4 + 3];
}

// ---------------------------------------------------------------------
// statistics, features
// ---------------------------------------------------------------------
__global__ void k_bn_zero(float* p, int n, int* flag, int clear_flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
    if (clear_flag && i == 0 && flag) *flag = 0;
}

// stats[shot*4 + 0] += sum |u|^2 over R_3
__global__ __launch_bounds__(256) void k_bn_norm(const float2* u, long ncell, Shell s3, float* stats)
{
    __shared__ float red[4];
    const int shot = blockIdx.y;
    float acc[1] = {0.f};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < s3.count; i += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        shell_coords(s3, i, x, y, z);
        const float2 v = u[(long)shot * ncell + ((long)z * s3.ny + y) * s3.nx + x];
        acc[0] += v.x * v.x + v.y * v.y;
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) atomicAdd(&stats[shot * 4 + 0], acc[0]);
}

struct FeatArgs {
    const float2* u;
    long ncell;
    const float* vr;
    float vscale;      // R / vmax: v / vmax = vr * vscale
    Shell s3;
    int bl, bf, bu, nx, ny, nz;
    const float* stats;
    uint16_t* feat;
    long feat_shot;
};

__device__ __forceinline__ void axis_feat(int c, int lo, int n, int nb, float& dist, float& side)
{
    const float w = (float)(nb - n);
    if (c >= lo + n) {
        dist = (float)(c - (lo + n - 1)) / w;
        side = 1.f;
    } else if (c < lo) {
        dist = (float)(lo - c) / w;
        side = -1.f;
    } else {
        dist = 0.f;
        side = 0.f;
    }
}

__global__ __launch_bounds__(256) void k_bn_features(FeatArgs a)
{
    const int shot = blockIdx.y;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.s3.count) return;
    int x, y, z;
    shell_coords(a.s3, i, x, y, z);
    const long cell = ((long)z * a.s3.ny + y) * a.s3.nx + x;
    const float sum = a.stats[shot * 4 + 0];
    const float s = sqrtf(sum / (float)a.s3.count);
    const float inv = s > 0.f ? 1.f / s : 0.f;
    const float2 v = a.u[(long)shot * a.ncell + cell];
    float f[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) f[k] = 0.f;
    f[0] = v.x * inv;
    f[1] = v.y * inv;
    f[2] = a.vr[cell] * a.vscale;
    axis_feat(x, a.bl, a.nx, a.s3.nx, f[3], f[4]);
    axis_feat(y, a.bf, a.ny, a.s3.ny, f[5], f[6]);
    axis_feat(z, a.bu, a.nz, a.s3.nz, f[7], f[8]);
    uint16_t* o = a.feat + (long)shot * a.feat_shot + i * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) o[k] = f2bf(f[k]);
}

// ---------------------------------------------------------------------
// forward convolution: out[c][o] = sum_tap sum_i in[c + off(tap)][i] W[tap][o][i]
// A = activations (16 cells x 16 in), B = W[tap] (16 in x 16 out), one
// MFMA 16x16x16 per tap; a wave owns 16 output cells at a time.
// ---------------------------------------------------------------------
enum { EPI_HIDDEN = 0, EPI_OUT = 1 };

struct ConvArgs {
    const uint16_t* in;
    long in_shot;
    Shell sin, sout;
    const uint16_t* W;      // [27][16 out][16 in]
    const float* bias;      // [16]
    uint16_t* hout;         // EPI_HIDDEN: gelu(y) bf16 [n_out][16]
    long h_shot;
    float* yout;            // EPI_HIDDEN: y fp32 [n_out][16] (optional); EPI_OUT: [n_out][2]
    long y_shot;
};

// Two taps per MFMA (16x16x32, K = 2 taps x 16 channels): lane group g
// (lanes 16g..16g+15) supplies channels 8(g%2)..8(g%2)+7 of tap 2p + g/2 for
// the A operand (one 16-byte load) and the matching weights for B.  Tap 27
// does not exist: its weights are zero.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
constexpr int BN_PAIRS = (BN_TAPS + 1) / 2;

__device__ __forceinline__ int tap_offset(int tap, int oy, int oz)
{
    return (tap / 9 - 1) * oz + ((tap / 3) % 3 - 1) * oy + (tap % 3 - 1);
}

template <int EPI>
__global__ __launch_bounds__(256, 4) void k_bn_conv(ConvArgs a)
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int r16 = lane & 15;
    const int half = g >> 1;                 // which tap of the pair
    const int ch0 = 8 * (g & 1);             // first channel of this lane
    const int shot = blockIdx.y;
    // the layer's weights in LDS (tap 27 zero), read per MFMA: holding all
    // 14 pairs in registers took 56 VGPRs and capped the kernel at 3 waves
    __shared__ __align__(16) uint16_t wl[2 * BN_PAIRS * 16 * 16];
    for (int i = threadIdx.x; i < 2 * BN_PAIRS * 16 * 16; i += blockDim.x)
        wl[i] = i < BN_WSZ ? a.W[i] : (uint16_t)0;
    __syncthreads();
    auto wb = [&](int p) { return *(const s16x8*)(wl + ((2 * p + half) * 16 + r16) * 16 + ch0); };
    const float bias = a.bias[r16];
    const long ntiles = (a.sout.count + 15) / 16;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    const uint16_t* in = a.in + (long)shot * a.in_shot;
    const bool fast_ok = a.sin.tz == a.sout.tz + 2 && a.sin.ty == a.sout.ty + 2 && a.sin.tx == a.sout.tx + 2;
    for (long tile = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < ntiles; tile += nwaves) {
        const long cell = tile * 16 + r16;
        const bool valid = cell < a.sout.count;
        NbFast nf{0u, 0, 0, false};
        int x = 0, y = 0, z = 0;
        if (valid) {
            nf = nb_fast(a.sout, a.sin, (unsigned)cell, fast_ok);
            if (!nf.fast) shell_coords(a.sout, cell, x, y, z);
        }
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        if (__all(nf.fast || !valid)) {
            // the whole tile on constant offsets: the fourteen pair-loads are
            // issued together, then fourteen MFMAs
            const unsigned base = valid ? nf.base : 0u;
            s16x8 av[BN_PAIRS];
#pragma unroll
            for (int p = 0; p < BN_PAIRS; ++p) {
                const int tap = 2 * p + half;
                const bool ok = valid && tap < BN_TAPS;
                const unsigned nb = ok ? base + (unsigned)tap_offset(tap, nf.oy, nf.oz) : 0u;
                av[p] = *(const s16x8*)(in + (size_t)nb * 16 + ch0);
                if (!ok) av[p] = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
#pragma unroll
            for (int p = 0; p < BN_PAIRS; ++p)
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[p]),
                                                              __builtin_bit_cast(bf16x8, wb(p)), acc, 0, 0, 0);
        } else {
#pragma unroll
            for (int p = 0; p < BN_PAIRS; ++p) {
                const int tap = 2 * p + half;
                s16x8 av = {0, 0, 0, 0, 0, 0, 0, 0};
                if (valid && tap < BN_TAPS) {
                    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                    long nb;
                    if (nf.fast)
                        nb = (long)(nf.base + (unsigned)(dz * nf.oz + dy * nf.oy + dx));
                    else
                        nb = shell_index(a.sin, wrapc(x + dx, a.sin.nx), wrapc(y + dy, a.sin.ny),
                                         wrapc(z + dz, a.sin.nz));
                    if (nb >= 0) av = *(const s16x8*)(in + nb * 16 + ch0);
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                              __builtin_bit_cast(bf16x8, wb(p)), acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long oc = tile * 16 + 4 * g + j;
            if (oc >= a.sout.count) continue;
            const float y = acc[j] + bias;
            if (EPI == EPI_HIDDEN) {
                a.hout[(long)shot * a.h_shot + oc * 16 + r16] = f2bf(gelu(y));
                if (a.yout) a.yout[(long)shot * a.y_shot + oc * 16 + r16] = y;
            } else if (r16 < 2) {
                a.yout[(long)shot * a.y_shot + oc * 2 + r16] = y;
            }
        }
    }
}

// ---------------------------------------------------------------------
// guard + apply: u' = u G(y) on the pad unless the pad energy grows by more
// than `guard` or turns non-finite; the flag is sticky.  (|G| <= 1, so the
// guard only catches non-finite outputs.)
// stats[shot*4 + 1] += sum |u G|^2, stats[shot*4 + 2] += sum |u|^2 (R_0)
// passive gain G = q / max(1, |q|), q = 1 + a + i b, applied to v
__device__ __forceinline__ float2 gain_q(float a, float b)
{
    float qx = 1.f + a, qy = b;
    const float m2 = qx * qx + qy * qy;
    if (m2 > 1.f) {
        const float r = rsqrtf(m2);
        qx *= r;
        qy *= r;
    }
    return make_float2(qx, qy);
}

__device__ __forceinline__ float2 pad_gain(float2 v, float a, float b)
{
    const float2 g = gain_q(a, b);
    return make_float2(v.x * g.x - v.y * g.y, v.y * g.x + v.x * g.y);
}

__global__ __launch_bounds__(256) void k_bn_guard(const float2* u, long ncell, Shell s0, const float* y3, long y_shot,
                                                  float* stats)
{
    __shared__ float red[8];
    const int shot = blockIdx.y;
    float acc[2] = {0.f, 0.f};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < s0.count; i += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        shell_coords(s0, i, x, y, z);
        const float2 v = u[(long)shot * ncell + ((long)z * s0.ny + y) * s0.nx + x];
        const float2 w = pad_gain(v, y3[(long)shot * y_shot + i * 2], y3[(long)shot * y_shot + i * 2 + 1]);
        acc[0] += w.x * w.x + w.y * w.y;
        acc[1] += v.x * v.x + v.y * v.y;
    }
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
        atomicAdd(&stats[shot * 4 + 1], acc[0]);
        atomicAdd(&stats[shot * 4 + 2], acc[1]);
    }
}

__global__ __launch_bounds__(256) void k_bn_apply(float2* u, long ncell, Shell s0, const float* y3, long y_shot,
                                                  const float* stats, int nshot, float guard, int* tripped)
{
    // every block evaluates the same guard decision
    bool bad = *tripped != 0;
    for (int k = 0; k < nshot && !bad; ++k) {
        // energy growth of the pad beyond `guard`, ignoring pads that hold
        // less than 1e-4 of the receptive shell's energy (leakage noise)
        const float post = stats[k * 4 + 1], pre = stats[k * 4 + 2], shell = stats[k * 4 + 0];
        if (!(post <= guard * pre + 1e-4f * shell + 1e-30f)) bad = true;
    }
    if (bad) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *tripped = 1;
        return;
    }
    const int shot = blockIdx.y;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < s0.count; i += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        shell_coords(s0, i, x, y, z);
        float2* p = u + (long)shot * ncell + ((long)z * s0.ny + y) * s0.nx + x;
        *p = pad_gain(*p, y3[(long)shot * y_shot + i * 2], y3[(long)shot * y_shot + i * 2 + 1]);
    }
}

// taper of a learned-mode step: 1 while the network is on, the classical
// taper once the guard has tripped (device-side, so the switch takes effect
// from the next step even inside one graph replay)
__global__ __launch_bounds__(256) void k_bn_taper(float* eff, const float* gx, int nx, const float* gy, int ny,
                                                  const float* gz, int nz, const int* tripped)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool tr = *tripped != 0;
    if (i < nx) eff[i] = tr ? gx[i] : 1.f;
    else if (i < nx + ny) eff[i] = tr ? gy[i - nx] : 1.f;
    else if (i < nx + ny + nz) eff[i] = tr ? gz[i - nx - ny] : 1.f;
}

// ---------------------------------------------------------------------
// training kernels
// ---------------------------------------------------------------------
// loss on the pad: e = (u/s) G(y) - t/s; for y = a + ib
// dL/da = 2 Re(conj(e) (u/s) dG/da) / M, likewise b (bf16)
__global__ __launch_bounds__(256) void k_bn_loss(const float2* u, const float2* t, Shell s0, Shell s3,
                                                 const float* y3, const float* stats, uint16_t* dy3, float* loss)
{
    __shared__ float red[4];
    const float s = sqrtf(stats[0] / (float)s3.count);
    const float inv = s > 0.f ? 1.f / s : 0.f;
    const float scale = 2.f / (float)s0.count;
    float acc[1] = {0.f};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < s0.count; i += (long)gridDim.x * blockDim.x) {
        int x, y, z;
        shell_coords(s0, i, x, y, z);
        const long cell = ((long)z * s0.ny + y) * s0.nx + x;
        const float2 un = make_float2(u[cell].x * inv, u[cell].y * inv);
        const float2 tn = make_float2(t[cell].x * inv, t[cell].y * inv);
        const float a = y3[i * 2], b = y3[i * 2 + 1];
        const float2 w = pad_gain(un, a, b);
        const float ex = w.x - tn.x, ey = w.y - tn.y;
        acc[0] += ex * ex + ey * ey;
        // dG/da, dG/db: (1, i) inside the unit disc; on the clipped branch
        // G = q/|q|: dG/da = 1/|q| - q Re(q)/|q|^3, dG/db = i/|q| - q Im(q)/|q|^3
        const float qx = 1.f + a, qy = b, m2 = qx * qx + qy * qy;
        float2 ga = make_float2(1.f, 0.f), gb = make_float2(0.f, 1.f);
        if (m2 > 1.f) {
            const float r = rsqrtf(m2), r3 = r * r * r;
            ga = make_float2(r - qx * qx * r3, -qy * qx * r3);
            gb = make_float2(-qx * qy * r3, r - qy * qy * r3);
        }
        // Re(conj(e) u_n dG), e = (ex, ey)
        const float2 ua = make_float2(un.x * ga.x - un.y * ga.y, un.x * ga.y + un.y * ga.x);
        const float2 ub = make_float2(un.x * gb.x - un.y * gb.y, un.x * gb.y + un.y * gb.x);
        uint16_t* o = dy3 + i * 16;
        o[0] = f2bf(scale * (ex * ua.x + ey * ua.y));
        o[1] = f2bf(scale * (ex * ub.x + ey * ub.y));
#pragma unroll
        for (int k = 2; k < 16; ++k) o[k] = 0;
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) atomicAdd(loss, acc[0] / (float)s0.count);
}

// input gradient: dx[c'][i] = sum_tap sum_o dy[c' - off][o] W[tap][o][i], then
// times gelu'(ypre[c'][i]).  A = dy (cells x out), B = WT[tap] (out x in).
struct DgradArgs {
    const uint16_t* dy;
    Shell sdy;              // R_d
    Shell sx;               // R_{d+1}
    const uint16_t* WT;     // [27][16 in][16 out]
    const float* ypre;      // [n_{d+1}][16]
    uint16_t* dprev;        // [n_{d+1}][16]
};

__global__ __launch_bounds__(256) void k_bn_dgrad(DgradArgs a)
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int r16 = lane & 15;
    bf16x4 wb[BN_TAPS];
#pragma unroll
    for (int tap = 0; tap < BN_TAPS; ++tap) wb[tap] = *(const bf16x4*)(a.WT + (tap * 16 + r16) * 16 + 4 * g);
    const long ntiles = (a.sx.count + 15) / 16;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long tile = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < ntiles; tile += nwaves) {
        const long cell = tile * 16 + r16;
        const bool valid = cell < a.sx.count;
        int x = 0, y = 0, z = 0;
        if (valid) shell_coords(a.sx, cell, x, y, z);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < BN_TAPS; ++tap) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            bf16x4 av = {0, 0, 0, 0};
            if (valid) {
                const long nb = shell_index(a.sdy, wrapc(x - dx, a.sdy.nx), wrapc(y - dy, a.sdy.ny),
                                            wrapc(z - dz, a.sdy.nz));
                if (nb >= 0) av = *(const bf16x4*)(a.dy + nb * 16 + 4 * g);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, wb[tap], acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long oc = tile * 16 + 4 * g + j;
            if (oc >= a.sx.count) continue;
            a.dprev[oc * 16 + r16] = f2bf(acc[j] * gelu_grad(a.ypre[oc * 16 + r16]));
        }
    }
}

// weight gradient of one tap: dW[tap][o][i] = sum_{c in R_d} x[c + off][i] dy[c][o]
// A = dy^T (out x cells), B = x (cells x in); K runs over cells, 16 at a time.
struct WgradArgs {
    const uint16_t* x;
    Shell sx;               // R_{d+1}
    const uint16_t* dy;
    Shell sdy;              // R_d
    float* dW;              // [27][16][16]
    long chunk;             // cells per wave
};

__global__ __launch_bounds__(256) void k_bn_wgrad(WgradArgs a)
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int r16 = lane & 15;
    const int tap = blockIdx.y;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long c0 = wave * a.chunk;
    const long c1 = c0 + a.chunk < a.sdy.count ? c0 + a.chunk : a.sdy.count;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long base = c0; base < c1; base += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const long c = base + 4 * g + k;
            short ak = 0, bk = 0;
            if (c < c1) {
                int x, y, z;
                shell_coords(a.sdy, c, x, y, z);
                ak = (short)a.dy[c * 16 + r16];
                const long nb = shell_index(a.sx, wrapc(x + dx, a.sx.nx), wrapc(y + dy, a.sx.ny), wrapc(z + dz, a.sx.nz));
                if (nb >= 0) bk = (short)a.x[nb * 16 + r16];
            }
            av[k] = ak;
            bv[k] = bk;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) atomicAdd(&a.dW[(tap * 16 + 4 * g + j) * 16 + r16], acc[j]);
}

// db[o] += sum_c dy[c][o]
__global__ __launch_bounds__(256) void k_bn_bgrad(const uint16_t* dy, long n, float* db)
{
    __shared__ float red[16 * 4];
    float acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0.f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] += bf2f(dy[i * 16 + k]);
    block_sum<16>(acc, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 16; ++k) atomicAdd(&db[k], acc[k]);
}

__global__ void k_bn_convert(const float* p, uint16_t* W, uint16_t* WT, float* bias)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * BN_WSZ) return;
    const int l = i / BN_WSZ, r = i % BN_WSZ;
    const int tap = r / 256, o = (r / 16) % 16, in = r % 16;
    const float w = p[l * (BN_WSZ + 16) + r];
    W[i] = f2bf(w);
    WT[l * BN_WSZ + (tap * 16 + in) * 16 + o] = f2bf(w);
    if (r < 16) bias[l * 16 + r] = p[l * (BN_WSZ + 16) + BN_WSZ + r];
}

__global__ void k_bn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                          float c1, float c2, float eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------
Shell make_shell(int nxb, int nyb, int nzb, int bl, int bf, int bu, int nx, int ny, int nz, int d)
{
    Shell s{};
    s.nx = nxb;
    s.ny = nyb;
    s.nz = nzb;
    auto band = [d](int n, int lo, int nint, int& b0, int& t) {
        t = (n - nint) + 2 * d;
        if (t >= n) {
            t = n;
            b0 = 0;
        } else {
            b0 = ((lo + nint - d) % n + n) % n;
        }
    };
    band(nxb, bl, nx, s.bx0, s.tx);
    band(nyb, bf, ny, s.by0, s.ty);
    band(nzb, bu, nz, s.bz0, s.tz);
    s.offY = (long)s.tz * nyb * nxb;
    s.offX = s.offY + (long)(nzb - s.tz) * s.ty * nxb;
    s.count = s.offX + (long)(nzb - s.tz) * (nyb - s.ty) * s.tx;
    return s;
}

static inline unsigned blocks_for(long n, int per = 256, long cap = 4096)
{
    long b = (n + per - 1) / per;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

#define LCHK(expr)                                      \
    do {                                                \
        if ((expr) != hipSuccess) return OSE_EDEVICE;   \
    } while (0)

static int conv_layer(hipStream_t st, int epi, const uint16_t* in, long in_shot, const Shell& sin, const Shell& sout,
                      const uint16_t* W, const float* bias, uint16_t* hout, long h_shot, float* yout, long y_shot,
                      int nshot)
{
    ConvArgs a{in, in_shot, sin, sout, W, bias, hout, h_shot, yout, y_shot};
    void* args[] = {&a};
    const long tiles = (sout.count + 15) / 16;
    dim3 grid(blocks_for(tiles, 4, 2048), nshot);
    const void* fn = epi == EPI_HIDDEN ? (const void*)&k_bn_conv<EPI_HIDDEN> : (const void*)&k_bn_conv<EPI_OUT>;
    LCHK(hipLaunchKernel(fn, grid, dim3(256), args, 0, st));
    return OSE_OK;
}

int bnet_forward(hipStream_t st, const BnetGeom& g, const uint16_t* Wbf, const float* bias, float2* u,
                 const float* vr, long ncell, int nshot, BnetBufs& b, float guard, bool apply)
{
    const Shell& s0 = g.sh[0];
    const Shell& s1 = g.sh[1];
    const Shell& s2 = g.sh[2];
    const Shell& s3 = g.sh[3];
    {
        int n = 4 * nshot;
        int clear = 0;
        void* args[] = {&b.stats, &n, &b.tripped, &clear};
        LCHK(hipLaunchKernel((const void*)&k_bn_zero, dim3(1), dim3(256), args, 0, st));
    }
    {
        void* args[] = {&u, &ncell, (void*)&s3, &b.stats};
        LCHK(hipLaunchKernel((const void*)&k_bn_norm, dim3(blocks_for(s3.count, 256, 1024), nshot), dim3(256), args, 0, st));
    }
    {
        FeatArgs a{u, ncell, vr, g.R / g.vmax, s3, g.bl, g.bf, g.bu, g.nx, g.ny, g.nz, b.stats, b.feat, s3.count * 16};
        void* args[] = {&a};
        LCHK(hipLaunchKernel((const void*)&k_bn_features, dim3(blocks_for(s3.count, 256, 1L << 30), nshot), dim3(256),
                             args, 0, st));
    }
    int rc;
    if ((rc = conv_layer(st, EPI_HIDDEN, b.feat, s3.count * 16, s3, s2, Wbf, bias, b.h1, s2.count * 16, b.y1,
                         s2.count * 16, nshot)))
        return rc;
    if ((rc = conv_layer(st, EPI_HIDDEN, b.h1, s2.count * 16, s2, s1, Wbf + BN_WSZ, bias + 16, b.h2, s1.count * 16,
                         b.y2, s1.count * 16, nshot)))
        return rc;
    if ((rc = conv_layer(st, EPI_OUT, b.h2, s1.count * 16, s1, s0, Wbf + 2 * BN_WSZ, bias + 32, nullptr, 0, b.y3,
                         s0.count * 2, nshot)))
        return rc;
    if (apply) {
        long ys = s0.count * 2;
        {
            void* args[] = {&u, &ncell, (void*)&s0, &b.y3, &ys, &b.stats};
            LCHK(hipLaunchKernel((const void*)&k_bn_guard, dim3(blocks_for(s0.count, 256, 1024), nshot), dim3(256),
                                 args, 0, st));
        }
        {
            void* args[] = {&u, &ncell, (void*)&s0, &b.y3, &ys, &b.stats, &nshot, &guard, &b.tripped};
            LCHK(hipLaunchKernel((const void*)&k_bn_apply, dim3(blocks_for(s0.count, 256, 1024), nshot), dim3(256),
                                 args, 0, st));
        }
    }
    return OSE_OK;
}

int bnet_taper_select(hipStream_t st, float* eff, const float* gx, int nx, const float* gy, int ny, const float* gz,
                      int nz, const int* tripped)
{
    void* args[] = {&eff, &gx, &nx, &gy, &ny, &gz, &nz, &tripped};
    LCHK(hipLaunchKernel((const void*)&k_bn_taper, dim3((nx + ny + nz + 255) / 256), dim3(256), args, 0, st));
    return OSE_OK;
}

int bnet_backward(hipStream_t st, const BnetGeom& g, const uint16_t* Wbf, const uint16_t* WTbf, const float* bias,
                  const float2* u, const float2* t, const float* vr, long ncell, BnetBufs& b, uint16_t* dy_a,
                  uint16_t* dy_b, float* grad, float* loss)
{
    const Shell& s0 = g.sh[0];
    const Shell& s1 = g.sh[1];
    const Shell& s2 = g.sh[2];
    const Shell& s3 = g.sh[3];
    {
        int n = BN_NPAR;
        int clear = 0;
        int* nf = nullptr;
        void* args[] = {&grad, &n, &nf, &clear};
        LCHK(hipLaunchKernel((const void*)&k_bn_zero, dim3(blocks_for(n, 256, 1L << 30)), dim3(256), args, 0, st));
        int one = 1;
        void* args2[] = {&loss, &one, &nf, &clear};
        LCHK(hipLaunchKernel((const void*)&k_bn_zero, dim3(1), dim3(64), args2, 0, st));
    }
    int rc;
    if ((rc = bnet_forward(st, g, Wbf, bias, const_cast<float2*>(u), vr, ncell, 1, b, 0.f, false))) return rc;
    float* dW1 = grad;
    float* db1 = grad + BN_WSZ;
    float* dW2 = grad + BN_WSZ + 16;
    float* db2 = dW2 + BN_WSZ;
    float* dW3 = db2 + 16;
    float* db3 = dW3 + BN_WSZ;
    // loss -> dy3 (on R_0) in dy_a
    {
        void* args[] = {(void*)&u, (void*)&t, (void*)&s0, (void*)&s3, &b.y3, &b.stats, &dy_a, &loss};
        LCHK(hipLaunchKernel((const void*)&k_bn_loss, dim3(blocks_for(s0.count, 256, 1024)), dim3(256), args, 0, st));
    }
    auto wgrad = [&](const uint16_t* x, const Shell& sx, const uint16_t* dy, const Shell& sdy, float* dW) -> int {
        const long chunk = 256;
        WgradArgs a{x, sx, dy, sdy, dW, chunk};
        void* args[] = {&a};
        const long waves = (sdy.count + chunk - 1) / chunk;
        LCHK(hipLaunchKernel((const void*)&k_bn_wgrad, dim3((unsigned)((waves + 3) / 4), BN_TAPS), dim3(256), args, 0, st));
        return OSE_OK;
    };
    auto bgrad = [&](const uint16_t* dy, long n, float* db) -> int {
        void* args[] = {(void*)&dy, &n, &db};
        LCHK(hipLaunchKernel((const void*)&k_bn_bgrad, dim3(blocks_for(n, 256, 512)), dim3(256), args, 0, st));
        return OSE_OK;
    };
    auto dgrad = [&](const uint16_t* dy, const Shell& sdy, const Shell& sx, const uint16_t* WT, const float* ypre,
                     uint16_t* dprev) -> int {
        DgradArgs a{dy, sdy, sx, WT, ypre, dprev};
        void* args[] = {&a};
        const long tiles = (sx.count + 15) / 16;
        LCHK(hipLaunchKernel((const void*)&k_bn_dgrad, dim3(blocks_for(tiles, 4, 2048)), dim3(256), args, 0, st));
        return OSE_OK;
    };
    // layer 3
    if ((rc = wgrad(b.h2, s1, dy_a, s0, dW3)) || (rc = bgrad(dy_a, s0.count, db3))) return rc;
    if ((rc = dgrad(dy_a, s0, s1, WTbf + 2 * BN_WSZ, b.y2, dy_b))) return rc;   // dy2 on R_1
    // layer 2
    if ((rc = wgrad(b.h1, s2, dy_b, s1, dW2)) || (rc = bgrad(dy_b, s1.count, db2))) return rc;
    if ((rc = dgrad(dy_b, s1, s2, WTbf + BN_WSZ, b.y1, dy_a))) return rc;       // dy1 on R_2
    // layer 1
    if ((rc = wgrad(b.feat, s3, dy_a, s2, dW1)) || (rc = bgrad(dy_a, s2.count, db1))) return rc;
    return OSE_OK;
}

int bnet_convert(hipStream_t st, const float* params, uint16_t* Wbf, uint16_t* WTbf, float* bias)
{
    void* args[] = {(void*)&params, &Wbf, &WTbf, &bias};
    LCHK(hipLaunchKernel((const void*)&k_bn_convert, dim3((3 * BN_WSZ + 255) / 256), dim3(256), args, 0, st));
    return OSE_OK;
}

int bnet_adam(hipStream_t st, float* params, const float* grad, float* m, float* v, int step, float lr, float beta1,
              float beta2, float eps)
{
    int n = BN_NPAR;
    float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    void* args[] = {&params, (void*)&grad, &m, &v, &n, &lr, &beta1, &beta2, &c1, &c2, &eps};
    LCHK(hipLaunchKernel((const void*)&k_bn_adam, dim3((n + 255) / 256), dim3(256), args, 0, st));
    return OSE_OK;
}

}  // namespace ose
