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
// the same function as x * sigmoid(2 k (x + 0.044715 x^3)) through the
// hardware exp2 and reciprocal (~1e-6 relative; the result is rounded to bf16)
__device__ __forceinline__ float gelu_fast(float x)
{
    const float k2 = -2.f * 0.7978845608028654f * 1.4426950408889634f;   // -2 k log2(e)
    const float e = __builtin_amdgcn_exp2f(k2 * (x + 0.044715f * x * x * x));
    return x * __builtin_amdgcn_rcpf(1.f + e);
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
// BF16 implicit GEMM on MFMA 16x16x32 (two taps per instruction), FP32
// accumulation.  The output shell's three bands are each a dense box [A][B][C]
// (C fastest) in the compact layout, mapped to the grid by an offset:
// (z, y, x) = (zo + A, yo + B, xo + C) modulo the extents.  A workgroup owns a
// TA x TB x TC tile of one band: it stages the tile's (TA+2)(TB+2)(TC+2)
// neighbour rows in LDS (each row found through the input shell's index,
// zero where the neighbour is not in the shell) and reads every MFMA's A
// operand from there: 81 B of global reads per output cell instead of
// 27 x 32 B of gathers.
// ---------------------------------------------------------------------
enum { EPI_HIDDEN = 0, EPI_OUT = 1 };

// one band of the output shell
struct ConvBand {
    long off_o;                 // compact index of the band's first cell
    int nA, nB, nC;             // box extents
    int zo, yo, xo;             // grid coordinates of box cell (0, 0, 0)
    int tA, tB, tC;             // tiles along A, B, C
    long tile0;                 // first tile of this band
};

constexpr int CT_A = 4, CT_B = 4, CT_C = 16;
constexpr int CH_A = CT_A + 2, CH_B = CT_B + 2, CH_C = CT_C + 2;
constexpr int CH_ROWS = CH_A * CH_B * CH_C;
// halo row stride in bf16 values: 48 bytes, so the 16-byte reads of eight
// consecutive rows (one LDS cycle of a 128-bit read) hit disjoint banks
constexpr int CH_LD = 24;

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
    ConvBand band[3];
};

// Two taps per MFMA (16x16x32, K = 2 taps x 16 channels): lane group g
// (lanes 16g..16g+15) supplies channels 8(g%2)..8(g%2)+7 of tap 2p + g/2 for
// the A operand (one 16-byte read) and the matching weights for B.  Tap 27
// does not exist: its weights are zero.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
constexpr int BN_PAIRS = (BN_TAPS + 1) / 2;

// v mod n (the division only for the rare v outside [0, n))
__device__ __forceinline__ int modn(int v, int n)
{
    if (v < 0 || v >= n) {
        v %= n;
        if (v < 0) v += n;
    }
    return v;
}

template <int EPI>
__global__ __launch_bounds__(256) void k_bn_conv(ConvArgs a)
{
    __shared__ __align__(16) uint16_t halo[CH_ROWS * CH_LD];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int r16 = lane & 15;
    const int half = g >> 1;                 // which tap of the pair
    const int ch0 = 8 * (g & 1);             // first channel of this lane
    const int shot = blockIdx.y;
    const uint16_t* in = a.in + (long)shot * a.in_shot;

    // the workgroup's tile: band, then A, B, C tile coordinates
    const long t = blockIdx.x;
    const int bi = t >= a.band[2].tile0 ? 2 : (t >= a.band[1].tile0 ? 1 : 0);
    const ConvBand& B = a.band[bi];
    const long tl = t - B.tile0;
    const int tc = (int)(tl % B.tC);
    const long q = tl / B.tC;
    const int tb = (int)(q % B.tB);
    const int ta = (int)(q / B.tB);
    const int a0 = ta * CT_A, b0 = tb * CT_B, c0 = tc * CT_C;
    const int a1 = min(a0 + CT_A, B.nA), b1 = min(b0 + CT_B, B.nB), c1 = min(c0 + CT_C, B.nC);

    // stage the neighbour rows: halo (ia, ib, ic) is the input shell's row at
    // grid cell (zo + a0 + ia - 1, yo + b0 + ib - 1, xo + c0 + ic - 1)
    for (int row = threadIdx.x; row < CH_ROWS; row += blockDim.x) {
        const int ic = row % CH_C, ib = (row / CH_C) % CH_B, ia = row / (CH_C * CH_B);
        const int z = modn(B.zo + a0 + ia - 1, a.sin.nz);
        const int y = modn(B.yo + b0 + ib - 1, a.sin.ny);
        const int x = modn(B.xo + c0 + ic - 1, a.sin.nx);
        const long nb = shell_index(a.sin, x, y, z);
        s16x8 v0 = {0, 0, 0, 0, 0, 0, 0, 0}, v1 = v0;
        if (nb >= 0) {
            v0 = *(const s16x8*)(in + nb * 16);
            v1 = *(const s16x8*)(in + nb * 16 + 8);
        }
        *(s16x8*)(halo + row * CH_LD) = v0;
        *(s16x8*)(halo + row * CH_LD + 8) = v1;
    }
    s16x8 wb[BN_PAIRS];
#pragma unroll
    for (int p = 0; p < BN_PAIRS; ++p) {
        const int tap = 2 * p + half;
        wb[p] = tap < BN_TAPS ? *(const s16x8*)(a.W + (tap * 16 + r16) * 16 + ch0) : s16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
    const float bias = a.bias[r16];
    __syncthreads();

    // wave w owns the rows (A, B) = (a0 + w, b0 .. b0 + 3), one MFMA tile of
    // 16 C-consecutive cells each
    f32x4 acc[CT_B];
#pragma unroll
    for (int k = 0; k < CT_B; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int p = 0; p < BN_PAIRS; ++p) {
        const int tap = min(2 * p + half, BN_TAPS - 1);      // tap 27: zero weights
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
#pragma unroll
        for (int k = 0; k < CT_B; ++k) {
            const int row = ((wave + dz) * CH_B + (k + dy)) * CH_C + (r16 + dx);
            const s16x8 av = *(const s16x8*)(halo + row * CH_LD + ch0);
            acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                              __builtin_bit_cast(bf16x8, wb[p]), acc[k], 0, 0, 0);
        }
    }
    // bias, activation, store: lane (r16, g) holds cells c0 + 4g + j, channel r16
    const int A = a0 + wave;
    if (EPI == EPI_HIDDEN) {
        // through LDS to cell-major rows, so that each lane stores 16 bytes
        __shared__ __align__(16) uint16_t outl[4][CT_B][16][16];
#pragma unroll
        for (int k = 0; k < CT_B; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float yv = acc[k][j] + bias;
                outl[wave][k][4 * g + j][r16] = f2bf(gelu_fast(yv));
                if (a.yout && A < a1 && b0 + k < b1 && 4 * g + j < c1 - c0)
                    a.yout[(long)shot * a.y_shot + (B.off_o + ((long)A * B.nB + b0 + k) * B.nC + c0 + 4 * g + j) * 16 +
                           r16] = yv;
            }
        __syncthreads();
        if (A >= a1) return;
        // lane: tile kk = 2 i + lane / 32, cell (lane % 32) / 2, half lane % 2
#pragma unroll
        for (int i = 0; i < CT_B / 2; ++i) {
            const int kk = 2 * i + (lane >> 5), cell = (lane & 31) >> 1, hh = lane & 1;
            const int Bc = b0 + kk;
            const s16x8 v = *(const s16x8*)&outl[wave][kk][cell][8 * hh];
            if (Bc < b1 && cell < c1 - c0) {
                const long oc = B.off_o + ((long)A * B.nB + Bc) * B.nC + c0 + cell;
                *(s16x8*)(a.hout + (long)shot * a.h_shot + oc * 16 + 8 * hh) = v;
            }
        }
    } else {
        if (A >= a1) return;
#pragma unroll
        for (int k = 0; k < CT_B; ++k) {
            const int Bc = b0 + k;
            if (Bc >= b1) continue;
            const long idx0 = B.off_o + ((long)A * B.nB + Bc) * B.nC + c0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = 4 * g + j;
                if (c < c1 - c0 && r16 < 2) a.yout[(long)shot * a.y_shot + (idx0 + c) * 2 + r16] = acc[k][j] + bias;
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
    ConvArgs a{};
    a.in = in;
    a.in_shot = in_shot;
    a.sin = sin;
    a.sout = sout;
    a.W = W;
    a.bias = bias;
    a.hout = hout;
    a.h_shot = h_shot;
    a.yout = yout;
    a.y_shot = y_shot;
    const int nx = sout.nx, ny = sout.ny, nz = sout.nz;
    // z band [bz][y][x], y band [iz][by][x], x band [iz][iy][bx] (make_shell)
    const int dims[3][3] = {{sout.tz, ny, nx}, {nz - sout.tz, sout.ty, nx}, {nz - sout.tz, ny - sout.ty, sout.tx}};
    const int org[3][3] = {{sout.bz0, 0, 0}, {sout.bz0 + sout.tz, sout.by0, 0},
                           {sout.bz0 + sout.tz, sout.by0 + sout.ty, sout.bx0}};
    const long off_o[3] = {0, sout.offY, sout.offX};
    long tiles = 0;
    for (int b = 0; b < 3; ++b) {
        ConvBand& c = a.band[b];
        c.off_o = off_o[b];
        c.nA = dims[b][0], c.nB = dims[b][1], c.nC = dims[b][2];
        c.zo = org[b][0], c.yo = org[b][1], c.xo = org[b][2];
        c.tA = (c.nA + CT_A - 1) / CT_A, c.tB = (c.nB + CT_B - 1) / CT_B, c.tC = (c.nC + CT_C - 1) / CT_C;
        if (c.nA <= 0 || c.nB <= 0 || c.nC <= 0) c.tA = c.tB = c.tC = 0;
        c.tile0 = tiles;
        tiles += (long)c.tA * c.tB * c.tC;
    }
    if (tiles == 0) return OSE_OK;
    if (tiles > 0x7fffffffL) return OSE_EINVAL;
    void* args[] = {&a};
    dim3 grid((unsigned)tiles, nshot);
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
