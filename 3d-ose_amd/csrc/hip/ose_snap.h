// ose_snap.h -- wavefield snapshot encoding on the device.
//
// A snapshot is one shot's state u = P + iQ over a box of the padded grid
// (the whole grid or the interior), one or both lanes (bit 0: Re u = P,
// bit 1: Im u = Q), in one of four encodings:
//   F32  4 B/value, exact
//   BF16 2 B/value, round to nearest even of the float32 bits
//   Q16  per x-line of the box a float scale s = max |v| and int16 values
//        q = rint(v * 32767 / s), decoded as q * (s / 32767)
//   Q8   the same with int8 and 127
// Layout of one encoded snapshot: for the Q formats the scales first,
// float [lane][z][y], then the values [lane][z][y][x] (x fastest); F32 and
// BF16 hold only the values.  Lanes appear in order P, Q.
//
// The encoder reads 8 B/cell (the complex state) and writes lanes x 1..4
// B/cell: it is bound by HBM bandwidth.  One workgroup of 256 threads
// encodes one x-line (the line maximum is a workgroup reduction).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ose {

enum SnapFormat { SNAP_F32 = 0, SNAP_BF16 = 1, SNAP_Q16 = 2, SNAP_Q8 = 3 };

__host__ __device__ inline int snap_value_bytes(int format)
{
    return format == SNAP_F32 ? 4 : format == SNAP_Q8 ? 1 : 2;
}

__host__ __device__ inline bool snap_has_scales(int format) { return format == SNAP_Q16 || format == SNAP_Q8; }

// bytes of one encoded snapshot of nl lanes over an nx x ny x nz box
__host__ __device__ inline size_t snap_bytes(int format, int nl, int nx, int ny, int nz)
{
    const size_t cells = (size_t)nx * ny * nz;
    const size_t sc = snap_has_scales(format) ? sizeof(float) * (size_t)nl * ny * nz : 0;
    return sc + (size_t)nl * cells * snap_value_bytes(format);
}

// float32 -> bfloat16, round to nearest even; NaN stays a quiet NaN
__host__ __device__ inline uint16_t snap_bf16(float v)
{
    uint32_t b;
    __builtin_memcpy(&b, &v, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x0040u);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

struct SnapArgs {
    const float2* u;      // the shot's state, padded [nzb][nyb][nxb]
    int nxb, nyb;
    int x0, y0, z0;       // box origin in the padded grid
    int nx, ny, nz;       // box extents
    int lanes;            // bit 0: Re u, bit 1: Im u
    int format;
    void* out;
};

constexpr int kSnapThreads = 256;

__device__ inline float snap_lane(float2 w, int lane) { return lane == 0 ? w.x : w.y; }

template <int FORMAT>
__global__ void __launch_bounds__(kSnapThreads) k_snap(SnapArgs a)
{
    const int y = blockIdx.x, z = blockIdx.y;
    const int t = threadIdx.x;
    const float2* line = a.u + ((long)(a.z0 + z) * a.nyb + (a.y0 + y)) * a.nxb + a.x0;
    const size_t cells = (size_t)a.nx * a.ny * a.nz;
    const int nl = (a.lanes & 1) + ((a.lanes >> 1) & 1);
    const size_t lofs = ((size_t)z * a.ny + y) * a.nx;     // line offset inside one lane
    float* scales = (float*)a.out;
    char* vals = (char*)a.out + (snap_has_scales(FORMAT) ? sizeof(float) * (size_t)nl * a.ny * a.nz : 0);
    __shared__ float red[2][kSnapThreads / 64];
    float inv[2] = {0.f, 0.f};
    if (snap_has_scales(FORMAT)) {
        float m0 = 0.f, m1 = 0.f;
        for (int x = t; x < a.nx; x += kSnapThreads) {
            const float2 w = line[x];
            m0 = fmaxf(m0, fabsf(w.x));
            m1 = fmaxf(m1, fabsf(w.y));
        }
        for (int o = 32; o > 0; o >>= 1) {
            m0 = fmaxf(m0, __shfl_xor(m0, o));
            m1 = fmaxf(m1, __shfl_xor(m1, o));
        }
        if ((t & 63) == 0) {
            red[0][t >> 6] = m0;
            red[1][t >> 6] = m1;
        }
        __syncthreads();
        m0 = red[0][0];
        m1 = red[1][0];
#pragma unroll
        for (int k = 1; k < kSnapThreads / 64; ++k) {
            m0 = fmaxf(m0, red[0][k]);
            m1 = fmaxf(m1, red[1][k]);
        }
        const float qmax = FORMAT == SNAP_Q8 ? 127.f : 32767.f;
        int l = 0;
        for (int lane = 0; lane < 2; ++lane) {
            if (!((a.lanes >> lane) & 1)) continue;
            const float s = lane == 0 ? m0 : m1;
            inv[lane] = s > 0.f ? qmax / s : 0.f;
            if (t == 0) scales[(size_t)l * a.ny * a.nz + (size_t)z * a.ny + y] = s;
            ++l;
        }
    }
    for (int x = t; x < a.nx; x += kSnapThreads) {
        const float2 w = line[x];
        int l = 0;
        for (int lane = 0; lane < 2; ++lane) {
            if (!((a.lanes >> lane) & 1)) continue;
            const float v = snap_lane(w, lane);
            const size_t i = (size_t)l * cells + lofs + x;
            if (FORMAT == SNAP_F32) {
                ((float*)vals)[i] = v;
            } else if (FORMAT == SNAP_BF16) {
                ((uint16_t*)vals)[i] = snap_bf16(v);
            } else {
                const float qmax = FORMAT == SNAP_Q8 ? 127.f : 32767.f;
                const float q = fminf(fmaxf(rintf(v * inv[lane]), -qmax), qmax);
                if (FORMAT == SNAP_Q8) ((int8_t*)vals)[i] = (int8_t)q;
                else ((int16_t*)vals)[i] = (int16_t)q;
            }
            ++l;
        }
    }
}

}  // namespace ose
