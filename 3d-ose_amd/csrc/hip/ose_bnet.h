// ose_bnet.h -- learned thin-pad boundary network (internal interface).
//
// A 3-layer 3D convolution (9 input features padded to 16 channels -> 16 ->
// 16 -> 2, 3^3 taps, GELU) applied after every OSE step to the pad of a thin
// absorbing layer: a learned passive complex gain u' = u G, G = q/max(1,|q|),
// q = 1 + y, y = net(u / s, v, geometry) on the pad cells, s the RMS of u
// over the receptive shell.  Activations live
// only on the shell R_d = {cells within distance d of the pad} (d = 3..0),
// stored compactly as three boxes on the periodic grid (a z band, a y band
// of the remaining z, an x band of the remaining z and y).  Every GEMM -
// forward, input gradient and weight gradient - is a BF16 implicit GEMM on
// MFMA 16x16x16 with FP32 accumulation.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ose {

constexpr int BN_C = 16;                       // channels
constexpr int BN_TAPS = 27;
constexpr int BN_WSZ = BN_TAPS * BN_C * BN_C;  // weights of one layer
constexpr int BN_NPAR = 3 * (BN_WSZ + BN_C);   // W1 b1 W2 b2 W3 b3

// R_d on the periodic padded grid, as three boxes
struct Shell {
    int nx, ny, nz;       // padded extents
    int bx0, by0, bz0;    // first cell of each band (interior end - d, modulo)
    int tx, ty, tz;       // band thicknesses (pad width + 2d, at most the extent)
    long offY, offX, count;
};

Shell make_shell(int nxb, int nyb, int nzb, int bl, int bf, int bu, int nx, int ny, int nz, int d);

__host__ __device__ inline int wrapc(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }
__host__ __device__ inline int bandpos(int c, int b0, int n)
{
    const int r = c - b0;
    return r < 0 ? r + n : r;
}

// compact index of (x, y, z) in the shell, or -1
__host__ __device__ inline long shell_index(const Shell& s, int x, int y, int z)
{
    const int bz = bandpos(z, s.bz0, s.nz);
    if (bz < s.tz) return ((long)bz * s.ny + y) * s.nx + x;
    const int iz = bz - s.tz;
    const int by = bandpos(y, s.by0, s.ny);
    if (by < s.ty) return s.offY + ((long)iz * s.ty + by) * s.nx + x;
    const int iy = by - s.ty;
    const int bx = bandpos(x, s.bx0, s.nx);
    if (bx < s.tx) return s.offX + ((long)iz * (s.ny - s.ty) + iy) * s.tx + bx;
    return -1;
}

// grid coordinates of compact index i
__host__ __device__ inline void shell_coords(const Shell& s, long i, int& x, int& y, int& z)
{
    if (i < s.offY) {
        x = (int)(i % s.nx);
        const long r = i / s.nx;
        y = (int)(r % s.ny);
        z = wrapc(s.bz0 + (int)(r / s.ny), s.nz);
    } else if (i < s.offX) {
        const long j = i - s.offY;
        x = (int)(j % s.nx);
        const long r = j / s.nx;
        y = wrapc(s.by0 + (int)(r % s.ty), s.ny);
        z = wrapc(s.bz0 + s.tz + (int)(r / s.ty), s.nz);
    } else {
        const long j = i - s.offX;
        x = wrapc(s.bx0 + (int)(j % s.tx), s.nx);
        const long r = j / s.tx;
        const int ny2 = s.ny - s.ty;
        y = wrapc(s.by0 + s.ty + (int)(r % ny2), s.ny);
        z = wrapc(s.bz0 + s.tz + (int)(r / ny2), s.nz);
    }
}

struct BnetGeom {
    int nxb, nyb, nzb, bl, bf, bu, nx, ny, nz;
    float vmax;
    float R;              // the model's spectral radius: v = vr * R
    Shell sh[4];          // R_0 .. R_3
};

// Device buffers of one network evaluation over `nshot` shots.
struct BnetBufs {
    uint16_t* feat = nullptr;   // [shot][n3][16] bf16
    uint16_t* h1 = nullptr;     // [shot][n2][16] bf16 (post-GELU)
    uint16_t* h2 = nullptr;     // [shot][n1][16]
    float* y1 = nullptr;        // [shot][n2][16] pre-activation (training only)
    float* y2 = nullptr;        // [shot][n1][16] (training only)
    float* y3 = nullptr;        // [shot][n0][2]  network output
    float* stats = nullptr;     // [shot][4]: sum |u|^2 on R_3, sum |delta|^2, sum |u|^2 on R_0, guard
    int* tripped = nullptr;     // [1]: set by the guard
};

// Launch the forward pass for all shots of a field (u in physical space,
// full padded grid, shot stride ncell) and multiply the pad cells of u by
// G(y) unless the energy guard trips (then u is left as is, *tripped = 1).
// Wbf: bf16 weights [3][27][16][16]; bias: fp32 [3][16].
int bnet_forward(hipStream_t st, const BnetGeom& g, const uint16_t* Wbf, const float* bias,
                 float2* u, const float* vr, long ncell, int nshot, BnetBufs& b, float guard, bool apply);

// eff[0:nx+ny+nz] = 1 while *tripped == 0, else the classical tapers gx|gy|gz
int bnet_taper_select(hipStream_t st, float* eff, const float* gx, int nx, const float* gy, int ny, const float* gz,
                      int nz, const int* tripped);

// Training: one sample, forward + backward into grad (fp32 [BN_NPAR], zeroed
// here), loss (device scalar).  t: target field, same layout as u.
int bnet_backward(hipStream_t st, const BnetGeom& g, const uint16_t* Wbf, const uint16_t* WTbf,
                  const float* bias, const float2* u, const float2* t, const float* vr, long ncell,
                  BnetBufs& b, uint16_t* dy_a, uint16_t* dy_b, float* grad, float* loss);

// fp32 master -> bf16 [3][27][16][16] (W) and transposed WT[l][tap][ci][co]
int bnet_convert(hipStream_t st, const float* params, uint16_t* Wbf, uint16_t* WTbf, float* bias);
int bnet_adam(hipStream_t st, float* params, const float* grad, float* m, float* v, int step, float lr,
              float beta1, float beta2, float eps);

}  // namespace ose
