// ose_host_model.cpp -- velocity models: reference file order, edge padding
// and the seeded generators of the benchmark configurations.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "ose_host.h"

namespace {

inline long idx(int x, int y, int z, int nx, int ny) { return ((long)z * ny + y) * nx + x; }

// separable box filter of width w (odd), edge mode, along each axis
void box_smooth(std::vector<float>& v, int nx, int ny, int nz, int w)
{
    const int h = w / 2;
    std::vector<double> line;
    auto pass = [&](int n, long stride, long count, auto base_of) {
        line.resize(n);
        for (long l = 0; l < count; ++l) {
            const long b = base_of(l);
            for (int i = 0; i < n; ++i) line[i] = v[b + i * stride];
            for (int i = 0; i < n; ++i) {
                double acc = 0.0;
                for (int k = -h; k <= h; ++k) acc += line[std::min(std::max(i + k, 0), n - 1)];
                v[b + i * stride] = (float)(acc / w);
            }
        }
    };
    pass(nx, 1, (long)ny * nz, [&](long l) { return l * nx; });
    pass(ny, nx, (long)nx * nz, [&](long l) { return (l / nx) * (long)nx * ny + (l % nx); });
    pass(nz, (long)nx * ny, (long)nx * ny, [&](long l) { return l; });
}

// monotone jittered interface depths for nl layers over nz cells
std::vector<double> interfaces(uint64_t* rng, int nl, int nz)
{
    std::vector<double> d(nl + 1);
    d[0] = 0.0;
    d[nl] = nz;
    const double step = (double)nz / nl;
    for (int i = 1; i < nl; ++i) d[i] = i * step + (oseh_rng_uniform(rng) - 0.5) * 0.5 * step;
    return d;
}

}  // namespace

extern "C" {

int oseh_read_velocity_xyz(const char* path, int nx, int ny, int nz, float* v)
{
    if (!path || !v || nx <= 0 || ny <= 0 || nz <= 0) return OSE_EINVAL;
    FILE* f = std::fopen(path, "rb");
    if (!f) return OSE_EINVAL;
    std::vector<float> buf((size_t)nx * ny * nz);
    const size_t got = std::fread(buf.data(), sizeof(float), buf.size(), f);
    std::fclose(f);
    if (got != buf.size()) return OSE_EINVAL;   // short file
    // file [x][y][z], z fastest (ose_3d_lib.cu:898-906) -> [z][y][x]
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nz; ++z) {
                const float s = buf[((size_t)x * ny + y) * nz + z];
                if (!std::isfinite(s) || s <= 0.f) return OSE_ENUMERIC;
                v[idx(x, y, z, nx, ny)] = s;
            }
    return OSE_OK;
}

void oseh_pad_velocity(const float* vi, int nx, int ny, int nz, int bl, int br, int bf, int bb,
                       int bu, int bd, float* vp)
{
    // add_pml_layers_v_h (ose_3d_lib.cu:912-1032) copies the nearest
    // interior value into every pad cell: np.pad(mode="edge")
    const int nxb = bl + nx + br, nyb = bf + ny + bb, nzb = bu + nz + bd;
    for (int z = 0; z < nzb; ++z) {
        const int sz = std::min(std::max(z - bu, 0), nz - 1);
        for (int y = 0; y < nyb; ++y) {
            const int sy = std::min(std::max(y - bf, 0), ny - 1);
            float* o = vp + idx(0, y, z, nxb, nyb);
            const float* r = vi + idx(0, sy, sz, nx, ny);
            for (int x = 0; x < nxb; ++x) o[x] = r[std::min(std::max(x - bl, 0), nx - 1)];
        }
    }
}

int oseh_model(int kind, int nx, int ny, int nz, uint64_t seed, double v0, double v1, int iface,
               float* out)
{
    if (!out || nx <= 0 || ny <= 0 || nz <= 0) return OSE_EINVAL;
    uint64_t rng[4];
    oseh_rng_seed(rng, seed);
    const long n = (long)nx * ny * nz;
    std::vector<float> v((size_t)n);
    if (kind == 0) {
        std::fill(v.begin(), v.end(), (float)v0);
    } else if (kind == 1) {
        for (int z = 0; z < nz; ++z)
            for (long i = 0; i < (long)nx * ny; ++i) v[(long)z * nx * ny + i] = (float)(z < iface ? v0 : v1);
    } else if (kind == 2 || kind == 4) {
        int nl = 8, nanom = 6;
        double vlo = 1500.0, vhi = 4500.0, sigma = 15.0, amp = 0.15;
        if (kind == 4) {   // training models: random layer count and range
            nl = 3 + (int)(oseh_rng_uniform(rng) * 7.0);
            nanom = (int)(oseh_rng_uniform(rng) * 5.0);
            vlo = 1500.0 + 500.0 * oseh_rng_uniform(rng);
            vhi = 3000.0 + 1500.0 * oseh_rng_uniform(rng);
            sigma = std::max(2.0, 0.08 * std::min(std::min(nx, ny), nz));
        }
        const std::vector<double> d = interfaces(rng, nl, nz);
        for (int z = 0; z < nz; ++z) {
            int l = 0;
            while (l + 1 < nl && z + 0.5 >= d[l + 1]) ++l;
            const float val = (float)(vlo + (vhi - vlo) * (nl > 1 ? (double)l / (nl - 1) : 0.0));
            for (long i = 0; i < (long)nx * ny; ++i) v[(long)z * nx * ny + i] = val;
        }
        for (int a = 0; a < nanom; ++a) {
            const double cx = nx * (0.15 + 0.7 * oseh_rng_uniform(rng));
            const double cy = ny * (0.15 + 0.7 * oseh_rng_uniform(rng));
            const double cz = nz * (0.15 + 0.7 * oseh_rng_uniform(rng));
            const double s = (oseh_rng_uniform(rng) < 0.5 ? -amp : amp);
            const double inv = 1.0 / (2.0 * sigma * sigma);
            const int r = (int)(4 * sigma);
            for (int z = std::max(0, (int)cz - r); z < std::min(nz, (int)cz + r + 1); ++z)
                for (int y = std::max(0, (int)cy - r); y < std::min(ny, (int)cy + r + 1); ++y)
                    for (int x = std::max(0, (int)cx - r); x < std::min(nx, (int)cx + r + 1); ++x) {
                        const double r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
                        v[idx(x, y, z, nx, ny)] *= (float)(1.0 + s * std::exp(-r2 * inv));
                    }
        }
        box_smooth(v, nx, ny, nz, 5);
    } else if (kind == 3) {
        const int nl = 12;
        const std::vector<double> d = interfaces(rng, nl, nz);
        for (int z = 0; z < nz; ++z) {
            int l = 0;
            while (l + 1 < nl && z + 0.5 >= d[l + 1]) ++l;
            const float val = (float)(1500.0 + 2500.0 * (double)l / (nl - 1));
            for (long i = 0; i < (long)nx * ny; ++i) v[(long)z * nx * ny + i] = val;
        }
        // dipping ellipsoid: semi-axes 150/120/80 cells at 768^3, scaled
        const double sc = std::min(std::min(nx, ny), nz) / 768.0;
        const double a = 150.0 * sc, b = 120.0 * sc, c = 80.0 * sc;
        const double cx = nx * 0.5, cy = ny * 0.5, cz = nz * (0.5 + 0.1 * oseh_rng_uniform(rng));
        const double dip = (15.0 + 15.0 * oseh_rng_uniform(rng)) * M_PI / 180.0;
        const double cd = std::cos(dip), sd = std::sin(dip);
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const double dx = x - cx, dy = y - cy, dz = z - cz;
                    const double u = cd * dx + sd * dz, w = -sd * dx + cd * dz;   // rotate about y
                    if ((u * u) / (a * a) + (dy * dy) / (b * b) + (w * w) / (c * c) <= 1.0)
                        v[idx(x, y, z, nx, ny)] = 4500.0f;
                }
    } else {
        return OSE_EINVAL;
    }
    std::copy(v.begin(), v.end(), out);
    return OSE_OK;
}

}  // extern "C"
