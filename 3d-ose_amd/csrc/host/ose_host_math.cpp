// ose_host_math.cpp -- FP64 host arithmetic of the OSE driver: grid rule,
// Ricker and Hilbert source, Bessel coefficients, order rules, tapers and
// the seeded random generator.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "ose_host.h"

extern "C" {

int oseh_judge_odd(int* after, int* nb)
{
    // ose_3d_lib.cu:36-44
    if (*nb % 2 != 0) {
        *after += 1;
        *nb += 1;
    }
    return *nb;
}

void oseh_ricker(double* w, int nt, double fm, double dt, double amp, int delay)
{
    // ose_3d_kernel.cu:3-13: tmp = pi fm ((i - delay) dt - 1/fm); w = amp (1 - 2 tmp^2) e^{-tmp^2}
    for (int i = 0; i < nt; ++i) {
        double t = M_PI * fm * (((double)i - (double)delay) * dt - 1.0 / fm);
        t *= t;
        w[i] = amp * (1.0 - 2.0 * t) * std::exp(-t);
    }
}

void oseh_hilbert(const double* s, double* h, int nt, int normalize)
{
    // Half spectrum S[k] = sum_n s[n] e^{-2 pi i k n / nt}, k <= nt/2; bins
    // k < nt/2 are rotated by -i and the rest by +i (hilbert_1d); the
    // complex-to-real inverse keeps the real part of the DC and Nyquist
    // bins and doubles the others.  Direct sums with an exact root table.
    if (nt <= 0) return;
    std::vector<double> c(nt), sn(nt);
    for (int j = 0; j < nt; ++j) {
        const double a = 2.0 * M_PI * (double)j / (double)nt;
        c[j] = std::cos(a);
        sn[j] = std::sin(a);
    }
    const int nh = nt / 2 + 1;
    std::vector<double> hr(nh), hi(nh);
    for (int k = 0; k < nh; ++k) {
        double re = 0.0, im = 0.0;
        for (int n = 0; n < nt; ++n) {
            const int j = (int)(((long)k * n) % nt);
            re += s[n] * c[j];
            im -= s[n] * sn[j];
        }
        if (k < nt / 2) {   // times -i
            hr[k] = im;
            hi[k] = -re;
        } else {            // times +i
            hr[k] = -im;
            hi[k] = re;
        }
    }
    const int kmax = (nt % 2 == 0) ? nt / 2 - 1 : (nt - 1) / 2;
    const double scale = normalize ? 1.0 / (double)nt : 1.0;
    for (int n = 0; n < nt; ++n) {
        double acc = hr[0];
        if (nt % 2 == 0) acc += hr[nt / 2] * ((n & 1) ? -1.0 : 1.0);
        for (int k = 1; k <= kmax; ++k) {
            const int j = (int)(((long)k * n) % nt);
            acc += 2.0 * (hr[k] * c[j] - hi[k] * sn[j]);
        }
        h[n] = acc * scale;
    }
}

int oseh_bessel_all(double x, int nmax, double* J)
{
    if (nmax < 0 || !J) return OSE_EINVAL;
    if (!(x >= 0.0) || !std::isfinite(x)) return OSE_ENUMERIC;
    if (x < 1e-300) {
        J[0] = 1.0;
        for (int n = 1; n <= nmax; ++n) J[n] = 0.0;
        return OSE_OK;
    }
    // Miller: start well above both nmax and x, recur downwards
    //   J_{k-1} = (2k/x) J_k - J_{k+1},
    // then normalise with J0 + 2 sum_{k>=1} J_{2k} = 1.
    int start = (int)(std::max((double)nmax, x) + 20.0 + 6.0 * std::sqrt(std::max((double)nmax, x) + 1.0));
    start += start & 1;
    std::vector<double> t((size_t)start + 2, 0.0);
    t[start + 1] = 0.0;
    t[start] = 1e-300;
    for (int k = start; k >= 1; --k) {
        t[k - 1] = (2.0 * k / x) * t[k] - t[k + 1];
        if (std::fabs(t[k - 1]) > 1e250) {   // rescale to avoid overflow
            for (int j = k - 1; j <= start + 1; ++j) t[j] *= 1e-250;
        }
    }
    double norm = t[0];
    for (int k = 2; k <= start; k += 2) norm += 2.0 * t[k];
    for (int n = 0; n <= nmax; ++n) J[n] = (n <= start) ? t[n] / norm : 0.0;
    return OSE_OK;
}

int oseh_coefficients(double z, int K, double* coef)
{
    if (K < 0 || !coef) return OSE_EINVAL;
    std::vector<double> J((size_t)K + 1);
    int rc = oseh_bessel_all(z, K, J.data());
    if (rc) return rc;
    coef[0] = J[0];
    for (int k = 1; k <= K; ++k) coef[k] = 2.0 * J[k];
    return OSE_OK;
}

double oseh_spectral_radius(double vmax, double dx, double dy, double dz)
{
    return vmax * M_PI * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
}

// R dt beyond this has no usable series order (K <= 64 holds up to ~50)
constexpr double kZMax = 1e4;

int oseh_order_reference(double z)
{
    // ose_3d.cu:197: M = (int)dtR + 6; the loop runs to order M + 1
    if (!(z >= 0.0) || z > kZMax) return -1;
    return (int)z + 6 + 1;
}

int oseh_order_tail(double z, double tol)
{
    if (!(z >= 0.0) || z > kZMax || !(tol > 0.0)) return -1;
    const int nmax = (int)z + 80;
    std::vector<double> J((size_t)nmax + 1);
    if (oseh_bessel_all(z, nmax, J.data())) return -1;
    // tail(K) = 2 sum_{k>K} |J_k|
    double tail = 0.0;
    std::vector<double> tails((size_t)nmax + 1, 0.0);
    for (int k = nmax; k >= 0; --k) {
        tails[k] = tail;   // sum over k' > k
        tail += 2.0 * std::fabs(J[k]);
    }
    for (int K = 2; K <= nmax; ++K)
        if (tails[K] < tol) return K;
    return nmax;
}

void oseh_taper(float* a, int before, int nn, int after, int mode, double alpha)
{
    for (int i = 0; i < before + nn + after; ++i) a[i] = 1.0f;
    if (mode == OSEH_TAPER_NONE) return;
    for (int i = 0; i < before; ++i) {
        double t;
        if (mode == OSEH_TAPER_REF)
            t = (double)(before - 1 - i) / (double)before;   // ose_3d_kernel.cu:247-251
        else
            t = alpha * (double)(before - i);
        a[i] = (float)std::exp(-t * t);
    }
    for (int i = 0; i < after; ++i) {
        double t;
        if (mode == OSEH_TAPER_REF)
            t = (double)(i + 1) / (double)after;             // ose_3d_kernel.cu:256-260
        else
            t = alpha * (double)(i + 1);
        a[before + nn + i] = (float)std::exp(-t * t);
    }
}

// ---- random numbers: splitmix64 seeding, xoshiro256** ------------------
static uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void oseh_rng_seed(uint64_t* s, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
}

uint64_t oseh_rng_next(uint64_t* s)
{
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

double oseh_rng_uniform(uint64_t* s) { return (double)(oseh_rng_next(s) >> 11) * 0x1.0p-53; }

double oseh_rng_normal(uint64_t* s)
{
    // Box-Muller (one of the pair)
    double u1 = oseh_rng_uniform(s);
    const double u2 = oseh_rng_uniform(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

}  // extern "C"
