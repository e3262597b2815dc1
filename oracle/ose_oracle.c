/*
 * ose_oracle.c -- FP64 CPU restatement of the 3D-OSE time stepper.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see the header
 * for which reference function each routine restates).  Nothing in the
 * product links this file.
 *
 * Parity pins (each checked in tests/test_oracle.py):
 *   FFT       brute-force DFT on tiny grids; numpy.fft fixtures
 *   Bessel    scipy.special.jv fixtures; J0 + 2*sum J_2k = 1
 *   Ricker    closed form
 *   Hilbert   DHT of a sampled cosine is the sine (times nt)
 *   taper     closed form; accumulator power law g^(M-k+2)
 *   L, step   single Fourier mode: one step = exp(-i dt v|k|) to within the
 *             truncation tail; |u| conserved for constant v without taper;
 *             an independent numpy restatement on tiny grids (fixtures)
 */
#include "ose_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort, any length.                      */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;
    int nf;
    int f[64];
    cplx* w;   /* w[j] = exp(-2 pi i j / n) */
} fplan;

static int plan_init(fplan* p, int n)
{
    if (n < 1) return -1;
    p->n = n;
    p->nf = 0;
    int m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    for (int d = 3; d * d <= m; d += 2)
        while (m % d == 0) { p->f[p->nf++] = d; m /= d; }
    if (m > 1) p->f[p->nf++] = m;
    p->w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    if (!p->w) return -1;
    for (int j = 0; j < n; ++j) {
        double a = -2.0 * M_PI * (double)j / (double)n;
        p->w[j] = cos(a) + I * sin(a);
    }
    return 0;
}

static void plan_free(fplan* p) { free(p->w); p->w = NULL; }

static inline cplx wpow(const fplan* p, long e, int sign)
{
    cplx w = p->w[e % p->n];
    return sign < 0 ? w : conj(w);
}

/* x: data (n), work: n + n scratch.  Result left in x. */
static void fft_line(const fplan* p, cplx* x, cplx* work, int sign)
{
    const int N = p->n;
    if (N == 1) return;
    cplx* src = x;
    cplx* dst = work;
    cplx* v = work + N;   /* radix scratch, length <= N */
    int L = 1;
    for (int s = 0; s < p->nf; ++s) {
        const int P = p->f[s];
        const int Np = N / P;
        const int m = N / (L * P);
        for (int b = 0; b < Np; ++b) {
            const int k = b % L, j = b / L;
            for (int q = 0; q < P; ++q) {
                cplx t = src[b + q * Np];
                if (q && k) t *= wpow(p, (long)q * k * m, sign);
                v[q] = t;
            }
            cplx* o = dst + (long)j * P * L + k;
            if (P == 2) {
                o[0] = v[0] + v[1];
                o[L] = v[0] - v[1];
            } else if (P == 4) {
                cplx a0 = v[0] + v[2], a1 = v[0] - v[2];
                cplx b0 = v[1] + v[3], b1 = v[1] - v[3];
                cplx rb1 = (sign < 0) ? (-I * b1) : (I * b1);
                o[0] = a0 + b0;
                o[L] = a1 + rb1;
                o[2 * L] = a0 - b0;
                o[3 * L] = a1 - rb1;
            } else {
                for (int r = 0; r < P; ++r) {
                    cplx acc = 0;
                    for (int q = 0; q < P; ++q)
                        acc += v[q] * wpow(p, (long)q * r % P * Np, sign);
                    o[(long)r * L] = acc;
                }
            }
        }
        cplx* t = src; src = dst; dst = t;
        L *= P;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)N);
}

int orc_fft1d(double* a, int n, int sign)
{
    fplan p;
    if (plan_init(&p, n)) return -1;
    cplx* work = (cplx*)malloc(sizeof(cplx) * 2 * (size_t)n);
    if (!work) { plan_free(&p); return -1; }
    fft_line(&p, (cplx*)a, work, sign);
    free(work);
    plan_free(&p);
    return 0;
}

int orc_fft3d(double* a_, int nx, int ny, int nz, int sign)
{
    cplx* a = (cplx*)a_;
    fplan px, py, pz;
    if (plan_init(&px, nx)) return -1;
    if (plan_init(&py, ny)) { plan_free(&px); return -1; }
    if (plan_init(&pz, nz)) { plan_free(&px); plan_free(&py); return -1; }
    const long sxy = (long)nx * ny;
    int maxn = nx > ny ? nx : ny;
    if (nz > maxn) maxn = nz;
    const int BX = 8; /* columns gathered together for the strided axes */
#pragma omp parallel
    {
        cplx* buf = (cplx*)malloc(sizeof(cplx) * (size_t)maxn * (BX + 3));
        cplx* work = buf + (size_t)maxn * BX;
        /* x lines */
#pragma omp for schedule(static)
        for (long l = 0; l < (long)ny * nz; ++l)
            fft_line(&px, a + l * nx, work, sign);
        /* y lines: for each z, blocks of BX columns */
#pragma omp for schedule(static)
        for (long t = 0; t < (long)nz * ((nx + BX - 1) / BX); ++t) {
            const long z = t / ((nx + BX - 1) / BX);
            const int x0 = (int)(t % ((nx + BX - 1) / BX)) * BX;
            const int bw = (x0 + BX <= nx) ? BX : nx - x0;
            cplx* base = a + z * sxy + x0;
            for (int y = 0; y < ny; ++y)
                for (int c = 0; c < bw; ++c) buf[(size_t)c * ny + y] = base[(long)y * nx + c];
            for (int c = 0; c < bw; ++c) fft_line(&py, buf + (size_t)c * ny, work, sign);
            for (int y = 0; y < ny; ++y)
                for (int c = 0; c < bw; ++c) base[(long)y * nx + c] = buf[(size_t)c * ny + y];
        }
        /* z lines: for each y, blocks of BX columns */
#pragma omp for schedule(static)
        for (long t = 0; t < (long)ny * ((nx + BX - 1) / BX); ++t) {
            const long y = t / ((nx + BX - 1) / BX);
            const int x0 = (int)(t % ((nx + BX - 1) / BX)) * BX;
            const int bw = (x0 + BX <= nx) ? BX : nx - x0;
            cplx* base = a + y * nx + x0;
            for (int z = 0; z < nz; ++z)
                for (int c = 0; c < bw; ++c) buf[(size_t)c * nz + z] = base[(long)z * sxy + c];
            for (int c = 0; c < bw; ++c) fft_line(&pz, buf + (size_t)c * nz, work, sign);
            for (int z = 0; z < nz; ++z)
                for (int c = 0; c < bw; ++c) base[(long)z * sxy + c] = buf[(size_t)c * nz + z];
        }
        free(buf);
    }
    plan_free(&px); plan_free(&py); plan_free(&pz);
    return 0;
}

void orc_dft3d_brute(const double* in_, double* out_, int nx, int ny, int nz, int sign)
{
    const cplx* in = (const cplx*)in_;
    cplx* out = (cplx*)out_;
    const long n = (long)nx * ny * nz;
    for (long o = 0; o < n; ++o) {
        const int kx = (int)(o % nx), ky = (int)((o / nx) % ny), kz = (int)(o / ((long)nx * ny));
        cplx acc = 0;
        for (long i = 0; i < n; ++i) {
            const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
            double ph = (double)((long)kx * x % nx) / nx + (double)((long)ky * y % ny) / ny
                      + (double)((long)kz * z % nz) / nz;
            double a = sign * 2.0 * M_PI * ph;
            acc += in[i] * (cos(a) + I * sin(a));
        }
        out[o] = acc;
    }
}

/* ------------------------------------------------------------------ */
/* Source wavelet                                                       */
/* ------------------------------------------------------------------ */
void orc_ricker(double* w, int nt, double fm, double dt, double amp, int delay)
{
    /* generate_wavelet_gpu, ose_3d_kernel.cu:3-13 */
    for (int i = 0; i < nt; ++i) {
        double t = M_PI * fm * ((double)(i - delay) * dt - 1.0 / fm);
        t = t * t;
        w[i] = amp * (1.0 - 2.0 * t) * exp(-t);
    }
}

void orc_hilbert(const double* s, double* h, int nt)
{
    /* R2C (ose_3d_lib.cu:578) -> hilbert_1d (ose_3d_kernel.cu:154-172:
     * bins k < nt/2 times -i, others times +i) -> C2R without scaling
     * (ose_3d_lib.cu:581).  A C2R of Hermitian data takes only the real
     * part of the DC and (even nt) Nyquist bins.                        */
    const int nh = nt / 2 + 1;
    cplx* H = (cplx*)malloc(sizeof(cplx) * (size_t)nh);
    for (int k = 0; k < nh; ++k) {
        cplx acc = 0;
        for (int n = 0; n < nt; ++n) {
            double a = -2.0 * M_PI * (double)((long)k * n % nt) / nt;
            acc += s[n] * (cos(a) + I * sin(a));
        }
        H[k] = (k < nt / 2) ? (-I * acc) : (I * acc);
    }
    const int kmax = (nt % 2 == 0) ? nt / 2 - 1 : (nt - 1) / 2;
    for (int n = 0; n < nt; ++n) {
        double acc = creal(H[0]);
        if (nt % 2 == 0) acc += creal(H[nt / 2]) * ((n & 1) ? -1.0 : 1.0);
        for (int k = 1; k <= kmax; ++k) {
            double a = 2.0 * M_PI * (double)((long)k * n % nt) / nt;
            acc += 2.0 * creal(H[k] * (cos(a) + I * sin(a)));
        }
        h[n] = acc;
    }
    free(H);
}

/* ------------------------------------------------------------------ */
/* Bessel, taper, grid                                                  */
/* ------------------------------------------------------------------ */
double orc_bessel_jn(int n, double x) { return jn(n, x); }

void orc_taper_ref(double* a, int before, int nn, int after)
{
    /* generate_att, ose_3d_kernel.cu:242-261 */
    for (int i = 0; i < before; ++i) {
        double t = (double)(before - 1 - i) / (double)before;
        a[i] = exp(-t * t);
    }
    for (int i = before; i < before + nn; ++i) a[i] = 1.0;
    for (int i = 0; i < after; ++i) {
        double t = (double)(i + 1) / (double)after;
        a[before + nn + i] = exp(-t * t);
    }
}

void orc_taper_cerjan(double* a, int before, int nn, int after, double alpha)
{
    /* Cerjan profile exp(-(alpha*d)^2), d = 1..b cells from the interior */
    for (int i = 0; i < before; ++i) {
        double t = alpha * (double)(before - i);
        a[i] = exp(-t * t);
    }
    for (int i = before; i < before + nn; ++i) a[i] = 1.0;
    for (int i = 0; i < after; ++i) {
        double t = alpha * (double)(i + 1);
        a[before + nn + i] = exp(-t * t);
    }
}

int orc_judge_odd(int* after, int* nb)
{
    /* ose_3d_lib.cu:36-44 */
    if (*nb % 2 != 0) { *after += 1; *nb += 1; }
    return *nb;
}

static void pad_edges(double* v, int nx, int ny, int nz, int bl, int bf, int bu,
                      int nxb, int nyb, int nzb)
{
    /* add_pml_layers_v_h, ose_3d_lib.cu:912-1032: sides and corners inside
     * the interior z range, then whole planes above and below.          */
    const long sxy = (long)nxb * nyb;
    for (int iz = bu; iz < bu + nz; ++iz)
        for (int iy = 0; iy < nyb; ++iy) {
            const int sy = iy < bf ? bf : (iy >= bf + ny ? bf + ny - 1 : iy);
            for (int ix = 0; ix < nxb; ++ix) {
                const int sx = ix < bl ? bl : (ix >= bl + nx ? bl + nx - 1 : ix);
                if (sx == ix && sy == iy) continue;
                v[iz * sxy + (long)iy * nxb + ix] = v[iz * sxy + (long)sy * nxb + sx];
            }
        }
    for (int iz = 0; iz < bu; ++iz)
        memcpy(v + iz * sxy, v + (long)bu * sxy, sizeof(double) * (size_t)sxy);
    for (int iz = bu + nz; iz < nzb; ++iz)
        memcpy(v + iz * sxy, v + (long)(bu + nz - 1) * sxy, sizeof(double) * (size_t)sxy);
}

void orc_velocity_from_file_order(const float* f, int nx, int ny, int nz,
                                  int bl, int bf, int bu, int nxb, int nyb, int nzb,
                                  double* v)
{
    /* input_file_xyz_boundary, ose_3d_lib.cu:898-906: file is [x][y][z] */
    memset(v, 0, sizeof(double) * (size_t)nxb * nyb * nzb);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz)
                v[(long)(iz + bu) * nxb * nyb + (long)(iy + bf) * nxb + ix + bl] =
                    f[((long)ix * ny + iy) * nz + iz];
    pad_edges(v, nx, ny, nz, bl, bf, bu, nxb, nyb, nzb);
}

void orc_velocity_pad_interior(const float* vi, int nx, int ny, int nz,
                               int bl, int bf, int bu, int nxb, int nyb, int nzb,
                               double* v)
{
    memset(v, 0, sizeof(double) * (size_t)nxb * nyb * nzb);
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix)
                v[(long)(iz + bu) * nxb * nyb + (long)(iy + bf) * nxb + ix + bl] =
                    vi[((long)iz * ny + iy) * nx + ix];
    pad_edges(v, nx, ny, nz, bl, bf, bu, nxb, nyb, nzb);
}

/* ------------------------------------------------------------------ */
/* Operator and step                                                    */
/* ------------------------------------------------------------------ */
static double kfold(int i, int n, double dk)
{
    /* scale_kxkz_3D, ose_3d_kernel.cu:109-120: folded wavenumber */
    return dk * (double)(i <= n / 2 ? i : n - i);
}

int orc_apply_L(const orc_model* m, const double* in_c, double* out_c)
{
    const int nx = m->nxb, ny = m->nyb, nz = m->nzb;
    const long n = (long)nx * ny * nz;
    cplx* t = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    if (!t) return -1;
    memcpy(t, in_c, sizeof(cplx) * (size_t)n);
    if (orc_fft3d((double*)t, nx, ny, nz, -1)) { free(t); return -1; }
    const double dkx = 2.0 * M_PI / (m->dx * nx);
    const double dky = 2.0 * M_PI / (m->dy * ny);
    const double dkz = 2.0 * M_PI / (m->dz * nz);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
        const double kx = kfold(x, nx, dkx), ky = kfold(y, ny, dky), kz = kfold(z, nz, dkz);
        t[i] *= sqrt(kx * kx + ky * ky + kz * kz);
    }
    if (orc_fft3d((double*)t, nx, ny, nz, +1)) { free(t); return -1; }
    cplx* o = (cplx*)out_c;
    const double inv = 1.0 / ((double)n * m->R);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) o[i] = t[i] * (m->v[i] * inv);
    free(t);
    return 0;
}

int orc_step(const orc_model* m, double* P, double* Q)
{
    /* wavefield_extro_ori2, ose_3d_lib.cu:639-874, in the reference's lanes:
     * p2x,q2x / p1x,q1x hold Re,Im of (-i)^k Z_k for the two latest k;
     * p2y,q2y (here P,Q) accumulate.  L acts on real fields; two real
     * fields share one complex transform (L is real, so this is exact). */
    const long n = (long)m->nxb * m->nyb * m->nzb;
    const int M = m->K - 1;
    if (m->K < 2) return -1;
    double* p1 = (double*)malloc(sizeof(double) * (size_t)n);
    double* p2 = (double*)malloc(sizeof(double) * (size_t)n);
    double* q1 = (double*)malloc(sizeof(double) * (size_t)n);
    double* q2 = (double*)malloc(sizeof(double) * (size_t)n);
    cplx* z = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    int rc = 0;
    if (!p1 || !p2 || !q1 || !q2 || !z) { rc = -1; goto done; }
    const double c0 = m->coef[0], c1 = m->coef[1];
    /* step1to4_3d (ose_3d_kernel.cu:300-312), n = 0 */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        p2[i] = P[i]; P[i] = c0 * p2[i];
        q2[i] = Q[i]; Q[i] = c0 * q2[i];
        z[i] = q2[i] + I * p2[i];
    }
    /* p1.x = +L(q2.x), q1.x = -L(p2.x)  (ose_3d_lib.cu:662-728) */
    if (orc_apply_L(m, (double*)z, (double*)z)) { rc = -1; goto done; }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        p1[i] = creal(z[i]);
        q1[i] = -cimag(z[i]);
        /* add_bessel_3d n=1 (ose_3d_lib.cu:741-743) */
        P[i] += c1 * p1[i];
        Q[i] += c1 * q1[i];
    }
    for (int it = 1; it <= M; ++it) {
        const double cn = m->coef[it + 1];
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) z[i] = p1[i] + I * q1[i];
        if (orc_apply_L(m, (double*)z, (double*)z)) { rc = -1; goto done; }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            /* scal_vr_3d -2 into q2.x, +2 into p2.x (ose_3d_lib.cu:787,830) */
            q2[i] += -2.0 * creal(z[i]);
            p2[i] += 2.0 * cimag(z[i]);
            /* add_bessel_3d n+1 (:845-847) */
            P[i] += cn * p2[i];
            Q[i] += cn * q2[i];
            /* apply_ABC_Xt (:857-858) */
            const int x = (int)(i % m->nxb), y = (int)((i / m->nxb) % m->nyb);
            const int zz = (int)(i / ((long)m->nxb * m->nyb));
            const double g = m->gx[x] * m->gy[y] * m->gz[zz];
            P[i] *= g;
            Q[i] *= g;
            /* replace_2array_3d (:867-868) */
            double t = p1[i]; p1[i] = p2[i]; p2[i] = t;
            t = q1[i]; q1[i] = q2[i]; q2[i] = t;
        }
    }
done:
    free(p1); free(p2); free(q1); free(q2); free(z);
    return rc;
}

int orc_run_profile(const orc_model* m, int nt, const double* px, const double* py, const double* pz,
                    const double* s, const double* h, int rz, int bl, int bf, int nx, int ny,
                    double* traces, double* P, double* Q)
{
    const long sxy = (long)m->nxb * m->nyb;
    for (int it = 0; it < nt; ++it) {
        for (int z = 0; z < m->nzb; ++z) {
            if (pz[z] == 0.0) continue;
            for (int y = 0; y < m->nyb; ++y) {
                if (py[y] == 0.0) continue;
                const double a = py[y] * pz[z];
                for (int x = 0; x < m->nxb; ++x) {
                    if (px[x] == 0.0) continue;
                    const long i = z * sxy + (long)y * m->nxb + x;
                    P[i] += s[it] * (a * px[x]);
                    Q[i] += h[it] * (a * px[x]);
                }
            }
        }
        if (orc_step(m, P, Q)) return -1;
        if (traces)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix)
                    traces[((long)iy * nx + ix) * nt + it] =
                        Q[rz * sxy + (long)(bf + iy) * m->nxb + bl + ix];
    }
    return 0;
}

int orc_run(const orc_model* m, int nt, long src, const double* s, const double* h,
            int rz, int bl, int bf, int nx, int ny, double* traces,
            double* P, double* Q)
{
    /* ose_3d.cu:248-275: inject (Q6), step, record Q at depth plane (Q7) */
    const long sxy = (long)m->nxb * m->nyb;
    for (int it = 0; it < nt; ++it) {
        P[src] += s[it];
        Q[src] += h[it];
        if (orc_step(m, P, Q)) return -1;
        if (traces)
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix)
                    traces[((long)iy * nx + ix) * nt + it] =
                        Q[rz * sxy + (long)(bf + iy) * m->nxb + bl + ix];
    }
    return 0;
}
