"""Generate golden fixtures with numpy/scipy, independently of the C oracle.

This script is a second, independent restatement of the reference's
arithmetic (Littleke/3D-OSE) written with numpy.fft (pocketfft) and
scipy.special.jv.  It does not import the C oracle or the product.  Its
outputs pin the FP64 C oracle in tests/test_oracle.py:

  bessel.npz     scipy.special.jv(n, x)                (jnf, ose_3d_kernel.cu:310,388)
  fft.npz        numpy.fft.fftn / fft                  (cufftXtExecDescriptorC2C)
  wavelet.npz    Ricker closed form, Hilbert via rfft/irfft without the 1/nt
                 (generate_wavelet_gpu :3-13; Alloc_wavelet ose_3d_lib.cu:575-583;
                  hilbert_1d ose_3d_kernel.cu:154-172)
  step_tiny.npz  a tiny run of wavefield_extro_ori2 (ose_3d_lib.cu:639-874)
                 in the reference's lane order with separate real transforms,
                 plus the time loop of ose_3d.cu:248-275

Run:  python tests/golden/gen_fixtures.py
"""
import os

import numpy as np
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def bessel():
    ns = np.arange(0, 25)
    xs = np.array([0.1, 0.5, 1.0, 1.904, 2.177, 2.449, 2.816, 3.5, 5.0, 8.0, 12.5])
    J = sp.jv(ns[:, None], xs[None, :])
    np.savez(os.path.join(HERE, "bessel.npz"), n=ns, x=xs, J=J)


def fft():
    rng = np.random.default_rng(11)
    a = rng.standard_normal((6, 10, 14)) + 1j * rng.standard_normal((6, 10, 14))
    out = {"a3": a, "f3": np.fft.fftn(a), "i3": np.fft.ifftn(a) * a.size}
    for n in (19, 53, 140, 152, 336, 864):
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"a{n}"] = v
        out[f"f{n}"] = np.fft.fft(v)
    np.savez(os.path.join(HERE, "fft.npz"), **out)


def ricker(nt, fm, dt, amp=100.0, delay=40):
    i = np.arange(nt)
    t = np.pi * fm * ((i - delay) * dt - 1.0 / fm)
    t = t * t
    return amp * (1.0 - 2.0 * t) * np.exp(-t)


def hilbert_unnorm(s):
    nt = s.size
    H = np.fft.rfft(s)
    k = np.arange(H.size)
    H = np.where(k < nt // 2, -1j * H, 1j * H)
    # irfft keeps only the real part of the DC (and Nyquist) bins, as a
    # complex-to-real transform of Hermitian data does; cuFFT's C2R is
    # unnormalised, numpy's is normalised by 1/nt.
    return np.fft.irfft(H, n=nt) * nt


def wavelet():
    out = {}
    for (nt, fm, dt) in ((300, 40.0, 1e-3), (1000, 25.0, 1e-3), (301, 20.0, 1e-3)):
        s = ricker(nt, fm, dt)
        out[f"s_{nt}"] = s
        out[f"h_{nt}"] = hilbert_unnorm(s)
        out[f"meta_{nt}"] = np.array([nt, fm, dt])
    np.savez(os.path.join(HERE, "wavelet.npz"), **out)


def kabs_grid(nxb, nyb, nzb, dx, dy, dz):
    def fold(n, d):
        i = np.arange(n)
        return 2.0 * np.pi / (d * n) * np.minimum(i, n - i)
    kx, ky, kz = fold(nxb, dx), fold(nyb, dy), fold(nzb, dz)
    return np.sqrt(kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2)


def taper_ref(b1, n, b2):
    a = np.ones(b1 + n + b2)
    i = np.arange(b1)
    a[:b1] = np.exp(-(((b1 - 1 - i) / b1) ** 2))
    j = np.arange(b2)
    a[b1 + n:] = np.exp(-(((j + 1) / b2) ** 2))
    return a


def step_np(P, Q, v, R, kabs, coef, g):
    def L(f):
        F = np.fft.fftn(f) * kabs
        return np.real(np.fft.ifftn(F)) * v / R

    K = len(coef) - 1
    p2x, q2x = P.copy(), Q.copy()
    p2y, q2y = coef[0] * P, coef[0] * Q
    p1x = L(q2x)
    q1x = -L(p2x)
    p2y = p2y + coef[1] * p1x
    q2y = q2y + coef[1] * q1x
    for n in range(1, K):
        q2x = q2x - 2.0 * L(p1x)
        p2x = p2x + 2.0 * L(q1x)
        p2y = (p2y + coef[n + 1] * p2x) * g
        q2y = (q2y + coef[n + 1] * q2x) * g
        p1x, p2x = p2x, p1x
        q1x, q2x = q2x, q1x
    return p2y, q2y


def step_tiny():
    # interior 10 x 17 x 26, pads 3 on every side, then the odd-extent rule
    # (padded 16 x 24 x 32: three different extents catch axis mix-ups)
    nx, ny, nz = 10, 17, 26
    bl = br = bf = bb = bu = bd = 3
    nxb, nyb, nzb = bl + nx + br, bf + ny + bb, bu + nz + bd
    if nyb % 2:
        bb += 1; nyb += 1
    if nzb % 2:
        bd += 1; nzb += 1
    dx = dy = dz = 10.0
    dt = 1e-3
    rng = np.random.default_rng(5)
    vi = np.where(np.arange(nz)[:, None, None] < 9, 2500.0, 3500.0) * np.ones((nz, ny, nx))
    vi = vi * (1.0 + 0.05 * rng.standard_normal((nz, ny, nx)))
    vi = vi.astype(np.float32).astype(np.float64)
    v = np.pad(vi, ((bu, bd), (bf, bb), (bl, br)), mode="edge")
    vmax = v.max()
    R = vmax * np.pi * np.sqrt(dx ** -2 + dy ** -2 + dz ** -2)
    z = R * dt
    M = int(z) + 6
    K = M + 1
    coef = np.array([sp.jv(0, z)] + [2.0 * sp.jv(k, z) for k in range(1, K + 1)])
    g = (taper_ref(bu, nz, bd)[:, None, None] * taper_ref(bf, ny, bb)[None, :, None]
         * taper_ref(bl, nx, br)[None, None, :])
    kabs = kabs_grid(nxb, nyb, nzb, dx, dy, dz)
    nt = 12
    s = ricker(nt, 120.0, dt, amp=100.0, delay=2)
    h = hilbert_unnorm(s)
    sx, sy, sz = 5, 8, 4
    src = (bu + sz) * nyb * nxb + (bf + sy) * nxb + bl + sx
    P = np.zeros((nzb, nyb, nxb))
    Q = np.zeros((nzb, nyb, nxb))
    rz = bu
    tr = np.zeros((ny * nx, nt))
    for it in range(nt):
        P.flat[src] += s[it]
        Q.flat[src] += h[it]
        P, Q = step_np(P, Q, v, R, kabs, coef, g)
        tr[:, it] = Q[rz, bf:bf + ny, bl:bl + nx].reshape(-1)
    np.savez(os.path.join(HERE, "step_tiny.npz"),
             dims=np.array([nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb]),
             d=np.array([dx, dy, dz, dt]), v=v, vi=vi, R=R, coef=coef,
             gx=taper_ref(bl, nx, br), gy=taper_ref(bf, ny, bb), gz=taper_ref(bu, nz, bd),
             s=s, h=h, src=src, rz=rz, traces=tr, P=P, Q=Q)


if __name__ == "__main__":
    bessel()
    fft()
    wavelet()
    step_tiny()
    print("fixtures written to", HERE)
