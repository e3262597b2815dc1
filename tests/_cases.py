"""Shared set-up of parity cases (checker side).

The parameters that the reference derives in main() (ose_3d.cu:75-199) are
computed here with numpy/scipy so that GPU and oracle get identical inputs.
"""
import numpy as np
import scipy.special as sp


def rel(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) else np.float64)
    b = np.asarray(b, dtype=np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def coefs(z, K):
    return np.array([sp.jv(0, z)] + [2.0 * sp.jv(k, z) for k in range(1, K + 1)])


def taper_ref(b1, n, b2):
    a = np.ones(b1 + n + b2)
    i = np.arange(b1)
    a[:b1] = np.exp(-(((b1 - 1 - i) / b1) ** 2))
    j = np.arange(b2)
    a[b1 + n:] = np.exp(-(((j + 1) / b2) ** 2))
    return a


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
    return np.fft.irfft(H, n=nt) * nt


class Case:
    """A single-shot model: interior (nx, ny, nz), pads, velocity, K."""

    def __init__(self, interior, pads, v_int, d=(10.0, 10.0, 10.0), dt=1e-3, K=None,
                 taper="ref", rec_z=0):
        nx, ny, nz = interior
        bl, br, bf, bb, bu, bd = pads
        self.nx, self.ny, self.nz = nx, ny, nz
        self.bl, self.bf, self.bu = bl, bf, bu
        self.nxb, self.nyb, self.nzb = bl + nx + br, bf + ny + bb, bu + nz + bd
        self.br, self.bb, self.bd = br, bb, bd
        self.d, self.dt = d, dt
        vi = np.asarray(v_int, dtype=np.float32).astype(np.float64)
        self.v = np.pad(vi, ((bu, bd), (bf, bb), (bl, br)), mode="edge")
        self.R = self.v.max() * np.pi * np.sqrt(sum(x ** -2 for x in d))
        z = self.R * dt
        self.K = (int(z) + 7) if K is None else K
        self.coef = coefs(z, self.K)
        if taper == "none":
            self.gx, self.gy, self.gz = np.ones(self.nxb), np.ones(self.nyb), np.ones(self.nzb)
        else:
            self.gx = taper_ref(bl, nx, br)
            self.gy = taper_ref(bf, ny, bb)
            self.gz = taper_ref(bu, nz, bd)
        self.rec_z = rec_z

    @property
    def shape(self):
        return (self.nzb, self.nyb, self.nxb)

    def add_to(self, ctx):
        return ctx.add_model(self.v, interior=(self.nx, self.ny, self.nz),
                             origin=(self.bl, self.bf, self.bu), d=self.d, dt=self.dt,
                             R=self.R, coef=self.coef, taper=(self.gx, self.gy, self.gz),
                             rec_z=self.rec_z)

    def oracle(self):
        import pyoracle
        return pyoracle.Model(self.v, self.d[0], self.d[1], self.d[2], self.dt, self.R,
                              self.coef, self.gx, self.gy, self.gz)

    def src_index(self, sx, sy, sz):
        return ((self.bu + sz) * self.nyb + (self.bf + sy)) * self.nxb + self.bl + sx


def layered(nx, ny, nz, seed=0, v0=2000.0, v1=3500.0, jitter=0.05):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None, None]
    v = v0 + (v1 - v0) * (z / max(nz - 1, 1)) * np.ones((nz, ny, nx))
    v *= 1.0 + jitter * rng.standard_normal((nz, ny, nx))
    return v.astype(np.float32)
