"""ctypes binding of the FP64 CPU oracle (liboseoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class OrcModel(ctypes.Structure):
    _fields_ = [
        ("nxb", ctypes.c_int), ("nyb", ctypes.c_int), ("nzb", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("dt", ctypes.c_double), ("R", ctypes.c_double), ("K", ctypes.c_int),
        ("coef", _dp), ("v", _dp), ("gx", _dp), ("gy", _dp), ("gz", _dp),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboseoracle.so")
        if not os.path.exists(path):
            raise RuntimeError("liboseoracle.so not built: run make -C oracle")
        L = ctypes.CDLL(path)
        L.orc_fft1d.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.orc_fft3d.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_dft3d_brute.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_ricker.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_hilbert.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_bessel_jn.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_bessel_jn.restype = ctypes.c_double
        L.orc_taper_ref.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_taper_cerjan.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_judge_odd.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_velocity_from_file_order.argtypes = [_fp] + [ctypes.c_int] * 9 + [_dp]
        L.orc_velocity_pad_interior.argtypes = [_fp] + [ctypes.c_int] * 9 + [_dp]
        L.orc_apply_L.argtypes = [ctypes.POINTER(OrcModel), _dp, _dp]
        L.orc_step.argtypes = [ctypes.POINTER(OrcModel), _dp, _dp]
        L.orc_run.argtypes = [ctypes.POINTER(OrcModel), ctypes.c_int, ctypes.c_long, _dp, _dp,
                              ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              _dp, _dp, _dp]
        L.orc_run_profile.argtypes = [ctypes.POINTER(OrcModel), ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _dp]
        L.orc_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def fft3d(a, sign=-1):
    """Unnormalised 3D DFT of a complex [nz][ny][nx] array."""
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nz, ny, nx = a.shape
    assert lib().orc_fft3d(a.view(np.float64).ctypes.data_as(_dp), nx, ny, nz, sign) == 0
    return a


def fft1d(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    assert lib().orc_fft1d(a.view(np.float64).ctypes.data_as(_dp), a.size, sign) == 0
    return a


def dft3d_brute(a, sign=-1):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.empty_like(a)
    nz, ny, nx = a.shape
    lib().orc_dft3d_brute(a.view(np.float64).ctypes.data_as(_dp),
                          out.view(np.float64).ctypes.data_as(_dp), nx, ny, nz, sign)
    return out


def ricker(nt, fm, dt, amp=100.0, delay=40):
    w = np.empty(nt)
    lib().orc_ricker(_d(w), nt, fm, dt, amp, delay)
    return w


def hilbert(s):
    s = _c(s)
    h = np.empty_like(s)
    lib().orc_hilbert(_d(s), _d(h), s.size)
    return h


def bessel_jn(n, x):
    return lib().orc_bessel_jn(n, x)


def taper_ref(before, nn, after):
    a = np.empty(before + nn + after)
    lib().orc_taper_ref(_d(a), before, nn, after)
    return a


def taper_cerjan(before, nn, after, alpha):
    a = np.empty(before + nn + after)
    lib().orc_taper_cerjan(_d(a), before, nn, after, alpha)
    return a


def judge_odd(after, nb):
    a = ctypes.c_int(after)
    n = ctypes.c_int(nb)
    lib().orc_judge_odd(ctypes.byref(a), ctypes.byref(n))
    return a.value, n.value


def velocity_from_file_order(f_xyz, nx, ny, nz, bl, bf, bu, nxb, nyb, nzb):
    f = np.ascontiguousarray(f_xyz, dtype=np.float32)
    v = np.empty((nzb, nyb, nxb))
    lib().orc_velocity_from_file_order(f.ctypes.data_as(_fp), nx, ny, nz, bl, bf, bu,
                                       nxb, nyb, nzb, _d(v))
    return v


def velocity_pad_interior(v_int, bl, bf, bu, nxb, nyb, nzb):
    vi = np.ascontiguousarray(v_int, dtype=np.float32)
    nz, ny, nx = vi.shape
    v = np.empty((nzb, nyb, nxb))
    lib().orc_velocity_pad_interior(vi.ctypes.data_as(_fp), nx, ny, nz, bl, bf, bu,
                                    nxb, nyb, nzb, _d(v))
    return v


class Model:
    """FP64 model: padded velocity, spacing, R, K, coefficients, taper."""

    def __init__(self, v, dx, dy, dz, dt, R, coef, gx, gy, gz):
        self.v = _c(v)
        self.coef = _c(coef)
        self.gx, self.gy, self.gz = _c(gx), _c(gy), _c(gz)
        nzb, nyb, nxb = self.v.shape
        assert self.gx.size == nxb and self.gy.size == nyb and self.gz.size == nzb
        self.s = OrcModel(nxb, nyb, nzb, dx, dy, dz, dt, R, len(self.coef) - 1,
                          _d(self.coef), _d(self.v), _d(self.gx), _d(self.gy), _d(self.gz))

    @property
    def shape(self):
        return self.v.shape

    def apply_L(self, u):
        u = np.ascontiguousarray(u, dtype=np.complex128)
        out = np.empty_like(u)
        assert lib().orc_apply_L(ctypes.byref(self.s), u.view(np.float64).ctypes.data_as(_dp),
                                 out.view(np.float64).ctypes.data_as(_dp)) == 0
        return out

    def step(self, P, Q):
        P = _c(P).copy()
        Q = _c(Q).copy()
        assert lib().orc_step(ctypes.byref(self.s), _d(P), _d(Q)) == 0
        return P, Q

    def run(self, nt, src, s, h, rz, bl, bf, nx, ny, P=None, Q=None, record=True):
        shp = self.v.shape
        P = np.zeros(shp) if P is None else _c(P).copy()
        Q = np.zeros(shp) if Q is None else _c(Q).copy()
        s, h = _c(s), _c(h)
        tr = np.zeros((ny * nx, nt)) if record else None
        rc = lib().orc_run(ctypes.byref(self.s), nt, int(src), _d(s), _d(h), rz, bl, bf, nx, ny,
                           _d(tr) if record else None, _d(P), _d(Q))
        assert rc == 0
        return tr, P, Q


def run_profile(model, nt, px, py, pz, s, h, rz, bl, bf, nx, ny, P=None, Q=None, record=True):
    """Model.run with the source P += s px py pz (full-length profiles)."""
    m = model
    shp = m.v.shape
    P = np.zeros(shp) if P is None else _c(P).copy()
    Q = np.zeros(shp) if Q is None else _c(Q).copy()
    px, py, pz, s, h = _c(px), _c(py), _c(pz), _c(s), _c(h)
    assert px.size == shp[2] and py.size == shp[1] and pz.size == shp[0]
    tr = np.zeros((ny * nx, nt)) if record else None
    rc = lib().orc_run_profile(ctypes.byref(m.s), nt, _d(px), _d(py), _d(pz), _d(s), _d(h), rz, bl, bf, nx, ny,
                               _d(tr) if record else None, _d(P), _d(Q))
    assert rc == 0
    return tr, P, Q


def num_threads():
    return lib().orc_num_threads()
