"""Python view of the C++ host library (include/ose_host.h)."""
import ctypes
import json

import numpy as np

from . import _lib
from ._lib import OseError

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


def _L():
    return _lib.host()


def ricker(nt, fm, dt, amp=100.0, delay=40):
    w = np.empty(nt)
    _L().oseh_ricker(w.ctypes.data_as(_dp), nt, fm, dt, amp, delay)
    return w


def hilbert(s, normalize=False):
    s = np.ascontiguousarray(s, dtype=np.float64)
    h = np.empty_like(s)
    _L().oseh_hilbert(s.ctypes.data_as(_dp), h.ctypes.data_as(_dp), s.size, int(normalize))
    return h


def bessel_all(x, nmax):
    J = np.empty(nmax + 1)
    rc = _L().oseh_bessel_all(x, nmax, J.ctypes.data_as(_dp))
    if rc:
        raise OseError(rc, "oseh_bessel_all")
    return J


def coefficients(z, K):
    c = np.empty(K + 1)
    rc = _L().oseh_coefficients(z, K, c.ctypes.data_as(_dp))
    if rc:
        raise OseError(rc, "oseh_coefficients")
    return c


def spectral_radius(vmax, dx, dy, dz):
    return _L().oseh_spectral_radius(vmax, dx, dy, dz)


def order_reference(z):
    return _L().oseh_order_reference(z)


def order_tail(z, tol=1e-7):
    return _L().oseh_order_tail(z, tol)


TAPER = {"ref": 0, "cerjan": 1, "none": 2}


def taper(before, nn, after, mode="ref", alpha=0.015):
    a = np.empty(before + nn + after, dtype=np.float32)
    _L().oseh_taper(a.ctypes.data_as(_fp), before, nn, after, TAPER[mode], alpha)
    return a


def judge_odd(after, nb):
    a, n = ctypes.c_int(after), ctypes.c_int(nb)
    _L().oseh_judge_odd(ctypes.byref(a), ctypes.byref(n))
    return a.value, n.value


def read_velocity_xyz(path, nx, ny, nz):
    v = np.empty((nz, ny, nx), dtype=np.float32)
    rc = _L().oseh_read_velocity_xyz(path.encode(), nx, ny, nz, v.ctypes.data_as(_fp))
    if rc:
        raise OseError(rc, f"cannot read {path}")
    return v


def pad_velocity(vi, pads):
    vi = np.ascontiguousarray(vi, dtype=np.float32)
    nz, ny, nx = vi.shape
    bl, br, bf, bb, bu, bd = pads
    out = np.empty((bu + nz + bd, bf + ny + bb, bl + nx + br), dtype=np.float32)
    _L().oseh_pad_velocity(vi.ctypes.data_as(_fp), nx, ny, nz, bl, br, bf, bb, bu, bd,
                           out.ctypes.data_as(_fp))
    return out


MODEL = {"const": 0, "twolayer": 1, "layered8": 2, "salt12": 3, "train": 4}


def model(kind, nx, ny, nz, seed=1, v0=2000.0, v1=3500.0, iface=None):
    v = np.empty((nz, ny, nx), dtype=np.float32)
    rc = _L().oseh_model(MODEL[kind] if isinstance(kind, str) else kind, nx, ny, nz, seed, v0, v1,
                         nz // 2 if iface is None else iface, v.ctypes.data_as(_fp))
    if rc:
        raise OseError(rc, "oseh_model")
    return v


def _write_traces(fn, path, traces, dt, rec_x, rec_y, rec_z, src, shot):
    tr = np.ascontiguousarray(traces, dtype=np.float32)
    nrec, nt = tr.shape
    coords = [None if c is None else np.ascontiguousarray(c, dtype=np.float64) for c in (rec_x, rec_y, rec_z)]
    ptrs = [None if c is None else c.ctypes.data_as(_dp) for c in coords]
    src = tuple(src) + (0.0,) * (3 - len(src))
    rc = fn(str(path).encode(), tr.ctypes.data_as(_fp), nrec, nt, dt, *ptrs, src[0], src[1], src[2], shot)
    if rc:
        raise OseError(rc, f"cannot write {path}")


def _read_traces(fn, path, cap):
    buf = np.empty(cap, dtype=np.float32)
    nrec, nt, dt = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    rc = fn(str(path).encode(), buf.ctypes.data_as(_fp), cap, ctypes.byref(nrec), ctypes.byref(nt),
            ctypes.byref(dt))
    if rc:
        raise OseError(rc, f"cannot read {path}")
    return buf[: nrec.value * nt.value].reshape(nrec.value, nt.value).copy(), dt.value


def write_su(path, traces, dt, rec_x=None, rec_y=None, src=(0.0, 0.0, 0.0), shot=0, rec_z=None):
    """rec_x/rec_y/rec_z: receiver coordinates in metres (z = depth);
    src: (x, y[, depth])."""
    _write_traces(_L().oseh_write_su, path, traces, dt, rec_x, rec_y, rec_z, src, shot)


def read_su(path, cap=1 << 26):
    return _read_traces(_L().oseh_read_su, path, cap)


def write_segy(path, traces, dt, rec_x=None, rec_y=None, src=(0.0, 0.0, 0.0), shot=0, rec_z=None):
    """rec_x/rec_y/rec_z: receiver coordinates in metres (z = depth);
    src: (x, y[, depth])."""
    _write_traces(_L().oseh_write_segy, path, traces, dt, rec_x, rec_y, rec_z, src, shot)


def read_segy(path, cap=1 << 26):
    return _read_traces(_L().oseh_read_segy, path, cap)


def write_snapshot(path, f_zyx):
    f = np.ascontiguousarray(f_zyx, dtype=np.float32)
    nzb, nyb, nxb = f.shape
    rc = _L().oseh_write_snapshot(path.encode(), f.ctypes.data_as(_fp), nxb, nyb, nzb)
    if rc:
        raise OseError(rc, f"cannot write {path}")


def read_params(path):
    """Boundary-network parameters from a .npy float32 array or raw float32 file."""
    n = ctypes.c_size_t()
    rc = _L().oseh_read_params(str(path).encode(), None, 0, ctypes.byref(n))
    if rc:
        raise OseError(rc, f"cannot read parameters from {path}")
    out = np.empty(n.value, dtype=np.float32)
    rc = _L().oseh_read_params(str(path).encode(), out.ctypes.data_as(_fp), out.size, ctypes.byref(n))
    if rc:
        raise OseError(rc, f"cannot read parameters from {path}")
    return out


def gaussian_profile(sigma, center, n, cap=64):
    """(start, values): normalised Gaussian source profile along one axis."""
    out = np.empty(cap, dtype=np.float32)
    start = ctypes.c_int()
    w = _L().oseh_gaussian_profile(float(sigma), int(center), int(n), ctypes.byref(start), out.ctypes.data_as(_fp),
                                   cap)
    if w < 1:
        raise OseError(1, f"gaussian profile sigma={sigma} does not fit {cap} cells")
    return start.value, out[:w].copy()


SNAP_META = ("format", "lane", "step", "shot", "x0", "y0", "z0", "nx", "ny", "nz")


def read_osnap(path):
    """Decode a .osnap file: (float32 [nz][ny][nx], header dict)."""
    meta = (ctypes.c_int * 10)()
    rc = _L().oseh_read_osnap(str(path).encode(), None, 0, meta)
    if rc:
        raise OseError(rc, f"cannot read {path}")
    m = dict(zip(SNAP_META, list(meta)))
    out = np.empty((m["nz"], m["ny"], m["nx"]), dtype=np.float32)
    rc = _L().oseh_read_osnap(str(path).encode(), out.ctypes.data_as(_fp), out.size, meta)
    if rc:
        raise OseError(rc, f"cannot read {path}")
    return out, m


def snapshot_lane(fmt, nl, lane, nx, ny, nz, data):
    """One lane of a delivered snapshot (bytes) as its encoded bytes."""
    n = _L().oseh_snapshot_lane_bytes(fmt, nx, ny, nz)
    out = ctypes.create_string_buffer(n)
    rc = _L().oseh_snapshot_lane(fmt, nl, lane, nx, ny, nz, data, out)
    if rc:
        raise OseError(rc, "oseh_snapshot_lane")
    return out.raw


def decode_snapshot(fmt, nx, ny, nz, lane_data):
    out = np.empty((nz, ny, nx), dtype=np.float32)
    rc = _L().oseh_decode_snapshot(fmt, nx, ny, nz, lane_data, out.ctypes.data_as(_fp))
    if rc:
        raise OseError(rc, "oseh_decode_snapshot")
    return out


class Simulation:
    """A preset or key=value configuration driven through the HIP backend."""

    def __init__(self, params):
        self._L = _L()
        self._hip = _lib.hip()
        h = ctypes.c_void_p()
        rc = self._L.oseh_sim_create(ctypes.byref(h), params.encode())
        if rc:
            raise OseError(rc, f"bad parameters: {params}")
        self._h = h
        self._setup = False

    def _check(self, rc):
        if rc:
            raise OseError(rc, self._L.oseh_sim_error(self._h).decode())

    def build(self):
        """Host-side construction only (no device)."""
        self._check(self._L.oseh_sim_build(self._h))
        return self

    def setup(self):
        self._check(self._L.oseh_sim_setup(self._h))
        self._setup = True
        return self

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.oseh_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        ms = ctypes.c_double()
        self._check(self._L.oseh_sim_run(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def info(self):
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self._L.oseh_sim_info(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def write(self, outdir):
        self._check(self._L.oseh_sim_write(self._h, outdir.encode()))

    def stream_snapshots(self, d, every):
        """Snapshots every `every` steps into d while the run goes on (keys
        snap_format, snap_region, snap_lanes); 0 stops."""
        self._check(self._L.oseh_sim_stream_snapshots(self._h, str(d).encode(), int(every)))

    def checkpoint(self, d):
        self._check(self._L.oseh_sim_checkpoint(self._h, str(d).encode()))

    def restore(self, d):
        self._check(self._L.oseh_sim_restore(self._h, str(d).encode()))

    @property
    def ctx(self):
        return ctypes.c_void_p(self._L.oseh_sim_ctx(self._h))

    def set_options(self, use_graph=True, profile=False):
        rc = self._hip.ose_set_options(self.ctx, int(use_graph), int(profile))
        self._check(rc)

    def timed(self, nsteps):
        ms = ctypes.c_double()
        rc = self._hip.ose_timed_run(self.ctx, int(nsteps), ctypes.byref(ms))
        if rc:
            raise OseError(rc, self._hip.ose_last_error(self.ctx).decode())
        return ms.value

    def reset_stats(self):
        self._hip.ose_reset_stats(self.ctx)

    def pass_stats(self):
        names = ["fwd_y", "z_symbol", "inv_y", "x_epilogue", "x_first", "x_last", "inject", "x_recur", "x_acc_first"]
        out = {}
        for i, n in enumerate(names):
            ms, cnt, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
            self._hip.ose_pass_stats(self.ctx, i, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by))
            out[n] = (ms.value, cnt.value, by.value)
        return out

    def receivers(self, shot):
        """Interior (x, y, z) of the shot's receivers, one row each, in trace order."""
        n = ctypes.c_int(0)
        self._check(self._L.oseh_sim_receivers(self._h, shot, None, ctypes.byref(n)))
        out = np.empty((n.value, 3), dtype=np.int32)
        self._check(self._L.oseh_sim_receivers(self._h, shot, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                               ctypes.byref(n)))
        return out

    def traces(self, shot, it0, nsteps):
        out = np.empty((len(self.receivers(shot)), nsteps), dtype=np.float32)
        rc = self._hip.ose_get_traces(self.ctx, shot, it0, nsteps, out.ctypes.data_as(_fp))
        if rc:
            raise OseError(rc, self._hip.ose_last_error(self.ctx).decode())
        return out

    def wavefield(self, shot):
        d = self.model_dims(self.shot(shot)["model"])
        re = np.empty((d[11], d[10], d[9]), dtype=np.float32)
        im = np.empty_like(re)
        rc = self._hip.ose_get_wavefield(self.ctx, shot, re.ctypes.data_as(_fp), im.ctypes.data_as(_fp))
        if rc:
            raise OseError(rc, self._hip.ose_last_error(self.ctx).decode())
        return re, im

    def set_wavefield(self, shot, re, im):
        """Overwrite a shot's state (initial conditions) after setup()."""
        re = np.ascontiguousarray(re, dtype=np.float32)
        im = np.ascontiguousarray(im, dtype=np.float32)
        rc = self._hip.ose_set_wavefield(self.ctx, shot, re.ctypes.data_as(_fp), im.ctypes.data_as(_fp))
        if rc:
            raise OseError(rc, self._hip.ose_last_error(self.ctx).decode())

    def boundary_status(self, model):
        a, t = ctypes.c_int(), ctypes.c_int()
        rc = self._hip.ose_boundary_status(self.ctx, model, ctypes.byref(a), ctypes.byref(t))
        if rc:
            raise OseError(rc, self._hip.ose_last_error(self.ctx).decode())
        return dict(active=bool(a.value), tripped=bool(t.value))

    def trip_boundary_net(self, model):
        """Force a model's learned pad into its tripped state (plain taper)."""
        rc = self._hip.ose_trip_boundary_net(self.ctx, model)
        if rc:
            raise OseError(rc, self._hip.ose_last_error(self.ctx).decode())

    # model/shot data (for the oracle)
    def nmodels(self):
        return self._L.oseh_sim_nmodels(self._h)

    def model_dims(self, m):
        d = (ctypes.c_int * 12)()
        self._check(self._L.oseh_sim_model_dims(self._h, m, d))
        return list(d)

    def model(self, m):
        d = self.model_dims(m)
        nxb, nyb, nzb = d[9], d[10], d[11]
        v = np.empty((nzb, nyb, nxb), dtype=np.float32)
        self._check(self._L.oseh_sim_model_velocity(self._h, m, v.ctypes.data_as(_fp)))
        R, K, rz = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.oseh_sim_model_params(self._h, m, ctypes.byref(R), ctypes.byref(K), None,
                                                  None, None, None, ctypes.byref(rz)))
        coef = np.empty(K.value + 1)
        gx = np.empty(nxb, dtype=np.float32)
        gy = np.empty(nyb, dtype=np.float32)
        gz = np.empty(nzb, dtype=np.float32)
        self._check(self._L.oseh_sim_model_params(self._h, m, None, None, coef.ctypes.data_as(_dp),
                                                  gx.ctypes.data_as(_fp), gy.ctypes.data_as(_fp),
                                                  gz.ctypes.data_as(_fp), None))
        return dict(dims=d, v=v, R=R.value, K=K.value, coef=coef, gx=gx, gy=gy, gz=gz, rec_z=rz.value)

    def nshots(self):
        return self._L.oseh_sim_nshots(self._h)

    def shot(self, k):
        m, nt = ctypes.c_int(), ctypes.c_int()
        src = (ctypes.c_int * 3)()
        self._check(self._L.oseh_sim_shot(self._h, k, ctypes.byref(m), src, ctypes.byref(nt), None, None))
        s = np.empty(nt.value, dtype=np.float32)
        h = np.empty(nt.value, dtype=np.float32)
        self._check(self._L.oseh_sim_shot(self._h, k, None, None, None, s.ctypes.data_as(_fp),
                                          h.ctypes.data_as(_fp)))
        return dict(model=m.value, src=list(src), nt=nt.value, s=s, h=h)
