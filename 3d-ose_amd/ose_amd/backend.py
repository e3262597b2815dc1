"""Python view of the HIP backend's C-ABI (include/ose_api.h).

Context mirrors the reference's device-side objects: a model is the padded
velocity plus the step parameters wavefield_extro_ori2 reads from globals
(ose_3d.cu:42-62), a shot is a source position with its two wavelets
(Xt_add_wavelet_sxyz, ose_3d_lib.cu:610-633), run() is the time loop body
(ose_3d.cu:248-275) and traces() the seismogram (ose_3d.cu:271-276).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import ModelDesc, OseError, c_double, c_int, c_long

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(_fp)


class Context:
    def __init__(self, device=0):
        self._L = _lib.hip()
        h = ctypes.c_void_p()
        rc = self._L.ose_create(ctypes.byref(h), device)
        if rc:
            raise OseError(rc, f"ose_create(device={device}) failed")
        self._h = h
        self._keep = []
        self.models = []
        self.shots = []

    # -- lifetime -----------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.ose_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc:
            msg = self._L.ose_last_error(self._h)
            raise OseError(rc, msg.decode() if msg else "")

    # -- set-up -------------------------------------------------------
    def add_model(self, v_padded, *, interior, origin, d, dt, R, coef, taper, rec_z=0):
        """v_padded: [nzb][nyb][nxb] velocity; interior = (nx, ny, nz);
        origin = (bl, bf, bu); d = (dx, dy, dz); coef: K+1 doubles;
        taper = (gx, gy, gz)."""
        v = _f32(v_padded)
        nzb, nyb, nxb = v.shape
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        gx, gy, gz = (_f32(t) for t in taper)
        desc = ModelDesc(nxb, nyb, nzb, origin[0], origin[1], origin[2],
                         interior[0], interior[1], interior[2], d[0], d[1], d[2], dt,
                         float(R), len(coef) - 1, coef.ctypes.data_as(_dp),
                         _ptr(gx), _ptr(gy), _ptr(gz), int(rec_z))
        mid = c_int()
        self._check(self._L.ose_add_model(self._h, ctypes.byref(desc), _ptr(v), ctypes.byref(mid)))
        self.models.append(dict(shape=v.shape, interior=tuple(interior), origin=tuple(origin)))
        return mid.value

    def add_shot(self, model, src, s, h):
        """src: padded (sx, sy, sz); s, h: wavelets added to Re u and Im u."""
        s, h = _f32(s), _f32(h)
        assert s.size == h.size
        sid = c_int()
        self._check(self._L.ose_add_shot(self._h, model, int(src[0]), int(src[1]), int(src[2]),
                                         _ptr(s), _ptr(h), s.size, ctypes.byref(sid)))
        self.shots.append(dict(model=model, nt=s.size))
        return sid.value

    def set_receivers(self, shot, xyz):
        """Record the shot at listed padded cells (an (n, 3) array of x, y, z)
        instead of the plane rec_z; an empty list restores the plane."""
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.int32).reshape(-1, 3))
        self._check(self._L.ose_set_receivers(self._h, shot, a.ctypes.data_as(ctypes.POINTER(c_int)), len(a)))

    def set_source_profile(self, shot, origin=None, px=None, py=None, pz=None):
        """Extended source: (s + i h) px py pz over the padded box starting at
        origin = (x0, y0, z0); origin None restores the point source."""
        if origin is None:
            self._check(self._L.ose_set_source_profile(self._h, shot, None, None, None, None))
            return
        px, py, pz = _f32(px), _f32(py), _f32(pz)
        box = (c_int * 6)(int(origin[0]), int(origin[1]), int(origin[2]), px.size, py.size, pz.size)
        self._check(self._L.ose_set_source_profile(self._h, shot, box, _ptr(px), _ptr(py), _ptr(pz)))

    def receiver_count(self, shot):
        n = c_int()
        self._check(self._L.ose_receiver_count(self._h, shot, ctypes.byref(n)))
        return n.value

    def set_snapshots(self, every, callback=None, lanes=3, fmt="f32", interior=False):
        """Stream snapshots every `every` steps (0 stops).  callback(info,
        data) runs on the library's snapshot thread with info a dict (shot_id,
        step, lanes, format, x0..nz, bytes) and data a bytes copy of the
        encoded snapshot (layout of ose_set_snapshots)."""
        if every == 0:
            self._check(self._L.ose_set_snapshots(self._h, 0, 0, 0, 0, _lib.SNAPSHOT_FN(), None))
            self._snap_cb = None
            return

        def trampoline(user, info, data):
            i = info.contents
            d = {k: getattr(i, k) for k, _ in _lib.SnapshotInfo._fields_}
            callback(d, ctypes.string_at(data, i.bytes))

        cb = _lib.SNAPSHOT_FN(trampoline)
        self._check(self._L.ose_set_snapshots(self._h, int(every), int(lanes), _lib.SNAP_FORMATS[fmt],
                                              int(bool(interior)), cb, None))
        self._snap_cb = cb          # keep the C callback alive while streaming

    def snapshot_drain(self):
        self._check(self._L.ose_snapshot_drain(self._h))

    def copy_probe(self, nbytes=1 << 30, reps=10):
        """GB/s of a plain device copy (read + write bytes)."""
        g = c_double()
        self._check(self._L.ose_copy_probe(self._h, int(nbytes), int(reps), ctypes.byref(g)))
        return g.value

    def set_autotune(self, on=True):
        self._check(self._L.ose_set_autotune(self._h, int(on)))

    def tuning(self, model):
        """Launch shapes in use: T per strided pass, P of the x pass
        (negative: half-LDS split exchange)."""
        out = (c_int * 4)()
        self._check(self._L.ose_get_tuning(self._h, model, out))
        return dict(zip(("T_fwd_y", "T_z", "T_inv_y", "P_x"), list(out)))

    def set_tuning(self, model, t_fwd_y, t_z, t_inv_y, p_x):
        arr = (c_int * 4)(t_fwd_y, t_z, t_inv_y, p_x)
        self._check(self._L.ose_set_tuning(self._h, model, arr))

    def tuning_modes(self, model):
        """All eight launch-shape codes of a model (ose_get_tuning_modes)."""
        out = (c_int * 8)()
        self._check(self._L.ose_get_tuning_modes(self._h, model, out))
        return list(out)

    def set_tuning_modes(self, model, shape):
        arr = (c_int * 8)(*shape)
        self._check(self._L.ose_set_tuning_modes(self._h, model, arr))

    def launch_shapes(self, model):
        """(strided codes usable on y and z, x-pass codes) compiled for a model."""
        ns, nx = c_int(16), c_int(16)
        st, xs = (c_int * 16)(), (c_int * 16)()
        self._check(self._L.ose_launch_shapes(self._h, model, st, ctypes.byref(ns), xs, ctypes.byref(nx)))
        return list(st)[:ns.value], list(xs)[:nx.value]

    def set_boundary_net(self, model, params, guard=2.0, replace_taper=True):
        """Enable the learned thin pad on a model (None removes it)."""
        if params is None:
            self._check(self._L.ose_set_boundary_net(self._h, model, None, 0, float(guard), 0))
            return
        p = _f32(params)
        self._check(self._L.ose_set_boundary_net(self._h, model, _ptr(p), p.size, float(guard),
                                                 int(replace_taper)))

    def boundary_status(self, model):
        a, t = c_int(), c_int()
        self._check(self._L.ose_boundary_status(self._h, model, ctypes.byref(a), ctypes.byref(t)))
        return dict(active=bool(a.value), tripped=bool(t.value))

    def trip_boundary_net(self, model):
        """Force the network's fallback to the plain taper (as a guard trip)."""
        self._check(self._L.ose_trip_boundary_net(self._h, model))

    def set_options(self, use_graph=True, profile=False):
        self._check(self._L.ose_set_options(self._h, int(use_graph), int(profile)))

    def set_fft_path(self, path):
        """1: runtime-length kernels on every axis of the models added
        afterwards; 0: compiled kernels where they exist (default)."""
        self._check(self._L.ose_set_fft_path(self._h, int(path)))

    def set_streams(self, n):
        """2: the step's shots run as two concurrent lanes; 1: one stream."""
        self._check(self._L.ose_set_streams(self._h, int(n)))

    # -- stepping -----------------------------------------------------
    def run(self, nsteps):
        self._check(self._L.ose_run(self._h, int(nsteps)))

    def sync(self):
        self._check(self._L.ose_sync(self._h))

    def timed_run(self, nsteps):
        ms = c_double()
        self._check(self._L.ose_timed_run(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    @property
    def step(self):
        it = c_int()
        self._check(self._L.ose_get_step(self._h, ctypes.byref(it)))
        return it.value

    @step.setter
    def step(self, it):
        self._check(self._L.ose_set_step(self._h, int(it)))

    # -- results ------------------------------------------------------
    def traces(self, shot, it0=0, nsteps=None):
        if nsteps is None:
            nsteps = self.shots[shot]["nt"] - it0
        out = np.empty((self.receiver_count(shot), nsteps), dtype=np.float32)
        self._check(self._L.ose_get_traces(self._h, shot, it0, nsteps, _ptr(out)))
        return out

    def wavefield(self, shot):
        shape = self.models[self.shots[shot]["model"]]["shape"]
        re = np.empty(shape, dtype=np.float32)
        im = np.empty(shape, dtype=np.float32)
        self._check(self._L.ose_get_wavefield(self._h, shot, _ptr(re), _ptr(im)))
        return re, im

    def set_wavefield(self, shot, re, im):
        re, im = _f32(re), _f32(im)
        self._check(self._L.ose_set_wavefield(self._h, shot, _ptr(re), _ptr(im)))

    def apply_operator(self, model, u):
        u = np.ascontiguousarray(u, dtype=np.complex64)
        out = np.empty_like(u)
        self._check(self._L.ose_apply_operator(self._h, model, u.view(np.float32).ctypes.data_as(_fp),
                                               out.view(np.float32).ctypes.data_as(_fp)))
        return out

    def apply_operator_noalloc(self, model, u):
        """apply_operator through the entry point that allocates nothing."""
        u = np.ascontiguousarray(u, dtype=np.complex64)
        out = np.empty_like(u)
        self._check(self._L.ose_apply_operator_noalloc(self._h, model, u.view(np.float32).ctypes.data_as(_fp),
                                                       out.view(np.float32).ctypes.data_as(_fp)))
        return out

    def fft3d(self, a, sign=-1):
        a = np.ascontiguousarray(a, dtype=np.complex64)
        out = np.empty_like(a)
        nz, ny, nx = a.shape
        self._check(self._L.ose_fft3d(self._h, nx, ny, nz, a.view(np.float32).ctypes.data_as(_fp),
                                      out.view(np.float32).ctypes.data_as(_fp), sign))
        return out

    def pass_stats(self):
        """{pass: (avg_ms, launches, bytes_per_launch)} for the 9 pass kinds."""
        out = {}
        names = ["fwd_y", "z_symbol", "inv_y", "x_epilogue", "x_first", "x_last", "inject", "x_recur", "x_acc_first"]
        for i, n in enumerate(names):
            ms, cnt, by = c_double(), c_long(), c_double()
            self._check(self._L.ose_pass_stats(self._h, i, ctypes.byref(ms), ctypes.byref(cnt), ctypes.byref(by)))
            out[n] = (ms.value, cnt.value, by.value)
        return out

    def reset_stats(self):
        self._check(self._L.ose_reset_stats(self._h))


def supported_extent(n):
    return _lib.hip().ose_supported_extent(int(n))


def extent_is_supported(n):
    return bool(_lib.hip().ose_extent_is_supported(int(n)))


def extent_kind(n):
    """1: compiled in several launch shapes, 3: compiled in one, 2: the
    runtime-length path, 0: unsupported."""
    return _lib.hip().ose_extent_kind(int(n))


def boundary_net_size():
    return _lib.hip().ose_boundary_net_size()


class BoundaryNet:
    """Training object of the learned thin-pad network (ose_bnet_*)."""

    def __init__(self, v_padded, origin, interior, device=0):
        self._L = _lib.hip()
        v = _f32(v_padded)
        nzb, nyb, nxb = v.shape
        h = ctypes.c_void_p()
        rc = self._L.ose_bnet_create(ctypes.byref(h), device, nxb, nyb, nzb, origin[0], origin[1], origin[2],
                                     interior[0], interior[1], interior[2], _ptr(v))
        if rc:
            raise OseError(rc, "ose_bnet_create failed")
        self._h = h
        self.shape = v.shape
        self.n = boundary_net_size()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.ose_bnet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise OseError(rc, self._L.ose_bnet_error(self._h).decode())

    @staticmethod
    def _c(a):
        return np.ascontiguousarray(a, dtype=np.complex64)

    def params(self):
        p = np.empty(self.n, dtype=np.float32)
        self._check(self._L.ose_bnet_get_params(self._h, _ptr(p), self.n))
        return p

    def set_params(self, p):
        p = _f32(p)
        self._check(self._L.ose_bnet_set_params(self._h, _ptr(p), p.size))

    def forward(self, u):
        u = self._c(u)
        d = np.empty_like(u)
        self._check(self._L.ose_bnet_forward(self._h, u.view(np.float32).ctypes.data_as(_fp),
                                             d.view(np.float32).ctypes.data_as(_fp)))
        return d

    def gradient(self, u, t):
        u, t = self._c(u), self._c(t)
        loss = ctypes.c_float()
        g = np.empty(self.n, dtype=np.float32)
        self._check(self._L.ose_bnet_gradient(self._h, u.view(np.float32).ctypes.data_as(_fp),
                                              t.view(np.float32).ctypes.data_as(_fp), ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def add_sample(self, u, t):
        u, t = self._c(u), self._c(t)
        cnt = c_int()
        self._check(self._L.ose_bnet_add_sample(self._h, u.view(np.float32).ctypes.data_as(_fp),
                                                t.view(np.float32).ctypes.data_as(_fp), ctypes.byref(cnt)))
        return cnt.value

    def train(self, iters, lr=1e-3, seed=0):
        losses = np.empty(iters, dtype=np.float32)
        self._check(self._L.ose_bnet_train(self._h, int(iters), float(lr), int(seed), _ptr(losses)))
        return losses
