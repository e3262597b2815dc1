"""ctypes loading of the in-tree native libraries.

libose_hip.so is the only way to the GPU: if it is missing or fails to load,
every call raises.  There is no CPU fallback.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")

_hip = None
_host = None

c_int, c_long, c_double, c_float = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float
P = ctypes.POINTER
fp, dp, ip = P(c_float), P(c_double), P(c_int)


class OseError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"OSE error {code}: {msg}")
        self.code = code


OSE_OK, OSE_EINVAL, OSE_ENOMEM, OSE_EDEVICE, OSE_ESTATE, OSE_ENUMERIC, OSE_EUNSUPPORTED = range(7)


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("nxb", c_int), ("nyb", c_int), ("nzb", c_int),
        ("bl", c_int), ("bf", c_int), ("bu", c_int),
        ("nx", c_int), ("ny", c_int), ("nz", c_int),
        ("dx", c_double), ("dy", c_double), ("dz", c_double), ("dt", c_double),
        ("R", c_double), ("K", c_int), ("coef", dp),
        ("taper_x", fp), ("taper_y", fp), ("taper_z", fp),
        ("rec_z", c_int),
    ]


class SnapshotInfo(ctypes.Structure):
    _fields_ = [
        ("shot_id", c_int), ("step", c_int), ("lanes", c_int), ("format", c_int),
        ("x0", c_int), ("y0", c_int), ("z0", c_int), ("nx", c_int), ("ny", c_int), ("nz", c_int),
        ("bytes", ctypes.c_size_t),
    ]


SNAPSHOT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, P(SnapshotInfo), ctypes.c_void_p)
SNAP_FORMATS = {"f32": 0, "bf16": 1, "q16": 2, "q8": 3}


def _proto(lib, name, args, res=c_int):
    f = getattr(lib, name)
    f.argtypes = args
    f.restype = res


def hip():
    """Load libose_hip.so (the HIP backend)."""
    global _hip
    if _hip is None:
        path = os.path.join(LIB_DIR, "libose_hip.so")
        if not os.path.exists(path):
            raise OseError(OSE_EDEVICE, f"{path} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        _proto(L, "ose_create", [P(vp), c_int])
        _proto(L, "ose_destroy", [vp])
        _proto(L, "ose_last_error", [vp], ctypes.c_char_p)
        _proto(L, "ose_add_model", [vp, P(ModelDesc), fp, ip])
        _proto(L, "ose_add_shot", [vp, c_int, c_int, c_int, c_int, fp, fp, c_int, ip])
        _proto(L, "ose_run", [vp, c_int])
        _proto(L, "ose_sync", [vp])
        _proto(L, "ose_timed_run", [vp, c_int, dp])
        _proto(L, "ose_get_traces", [vp, c_int, c_int, c_int, fp])
        _proto(L, "ose_set_traces", [vp, c_int, c_int, c_int, fp])
        _proto(L, "ose_set_receivers", [vp, c_int, ip, c_int])
        _proto(L, "ose_receiver_count", [vp, c_int, ip])
        _proto(L, "ose_set_source_profile", [vp, c_int, ip, fp, fp, fp])
        _proto(L, "ose_set_snapshots", [vp, c_int, c_int, c_int, c_int, SNAPSHOT_FN, vp])
        _proto(L, "ose_snapshot_drain", [vp])
        _proto(L, "ose_get_wavefield", [vp, c_int, fp, fp])
        _proto(L, "ose_set_wavefield", [vp, c_int, fp, fp])
        _proto(L, "ose_get_step", [vp, ip])
        _proto(L, "ose_set_step", [vp, c_int])
        _proto(L, "ose_apply_operator", [vp, c_int, fp, fp])
        _proto(L, "ose_apply_operator_noalloc", [vp, c_int, fp, fp])
        _proto(L, "ose_fft3d", [vp, c_int, c_int, c_int, fp, fp, c_int])
        _proto(L, "ose_set_options", [vp, c_int, c_int])
        _proto(L, "ose_pass_stats", [vp, c_int, dp, P(c_long), dp])
        _proto(L, "ose_reset_stats", [vp])
        _proto(L, "ose_copy_probe", [vp, ctypes.c_size_t, c_int, dp])
        _proto(L, "ose_pattern_probe", [vp, c_int, c_int, c_int, dp])
        _proto(L, "ose_pass_probe", [vp, c_int, c_int, c_int, dp])
        _proto(L, "ose_set_autotune", [vp, c_int])
        _proto(L, "ose_get_tuning", [vp, c_int, ip])
        _proto(L, "ose_set_tuning", [vp, c_int, ip])
        _proto(L, "ose_get_tuning_modes", [vp, c_int, ip])
        _proto(L, "ose_set_tuning_modes", [vp, c_int, ip])
        _proto(L, "ose_launch_shapes", [vp, c_int, ip, ip, ip, ip])
        _proto(L, "ose_boundary_net_size", [])
        _proto(L, "ose_set_boundary_net", [vp, c_int, fp, ctypes.c_size_t, c_double, c_int])
        _proto(L, "ose_boundary_status", [vp, c_int, ip, ip])
        _proto(L, "ose_trip_boundary_net", [vp, c_int])
        _proto(L, "ose_bnet_create", [P(vp), c_int] + [c_int] * 9 + [fp])
        _proto(L, "ose_bnet_destroy", [vp])
        _proto(L, "ose_bnet_error", [vp], ctypes.c_char_p)
        _proto(L, "ose_bnet_set_params", [vp, fp, ctypes.c_size_t])
        _proto(L, "ose_bnet_get_params", [vp, fp, ctypes.c_size_t])
        _proto(L, "ose_bnet_forward", [vp, fp, fp])
        _proto(L, "ose_bnet_gradient", [vp, fp, fp, fp, fp])
        _proto(L, "ose_bnet_add_sample", [vp, fp, fp, ip])
        _proto(L, "ose_bnet_train", [vp, c_int, c_float, ctypes.c_ulonglong, fp])
        _proto(L, "ose_supported_extent", [c_int])
        _proto(L, "ose_extent_is_supported", [c_int])
        _proto(L, "ose_extent_kind", [c_int])
        _proto(L, "ose_set_fft_path", [vp, c_int])
        _proto(L, "ose_accumulation_plan", [c_int, ip])
        _proto(L, "ose_set_streams", [vp, c_int])
        _hip = L
    return _hip


def host():
    """Load libose_host.so (C++ host library: wavelets, Bessel, models, I/O)."""
    global _host
    if _host is None:
        path = os.path.join(LIB_DIR, "libose_host.so")
        if not os.path.exists(path):
            raise OseError(OSE_EINVAL, f"{path} is not built (run __graft_entry__.build())")
        _host = ctypes.CDLL(path)
        from . import _host_protos
        _host_protos.declare(_host)
    return _host


def hip_symbols():
    """Names declared in include/ose_api.h (checked against the .so exports)."""
    return [
        "ose_create", "ose_destroy", "ose_last_error", "ose_add_model", "ose_add_shot",
        "ose_run", "ose_sync", "ose_timed_run", "ose_get_traces", "ose_set_traces", "ose_set_receivers",
        "ose_receiver_count", "ose_set_source_profile", "ose_set_snapshots", "ose_snapshot_drain", "ose_get_wavefield",
        "ose_set_wavefield", "ose_get_step", "ose_set_step", "ose_apply_operator", "ose_apply_operator_noalloc",
        "ose_fft3d", "ose_set_options", "ose_pass_stats", "ose_reset_stats",
        "ose_supported_extent", "ose_extent_is_supported", "ose_extent_kind", "ose_set_fft_path", "ose_set_autotune", "ose_get_tuning",
        "ose_set_tuning", "ose_get_tuning_modes", "ose_set_tuning_modes", "ose_launch_shapes",
        "ose_copy_probe", "ose_pattern_probe", "ose_pass_probe", "ose_boundary_net_size", "ose_set_boundary_net",
        "ose_boundary_status", "ose_trip_boundary_net", "ose_bnet_create", "ose_bnet_destroy", "ose_bnet_error",
        "ose_bnet_set_params", "ose_bnet_get_params", "ose_bnet_forward", "ose_bnet_gradient",
        "ose_bnet_add_sample", "ose_bnet_train", "ose_accumulation_plan", "ose_set_streams",
    ]
