"""ctypes prototypes of include/ose_host.h."""
import ctypes

c_int, c_double, c_float, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_size_t
c_u64 = ctypes.c_uint64
P = ctypes.POINTER
fp, dp, ip, vp, cp = P(c_float), P(c_double), P(c_int), ctypes.c_void_p, ctypes.c_char_p

PROTOS = {
    "oseh_judge_odd": ([ip, ip], c_int),
    "oseh_ricker": ([dp, c_int, c_double, c_double, c_double, c_int], None),
    "oseh_hilbert": ([dp, dp, c_int, c_int], None),
    "oseh_bessel_all": ([c_double, c_int, dp], c_int),
    "oseh_coefficients": ([c_double, c_int, dp], c_int),
    "oseh_spectral_radius": ([c_double, c_double, c_double, c_double], c_double),
    "oseh_order_reference": ([c_double], c_int),
    "oseh_order_tail": ([c_double, c_double], c_int),
    "oseh_taper": ([fp, c_int, c_int, c_int, c_int, c_double], None),
    "oseh_read_velocity_xyz": ([cp, c_int, c_int, c_int, fp], c_int),
    "oseh_pad_velocity": ([fp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, fp], None),
    "oseh_model": ([c_int, c_int, c_int, c_int, c_u64, c_double, c_double, c_int, fp], c_int),
    "oseh_rng_next": ([P(c_u64)], c_u64),
    "oseh_rng_uniform": ([P(c_u64)], c_double),
    "oseh_rng_normal": ([P(c_u64)], c_double),
    "oseh_rng_seed": ([P(c_u64), c_u64], None),
    "oseh_write_raw": ([cp, fp, c_size_t], c_int),
    "oseh_write_su": ([cp, fp, c_int, c_int, c_double, dp, dp, dp, c_double, c_double, c_double, c_int], c_int),
    "oseh_read_su": ([cp, fp, c_size_t, ip, ip, dp], c_int),
    "oseh_write_segy": ([cp, fp, c_int, c_int, c_double, dp, dp, dp, c_double, c_double, c_double, c_int], c_int),
    "oseh_read_segy": ([cp, fp, c_size_t, ip, ip, dp], c_int),
    "oseh_write_snapshot": ([cp, fp, c_int, c_int, c_int], c_int),
    "oseh_snapshot_lane_bytes": ([c_int, c_int, c_int, c_int], c_size_t),
    "oseh_snapshot_lane": ([c_int, c_int, c_int, c_int, c_int, c_int, vp, vp], c_int),
    "oseh_decode_snapshot": ([c_int, c_int, c_int, c_int, vp, fp], c_int),
    "oseh_write_osnap": ([cp, c_int, c_int, c_int, c_int, ip, vp], c_int),
    "oseh_read_osnap": ([cp, fp, c_size_t, ip], c_int),
    "oseh_sim_stream_snapshots": ([vp, cp, c_int], c_int),
    "oseh_sim_create": ([P(vp), cp], c_int),
    "oseh_sim_build": ([vp], c_int),
    "oseh_sim_setup": ([vp], c_int),
    "oseh_sim_run": ([vp, c_int, dp], c_int),
    "oseh_sim_ctx": ([vp], vp),
    "oseh_read_params": ([cp, fp, c_size_t, P(c_size_t)], c_int),
    "oseh_sim_info": ([vp, ctypes.c_char_p, c_size_t], c_int),
    "oseh_sim_write": ([vp, cp], c_int),
    "oseh_sim_snapshot": ([vp, cp], c_int),
    "oseh_sim_checkpoint": ([vp, cp], c_int),
    "oseh_sim_restore": ([vp, cp], c_int),
    "oseh_sim_error": ([vp], cp),
    "oseh_sim_destroy": ([vp], None),
    "oseh_sim_nmodels": ([vp], c_int),
    "oseh_sim_model_dims": ([vp, c_int, ip], c_int),
    "oseh_sim_model_velocity": ([vp, c_int, fp], c_int),
    "oseh_sim_model_params": ([vp, c_int, dp, ip, dp, fp, fp, fp, ip], c_int),
    "oseh_sim_nshots": ([vp], c_int),
    "oseh_gaussian_profile": ([c_double, c_int, c_int, ip, fp, c_int], c_int),
    "oseh_sim_receivers": ([vp, c_int, ip, ip], c_int),
    "oseh_sim_shot": ([vp, c_int, ip, ip, ip, fp, fp], c_int),
}


def declare(lib):
    for name, (args, res) in PROTOS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
