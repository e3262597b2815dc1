"""Extended (profile) sources on the GPU (ose_set_source_profile): the unit
profile is the point source, a Gaussian profile matches the FP64 oracle's
profile run, and with a profile band-limited below the temporal Nyquist a
large time step reproduces the small-step seismogram (a point source does
not: its per-step impulses alias into the grid's high-frequency modes)."""
import numpy as np
import pytest

import pyoracle as orc
from _cases import Case, hilbert_unnorm, layered, rel, ricker

pytestmark = pytest.mark.gpu


def _profiles(H, case, cell, sigma):
    n = (case.nxb, case.nyb, case.nzb)
    return [H.gaussian_profile(sigma, cell[a], n[a]) for a in range(3)]


def test_unit_profile_is_the_point_source():
    import ose_amd
    print("unit profile vs point", flush=True)
    nt = 40
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=41))
    s = ricker(nt, 30.0, case.dt, delay=8)
    h = hilbert_unnorm(s)
    cell = (case.bl + 18, case.bf + 10, case.bu + 5)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        a = c.add_shot(m, cell, s, h)
        b = c.add_shot(m, (0, 0, 0), s, h)            # the point comes from the profile
        c.set_source_profile(b, cell, [1.0], [1.0], [1.0])
        c.run(nt)
        ta, tb = c.traces(a), c.traces(b)
        (ra, ia), (rb, ib) = c.wavefield(a), c.wavefield(b)
    assert np.abs(tb - ta).max() <= 1e-6 * np.abs(ta).max()
    assert np.abs(ib - ia).max() <= 1e-6 * np.abs(ia).max()
    assert np.abs(rb - ra).max() <= 1e-6 * np.abs(ra).max()


def test_gaussian_profile_matches_oracle():
    import ose_amd
    from ose_amd import host as H
    print("gaussian profile vs oracle", flush=True)
    nt = 60
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=42))
    s = ricker(nt, 30.0, case.dt, delay=8)
    h = hilbert_unnorm(s)
    cell = (case.bl + 18, case.bf + 10, case.bu + 8)
    prof = _profiles(H, case, cell, 1.6)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, cell, s, h)
        c.set_source_profile(sid, tuple(p[0] for p in prof), *(p[1] for p in prof))
        with pytest.raises(ose_amd.OseError):
            c.set_source_profile(sid, (0, 0, 0), np.ones(65), [1.0], [1.0])      # wider than 64
        c.run(nt)
        tr = c.traces(sid)
        re, im = c.wavefield(sid)
    full = []
    for a, (start, vals) in enumerate(prof):
        v = np.zeros((case.nxb, case.nyb, case.nzb)[a])
        v[start:start + vals.size] = vals
        full.append(v)
    otr, P, Q = orc.run_profile(case.oracle(), nt, full[0], full[1], full[2], s, h, case.bu, case.bl, case.bf,
                                case.nx, case.ny)
    assert rel(tr, otr) < 1e-5
    assert rel(re + 1j * im, P + 1j * Q) < 1e-5


def test_band_limited_source_allows_large_time_steps():
    """Seismogram and interior field at dt = 4 and 8 ms (R dt up to 19.6)
    against dt = 1 ms, the source a Gaussian profile whose spectrum is 1e-3
    at the 8 ms Nyquist (the same in every run): within 1 %; the point source
    at 8 ms is swamped by its aliases."""
    from ose_amd import dtsweep as D
    print("large steps with a band-limited source", flush=True)
    rows = D.sweep(n=128, T=0.224, dts=(1e-3, 4e-3, 8e-3), sigma="auto", log=lambda m: print(m, flush=True))
    for r in rows[1:]:
        print(r["dt_ms"], r["K"], r["rel_l2_vs_smallest_dt"], r["field_rel_l2_vs_smallest_dt"], flush=True)
        assert r["rel_l2_vs_smallest_dt"] < 0.01 and r["field_rel_l2_vs_smallest_dt"] < 0.01
    assert rows[2]["K"] >= 30 and rows[2]["src_sigma_cells"] > 1.0
    point = D.sweep(n=128, T=0.224, dts=(1e-3, 8e-3), log=lambda m: None)
    assert point[1]["rel_l2_vs_smallest_dt"] > 0.5


def test_driver_large_step_keys(tmp_path):
    """ose3d with 4 ms steps, an automatic source profile and the analytic
    wavelet pair: the order follows R dt and the run completes."""
    import json
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "3d-ose_amd", "bin", "ose3d")
    print("ose3d dt=0.004 src_sigma=auto hilbert=analytic", flush=True)
    # the Ricker is delayed 40 samples plus 1/fm: it peaks near step 48 at 4 ms
    p = subprocess.run([exe, "preset=ref", "nt=90", "dt=0.004", "src_sigma=auto", "hilbert=analytic",
                        f"out={tmp_path}"], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    out = json.loads(p.stdout.strip().splitlines()[-1])
    mod = out["setup"]["models"][0]
    assert out["steps"] == 90 and mod["K"] > mod["R"] * 0.004 and mod["K"] >= 12
    tr = np.fromfile(tmp_path / "shot0.bin", dtype=np.float32)
    assert tr.size == 101 * 101 * 90 and np.all(np.isfinite(tr)) and np.abs(tr).max() > 0
    bad = subprocess.run([exe, "preset=ref", "nt=5", "hilbert=upside", f"out={tmp_path}"], capture_output=True,
                         text=True, timeout=120)
    assert bad.returncode != 0
