"""Large-time-step study: the same physical run at several time steps.

The one-step extrapolation evaluates exp(-i dt L) to the series tolerance for
any dt; the order K grows with R dt (ose_3d.cu:196-197), so a step costs
more but far fewer steps are needed.  This sweeps dt over one seeded layered
model and reports, per dt, the order K, the wall time per simulated second and
the surface seismogram's rel-L2 against the smallest dt at common times.

The source is the same physical signal at every dt: a Ricker wavelet with
its delay in seconds and amplitude proportional to dt (a per-step increment
of a rate), injected as s - i H[s] (H the normalised discrete Hilbert
transform, H[cos] = sin).  The step propagates u by exp(-i L dt), under
which a mode of frequency w > 0 evolves as exp(-i w t): s - i H[s] is the
analytic signal of that sign.  The reference pairs P += s with Q += H[s]
(times nt, its FFT being unnormalised; ose_3d_lib.cu:575-583): with the
normalised transform that pairing is the opposite analytic signal, which
the propagator barely radiates, leaving a near-DC residual that depends on
dt; the unnormalised one is dominated by i nt H[s], half of which radiates.

With --sigma the source is a normalised Gaussian profile of that width in
cells ("auto": wide enough for the largest dt, sigma = 3.72 v dt / (pi d),
so its spectrum is below 1e-3 at that step's temporal Nyquist) at every dt:
a point injected every dt excites the grid's modes above pi/dt, whose
aliases swamp the signal at large dt; the band-limited profile does not.

    python -m ose_amd.dtsweep [--n 128] [--T 0.512] [--dts 0.5,1,2,4,8] [--sigma auto]
"""
import argparse
import json
import time

import numpy as np

from . import host as H
from .backend import Context
from .boundary import Grid, extent


def run_one(vi, dt, T, fm, pad, delay_s, log=print, sigma=0.0, mid=False, field=False, src=None, rec_z=0,
            taper="ref"):
    """One run; mid samples the wavelet at the middle of each step (the
    injection before a step stands for the source over that step), field
    also returns Im u at the end (interior); src: interior source cell
    (default: centre of the top, depth 2), rec_z: receiver plane depth."""
    n = vi.shape[::-1]
    after = tuple(extent(n[a], pad, pad) for a in range(3))
    g = Grid(vi, (pad,) * 3, after, dt=dt, taper=taper)
    nt = int(round(T / dt))
    t = (np.arange(nt) + (0.5 if mid else 0.0)) * dt - delay_s
    arg = (np.pi * fm * t) ** 2
    s = (1.0 - 2.0 * arg) * np.exp(-arg) * (dt / 1e-3)
    h = -H.hilbert(s.astype(np.float64), normalize=True)
    src = (n[0] // 2, n[1] // 2, 2) if src is None else src
    with Context(0) as c:
        m = g.add_to(c, rec_z=rec_z)
        cell = (g.origin[0] + src[0], g.origin[1] + src[1], g.origin[2] + src[2])
        sid = c.add_shot(m, cell, s, h)
        if sigma > 0:
            prof = [H.gaussian_profile(sigma, cell[a], g.shape[2 - a]) for a in range(3)]
            c.set_source_profile(sid, tuple(p[0] for p in prof), *(p[1] for p in prof))
        c.run(1)                                      # set-up, tuning and graph outside the timing
        c.step = 0
        re, im = c.wavefield(sid)
        c.set_wavefield(sid, np.zeros_like(re), np.zeros_like(im))
        t0 = time.time()
        c.run(nt)
        c.sync()
        wall = time.time() - t0
        tr = c.traces(sid)
        final = None
        if field:
            o = g.origin
            final = c.wavefield(sid)[1][o[2]:o[2] + n[2], o[1]:o[1] + n[1], o[0]:o[0] + n[0]]
    padded = int(np.prod(g.shape))
    row = {"dt_ms": dt * 1e3, "K": g.K, "steps": nt, "wall_s": round(wall, 4), "src_sigma_cells": sigma,
           "wall_per_simulated_s": wall / T, "padded": list(g.shape[::-1]),
           "padded_gcell_steps_per_s": padded * nt / wall / 1e9}
    log(f"[dtsweep] dt {dt * 1e3:g} ms: K={g.K}, {nt} steps in {wall:.3f} s")
    return (row, tr, final) if field else (row, tr)


def auto_sigma(v_src, dt, d=10.0):
    """Gaussian width (cells) whose spectrum is 1e-3 at the temporal Nyquist
    pi / (v dt): exp(-k^2 s^2 / 2) = 1e-3 at k = pi / (v dt)."""
    return 3.72 * v_src * dt / (np.pi * d)


def sweep(n=128, T=0.224, dts=(0.5e-3, 1e-3, 2e-3, 4e-3, 8e-3), fm=15.0, pad=40, seed=7, log=print, sigma=0.0,
          delay_s=0.08):
    """Runs at every dt with the source at the centre of the model and the
    receivers on the plane n/4 above it, over a window that ends before the
    pads' reflections return (the pad taper acts once per series order, so
    its absorption per unit time depends on dt).  Each row: the order K, the
    cost per simulated second, and the rel-L2 of the seismogram and of the
    interior field at T against the smallest dt."""
    vi = H.model("layered8", n, n, n, seed=seed)
    src = (n // 2, n // 2, n // 2)
    if sigma == "auto":
        sigma = auto_sigma(float(vi[src[2], src[1], src[0]]), max(dts))
    runs = [run_one(vi, dt, T, fm, pad, delay_s, log, sigma=float(sigma), field=True, src=src, rec_z=n // 4)
            for dt in dts]
    # compare at the common times t = m * max(dt), m = 1, 2, ...: trace column
    # i holds the state after step i + 1, i.e. at (i + 1) dt
    coarse = max(dts)
    ref_row, ref_tr, ref_f = runs[int(np.argmin(dts))]
    ref_dt = min(dts)
    r = int(round(coarse / ref_dt))
    ref = ref_tr[:, r - 1::r]
    out = []
    for (row, tr, fin), dt in zip(runs, dts):
        r = int(round(coarse / dt))
        sub = tr[:, r - 1::r]
        k = min(sub.shape[1], ref.shape[1])
        row["rel_l2_vs_smallest_dt"] = float(np.linalg.norm(sub[:, :k] - ref[:, :k]) / np.linalg.norm(ref[:, :k]))
        row["field_rel_l2_vs_smallest_dt"] = float(np.linalg.norm(fin - ref_f) / np.linalg.norm(ref_f))
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser(description="accuracy and cost of large OSE time steps")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--T", type=float, default=0.224)
    ap.add_argument("--dts", default="0.5,1,2,4,8", help="time steps in ms")
    ap.add_argument("--fm", type=float, default=15.0)
    ap.add_argument("--sigma", default="0", help="source profile width in cells, or auto")
    args = ap.parse_args()
    dts = tuple(float(x) * 1e-3 for x in args.dts.split(","))
    sigma = args.sigma if args.sigma == "auto" else float(args.sigma)
    for r in sweep(n=args.n, T=args.T, dts=dts, fm=args.fm, log=lambda m: print(m, flush=True), sigma=sigma):
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
