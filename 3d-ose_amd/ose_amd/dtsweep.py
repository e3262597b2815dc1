"""Large-time-step study: the same physical run at several time steps.

The one-step extrapolation evaluates exp(-i dt L) to the series tolerance for
any dt; the order K grows with R dt (ose_3d.cu:196-197), so a step costs
more but far fewer steps are needed.  This sweeps dt over one seeded layered
model and reports, per dt, the order K, the wall time per simulated second and
the surface seismogram's rel-L2 against the smallest dt at common times.

The source is the same physical signal at every dt: a Ricker wavelet with
its delay in seconds, injected as s + i H[s] (H the normalised discrete
Hilbert transform, so u stays analytic) with amplitude proportional to dt
(a per-step increment of a rate).

    python -m ose_amd.dtsweep [--n 128] [--T 0.512] [--dts 0.5,1,2,4,8]
"""
import argparse
import json
import time

import numpy as np

from . import host as H
from .backend import Context
from .boundary import Grid, extent


def run_one(vi, dt, T, fm, pad, delay_s, log=print):
    n = vi.shape[::-1]
    after = tuple(extent(n[a], pad, pad) for a in range(3))
    g = Grid(vi, (pad,) * 3, after, dt=dt)
    nt = int(round(T / dt))
    t = np.arange(nt) * dt - delay_s
    arg = (np.pi * fm * t) ** 2
    s = (1.0 - 2.0 * arg) * np.exp(-arg) * (dt / 1e-3)
    h = H.hilbert(s.astype(np.float64), normalize=True)
    src = (n[0] // 2, n[1] // 2, 2)
    with Context(0) as c:
        m = g.add_to(c)
        sid = c.add_shot(m, (g.origin[0] + src[0], g.origin[1] + src[1], g.origin[2] + src[2]), s, h)
        c.run(1)                                      # set-up, tuning and graph outside the timing
        c.step = 0
        re, im = c.wavefield(sid)
        c.set_wavefield(sid, np.zeros_like(re), np.zeros_like(im))
        t0 = time.time()
        c.run(nt)
        c.sync()
        wall = time.time() - t0
        tr = c.traces(sid)
    padded = int(np.prod(g.shape))
    row = {"dt_ms": dt * 1e3, "K": g.K, "steps": nt, "wall_s": round(wall, 4),
           "wall_per_simulated_s": wall / T, "padded": list(g.shape[::-1]),
           "padded_gcell_steps_per_s": padded * nt / wall / 1e9}
    log(f"[dtsweep] dt {dt * 1e3:g} ms: K={g.K}, {nt} steps in {wall:.3f} s")
    return row, tr


def sweep(n=128, T=0.512, dts=(0.5e-3, 1e-3, 2e-3, 4e-3, 8e-3), fm=15.0, pad=40, seed=7, log=print):
    vi = H.model("layered8", n, n, n, seed=seed)
    delay_s = 0.1
    runs = [run_one(vi, dt, T, fm, pad, delay_s, log) for dt in dts]
    coarse = max(dts)
    ref_row, ref_tr = runs[int(np.argmin(dts))]
    ref_dt = min(dts)
    ref = ref_tr[:, ::int(round(coarse / ref_dt))]
    out = []
    for (row, tr), dt in zip(runs, dts):
        sub = tr[:, ::int(round(coarse / dt))]
        k = min(sub.shape[1], ref.shape[1])
        row["rel_l2_vs_smallest_dt"] = float(np.linalg.norm(sub[:, :k] - ref[:, :k]) / np.linalg.norm(ref[:, :k]))
        out.append(row)
    return out


def main():
    ap = argparse.ArgumentParser(description="accuracy and cost of large OSE time steps")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--T", type=float, default=0.512)
    ap.add_argument("--dts", default="0.5,1,2,4,8", help="time steps in ms")
    ap.add_argument("--fm", type=float, default=15.0)
    args = ap.parse_args()
    dts = tuple(float(x) * 1e-3 for x in args.dts.split(","))
    for r in sweep(n=args.n, T=args.T, dts=dts, fm=args.fm, log=lambda m: print(m, flush=True)):
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
