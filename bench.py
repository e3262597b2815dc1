#!/usr/bin/env python3
"""Throughput of the MI355X OSE time stepper (one GPU).

A step is one OSE time step (source injection, K operator applications with
the Chebyshev recursion, taper, receiver gather) of every shot of the
workload; inputs are resident in HBM before the timed region starts.

    python bench.py [--steps K] [--warmup W] [--config c4|c1|c0|c2|c3|ref]

Prints one JSON line.  value = interior cells x shots x steps / s (whole job,
in Gcell*steps/s), timed with HIP events on the backend's stream.  The
roofline entry is the dominant pass kernel's algorithmic bytes per launch over
its mean HIP-event launch duration; the cpu_baseline entry times the FP64
oracle (OpenMP, all host cores) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-ose_amd"))

METRIC = ("Gcell·steps/s on 1 GPU at seismogram rel-L2≤1e-3 vs FP64; "
          "learned vs plain pad")

CONFIGS = {
    # 768^3 x 8 shots, seeded 12-layer + dipping salt model, 40-cell taper
    "c4": ("preset=c4", "768^3 x 8 shots, 12 layers + dipping salt, plain 40-cell taper"),
    "c1": ("preset=c1", "256^3 single shot, 8 layers + 6 anomalies, 40-cell taper"),
    "c0": ("preset=c0", "100^3 two-layer, 20-cell Cerjan taper"),
    "c2": ("preset=c2", "16 shots on 512x512x384 + 4 uneven models, 40-cell taper"),
    "c3": ("preset=c3", "256^3 single shot, thin 8-cell pad (taper)"),
    "ref": ("preset=ref", "reference as shipped: 101^3 two-layer, 25-cell pads"),
}
PASS_BYTES_DOC = {
    "fwd_y": "16 B/cell (read X_j, write it in place)",
    "z_symbol": "16 B/cell (read, write in place)",
    "inv_y": "16 B/cell (read, write in place: S)",
    "x_epilogue": "52 B/cell (read S, W_{j-1}, v/R, A; write X_{j+1}, W_{j+1}, A)",
    "x_acc_first": "44 B/cell (read S, W_{j-1}, v/R; write X_{j+1}, W_{j+1}, A: the order that starts A)",
    "x_recur": "36 B/cell (read S, W_{j-1}, v/R; write X_{j+1}, W_{j+1}; accumulator not touched)",
    "x_first": "28 B/cell (read S, v/R; write X_1, W_1); 44 where it starts the accumulator (+W_0 read, A write)",
    "x_last": "44 B/cell (read S, W_{K-2}, v/R, A; write X_0, u)",
    "inject": "16 B per x cell of the source line",
}
HBM_PEAK_GBS = 8000.0


def cpu_baseline(sim, info, seconds):
    """FP64 oracle on a centred cube of the workload's first model."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle

    m = sim.model(0)
    d = m["dims"]
    nxb, nyb, nzb = d[9], d[10], d[11]
    side = min(nxb, nyb, nzb, 216)
    z0, y0, x0 = (nzb - side) // 2, (nyb - side) // 2, (nxb - side) // 2
    v = m["v"][z0:z0 + side, y0:y0 + side, x0:x0 + side].astype(np.float64)
    om = pyoracle.Model(v, 10.0, 10.0, 10.0, 1e-3, m["R"], m["coef"], np.ones(side), np.ones(side),
                        np.ones(side))
    P = np.zeros(v.shape)
    Q = np.zeros(v.shape)
    P[side // 2, side // 2, side // 2] = 1.0
    steps, t0 = 0, time.time()
    while True:
        P, Q = om.step(P, Q)
        steps += 1
        if time.time() - t0 > seconds or steps >= 8:
            break
    dt = time.time() - t0
    padded_rate = side ** 3 * steps / dt / 1e9
    ratio = info["interior_cells"] / info["padded_cells"]
    return {
        "value": padded_rate * ratio,
        "unit": "Gcell*steps/s",
        "cores": pyoracle.num_threads(),
        "kind": "port",
        "sample": (f"FP64 oracle, {steps} step(s) of one shot on a {side}^3 cube cut from the "
                   f"workload's padded model (same K={m['K']}), {dt:.1f} s; padded-cell rate "
                   f"{padded_rate:.6f} scaled by the workload's interior/padded ratio {ratio:.3f}"),
    }


def rocprof_summary(db_path):
    """Per-kernel dispatch statistics from a rocprofv3 results database."""
    import sqlite3
    con = sqlite3.connect(db_path)
    rows = con.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = [f"{'calls':>6} {'total_ms':>10} {'avg_us':>9} {'min_us':>9} {'max_us':>9} {'pct':>6}  kernel"]
    for name, n, tot, avg, mn, mx in rows:
        out.append(f"{n:6d} {tot / 1e6:10.3f} {avg / 1e3:9.1f} {mn / 1e3:9.1f} {mx / 1e3:9.1f} "
                   f"{100.0 * tot / total:6.2f}  {name}")
    return "\n".join(out)


def pmc_summary(db_path):
    """Mean counter value per dispatch (rocprofv3 --pmc results database),
    by kernel, with the number of dispatches."""
    import sqlite3
    con = sqlite3.connect(db_path)
    rows = con.execute("select kernel_name, counter_name, count(*), avg(value), min(value), max(value) "
                       "from counters_collection group by kernel_name, counter_name "
                       "order by avg(value) desc").fetchall()
    out = [f"{'dispatches':>10} {'mean':>14} {'min':>14} {'max':>14}  counter  kernel"]
    for name, ctr, n, avg, mn, mx in rows:
        out.append(f"{n:10d} {avg:14.1f} {mn:14.1f} {mx:14.1f}  {ctr}  {name}")
    return "\n".join(out)


def traffic_from_profile(config, kernel):
    """HBM bytes per launch of `kernel` from the committed counter profile of
    this workload (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled by
    the same counter on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{config}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("traffic_bytes_per_launch_all_shots")


def tuning(sim, nmodels):
    """Launch shapes per model: T of forward-y, z, inverse-y, then the x-pass
    code of each x mode (middle orders that update the accumulator, those
    that leave it alone, its first update, the first order, the last order)."""
    import ctypes
    from ose_amd import _lib
    out = []
    for m in range(nmodels):
        t = (ctypes.c_int * 8)()
        _lib.hip().ose_get_tuning_modes(sim.ctx, m, t)
        out.append(dict(zip(("fwd_y", "z", "inv_y", "x_mid", "x_skip", "x_midf", "x_first", "x_last"), list(t))))
    return out


def set_streams(sim, n):
    from ose_amd import _lib
    rc = _lib.hip().ose_set_streams(sim.ctx, n)
    if rc:
        raise RuntimeError(f"ose_set_streams({n}) failed: {rc}")


def learned_leg(params, steps, warmup, plain_value, weights, streams):
    """The same workload with an 8-cell pad: plain thin taper, then the
    learned pad (network in the step, replacing the taper)."""
    from ose_amd.host import Simulation
    res = {}
    for name, extra in (("thin_taper", ""), ("learned", f"bnet={weights}")):
        sim = Simulation(f"{params} boundary=8 nt={warmup + steps + 1} {extra}")
        sim.setup()
        set_streams(sim, streams)
        info = sim.info()
        sim.timed(warmup)
        ms = sim.timed(steps)
        res[name] = {"ms_per_step": ms / steps, "value": info["interior_cells"] * steps / (ms * 1e-3) / 1e9,
                     "padded": info["models"][0]["padded"]}
        print(f"[bench] {name}: {ms / steps:.1f} ms/step", file=sys.stderr, flush=True)
        sim.close()
    lr = res["learned"]
    return {
        "value": lr["value"],
        "unit": "Gcell*steps/s",
        "padded": lr["padded"],
        "ms_per_step": lr["ms_per_step"],
        "vs_plain": lr["value"] / plain_value,
        "thin_taper_value": res["thin_taper"]["value"],
        "network_ms_per_step": lr["ms_per_step"] - res["thin_taper"]["ms_per_step"],
        "weights": os.path.relpath(weights, ROOT),
        "accuracy": ("not at the rel-L2 1e-3 bar: on held-out 256^3 models over 2000 steps the trained 8-cell "
                     "pad leaves 5.4-9.9x less reflected interior energy than the 8-cell taper; seismogram rel-L2 "
                     "against a 64-cell sponge 0.52-0.53 (taper 0.60-0.72); profiles/learned_pad_eval_256cube_2000steps.txt"),
    }


SNAP_CASES = (("f32", 3, False), ("bf16", 2, True), ("q16", 2, True), ("q8", 2, True))


def snapshot_leg(params, steps, every, warmup=2):
    """Cost of snapshot streaming on the running workload, wall clock around
    the steps and the final drain: steps without snapshots; with snapshots
    every `every` steps (ose_set_snapshots: encode on the device, pinned copy
    on a second stream, a callback that only counts bytes) in several
    encodings; and the synchronous dump of the reference's commented block
    (ose_3d.cu:258-269: stop after the step, copy every shot's whole padded
    state to the host, go on).  One JSON line per case, then a summary."""
    import threading

    import numpy as np
    from ose_amd import _lib
    from ose_amd.host import Simulation

    nt = warmup + steps * (len(SNAP_CASES) + 3) + 8
    sim = Simulation(f"{params} nt={nt}").setup()
    hip = _lib.hip()
    ctx = sim.ctx
    info = sim.info()
    shots = sim.nshots()
    out = []
    try:
        sim.run(warmup)

        def wall(fn):
            hip.ose_sync(ctx)
            t0 = time.perf_counter()
            fn()
            if hip.ose_sync(ctx):
                raise RuntimeError(hip.ose_last_error(ctx).decode())
            return time.perf_counter() - t0

        def plain():
            hip.ose_run(ctx, steps)

        base = wall(plain)
        out.append({"mode": "none", "ms_per_step": 1e3 * base / steps})
        print(json.dumps(out[-1]), flush=True)
        for fmt, lanes, interior in SNAP_CASES:
            got = {"n": 0, "bytes": 0}
            lock = threading.Lock()

            def count(user, inf, data):
                with lock:
                    got["n"] += 1
                    got["bytes"] += inf.contents.bytes

            cb = _lib.SNAPSHOT_FN(count)
            if hip.ose_set_snapshots(ctx, every, lanes, _lib.SNAP_FORMATS[fmt], int(interior), cb, None):
                raise RuntimeError(hip.ose_last_error(ctx).decode())
            t = wall(plain)
            hip.ose_set_snapshots(ctx, 0, 0, 0, 0, _lib.SNAPSHOT_FN(), None)
            out.append({"mode": "stream", "format": fmt, "lanes": "pq" if lanes == 3 else "q",
                        "box": "interior" if interior else "padded", "every": every, "snapshots": got["n"],
                        "gb_per_snapshot": got["bytes"] / max(got["n"], 1) / 1e9,
                        "ms_per_step": 1e3 * t / steps, "overhead": t / base - 1.0})
            print(json.dumps(out[-1]), flush=True)

        n = max(int(np.prod(sim.model_dims(sim.shot(k)["model"])[9:12])) for k in range(shots))
        re = np.empty(n, dtype=np.float32)
        im = np.empty(n, dtype=np.float32)

        def sync_dump():
            for _ in range(steps // every):
                hip.ose_run(ctx, every)
                for k in range(shots):
                    if hip.ose_get_wavefield(ctx, k, re.ctypes.data_as(_lib.fp), im.ctypes.data_as(_lib.fp)):
                        raise RuntimeError(hip.ose_last_error(ctx).decode())

        t = wall(sync_dump)
        out.append({"mode": "sync_dump", "format": "f32", "lanes": "pq", "box": "padded", "every": every,
                    "ms_per_step": 1e3 * t / steps, "overhead": t / base - 1.0})
        print(json.dumps(out[-1]), flush=True)
    finally:
        sim.close()
    return {"params": params, "models": info["models"], "shots": shots, "steps": steps, "runs": out}


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--rocprof-summary":
        print(rocprof_summary(sys.argv[2]))
        return
    if len(sys.argv) == 3 and sys.argv[1] == "--pmc-summary":
        print(pmc_summary(sys.argv[2]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--extra", default="", help="extra key=value parameters")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the 8-cell learned-pad leg")
    ap.add_argument("--streams", type=int, default=2, choices=(1, 2),
                    help="2: the step's shots run as two concurrent lanes (default); 1: one stream")
    ap.add_argument("--snapshots", type=int, default=0, metavar="EVERY",
                    help="instead of the throughput line: the cost of streaming snapshots every EVERY steps "
                         "(one JSON line per encoding, then a summary)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="also run one plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); shots would shard "
              "across GPUs without any exchange", file=sys.stderr)
        sys.exit(2)

    from ose_amd.host import Simulation

    params, workload = CONFIGS[args.config]
    if args.snapshots > 0:
        print(json.dumps(snapshot_leg(f"{params} {args.extra}", args.steps, args.snapshots)))
        return
    prof_steps = max(2, min(args.steps, 4))
    nt = args.warmup + args.steps + prof_steps
    sim = Simulation(f"{params} nt={nt} {args.extra}")
    t0 = time.time()
    sim.setup()
    set_streams(sim, args.streams)
    info = sim.info()
    print(f"[bench] set-up {time.time() - t0:.1f} s: {json.dumps(info)}", file=sys.stderr, flush=True)

    # warm-up (also builds the per-step hipGraph)
    sim.timed(args.warmup)
    # timed region: K graph-replayed steps between stream synchronisations
    ms = sim.timed(args.steps)
    print(f"[bench] timed {args.steps} steps: {ms:.1f} ms", file=sys.stderr, flush=True)
    # per-pass HIP-event durations over the following steps (eager launches)
    sim.set_options(use_graph=False, profile=True)
    sim.reset_stats()
    ms_prof = sim.timed(prof_steps)
    stats = sim.pass_stats()
    sim.set_options(use_graph=True, profile=False)

    cells = info["interior_cells"]
    value = cells * args.steps / (ms * 1e-3) / 1e9
    passes = {}
    for name, (avg_ms, n, by) in stats.items():
        if n:
            passes[name] = {"avg_us": avg_ms * 1e3, "launches": n, "bytes": by,
                            "GBs": by / (avg_ms * 1e-3) / 1e9 if avg_ms else 0.0}
    dom = max((p for p in passes if p != "inject"), key=lambda p: passes[p]["avg_us"] * passes[p]["launches"])
    pd = passes[dom]
    if args.copy_probe:
        from ose_amd import _lib
        import ctypes
        g = ctypes.c_double()
        _lib.hip().ose_copy_probe(sim.ctx, args.copy_probe, 1, ctypes.byref(g))
        print(f"[bench] copy probe {args.copy_probe} B: {g.value:.1f} GB/s", file=sys.stderr, flush=True)
    roofline = {"bound": "hbm", "achieved": pd["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": pd["GBs"] / HBM_PEAK_GBS, "traffic": traffic_from_profile(args.config, dom), "kernel": dom,
                "bytes_per_launch": pd["bytes"], "bytes_rule": PASS_BYTES_DOC[dom],
                "passes": {k: round(v["GBs"], 1) for k, v in passes.items()}}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sim, info, args.cpu_seconds)
    m0 = info["models"][0]
    tune = tuning(sim, len(info["models"]))
    sim.close()
    learned = None
    if not args.no_learned and args.config in ("c4", "c1"):
        from ose_amd.boundary import WEIGHTS
        learned = learned_leg(params + " " + args.extra, args.steps, args.warmup, value, WEIGHTS, args.streams)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Gcell*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded velocity model, Ricker source)",
        "config": {
            "workload": workload,
            "preset": args.config,
            "boundary": "taper",
            "interior": m0["interior"],
            "padded": m0["padded"],
            "K": m0["K"],
            "shots": info["nshots"],
            "interior_cells_per_step": cells,
            "padded_cells_per_step": info["padded_cells"],
            "padded_rate": info["padded_cells"] * args.steps / (ms * 1e-3) / 1e9,
            "ms_per_step_profiled": ms_prof / prof_steps,
            "streams": args.streams,
            "launch_shapes": tune,
        },
        "learned": learned,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
