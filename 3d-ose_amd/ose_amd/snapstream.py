"""Cost of snapshot streaming on a running simulation.

For a preset this times, with the wall clock around the steps and the final
drain: steps without snapshots; steps with snapshots streamed every E steps
(ose_set_snapshots: encode on the device, pinned copy on a second stream, a
callback that only counts bytes) in several encodings; and the synchronous
dump of the reference's commented block (ose_3d.cu:258-269: stop after the
step, copy the whole padded state of every shot to the host, go on).

    python -m ose_amd.snapstream --preset c1 --steps 40 --every 10
"""
import argparse
import json
import threading
import time

import numpy as np

from . import _lib
from .host import Simulation

CASES = (("f32", 3, False), ("bf16", 2, True), ("q16", 2, True), ("q8", 2, True))


def run(preset, steps, every, warmup=2, log=print):
    nt = warmup + steps * (len(CASES) + 3) + 8
    sim = Simulation(f"preset={preset} nt={nt}").setup()
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
        row = {"mode": "none", "ms_per_step": 1e3 * base / steps}
        log(json.dumps(row))
        out.append(row)
        for fmt, lanes, interior in CASES:
            got = {"n": 0, "bytes": 0}
            lock = threading.Lock()

            def count(user, inf, data):
                with lock:
                    got["n"] += 1
                    got["bytes"] += inf.contents.bytes

            cb = _lib.SNAPSHOT_FN(count)
            rc = hip.ose_set_snapshots(ctx, every, lanes, _lib.SNAP_FORMATS[fmt], int(interior), cb, None)
            if rc:
                raise RuntimeError(hip.ose_last_error(ctx).decode())
            t = wall(plain)
            hip.ose_set_snapshots(ctx, 0, 0, 0, 0, _lib.SNAPSHOT_FN(), None)
            row = {"mode": "stream", "format": fmt, "lanes": "pq" if lanes == 3 else "q",
                   "box": "interior" if interior else "padded", "every": every, "snapshots": got["n"],
                   "gb_per_snapshot": got["bytes"] / max(got["n"], 1) / 1e9,
                   "ms_per_step": 1e3 * t / steps, "overhead": t / base - 1.0}
            log(json.dumps(row))
            out.append(row)

        n = max(int(np.prod(sim.model_dims(sim.shot(k)["model"])[9:12])) for k in range(shots))
        re = np.empty(n, dtype=np.float32)
        im = np.empty(n, dtype=np.float32)

        def sync_dump():
            for _ in range(steps // every):
                hip.ose_run(ctx, every)
                for k in range(shots):
                    rc = hip.ose_get_wavefield(ctx, k, re.ctypes.data_as(_lib.fp), im.ctypes.data_as(_lib.fp))
                    if rc:
                        raise RuntimeError(hip.ose_last_error(ctx).decode())

        t = wall(sync_dump)
        row = {"mode": "sync_dump", "format": "f32", "lanes": "pq", "box": "padded", "every": every,
               "ms_per_step": 1e3 * t / steps, "overhead": t / base - 1.0}
        log(json.dumps(row))
        out.append(row)
    finally:
        sim.close()
    return {"preset": preset, "models": info["models"], "shots": shots, "steps": steps, "runs": out}


def main():
    ap = argparse.ArgumentParser(description="snapshot streaming overhead")
    ap.add_argument("--preset", default="c1")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--every", type=int, default=10)
    a = ap.parse_args()
    res = run(a.preset, a.steps, a.every, log=lambda m: print(m, flush=True))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
