"""FP64 oracle run of the unreflected reference used to judge the learned
thin pad (ose_amd.boundary.compare_pads): held-out model seed 2018 (64^3
'train' model), taper-free, on the wide domain unreflected_pad gives for a
400-step window (67-cell pads -> 200^3), Ricker 20 Hz at (32, 32, 2), 64^2
surface receivers; the first nt steps.  The GPU test runs the same grid
and compares its FP32 traces (tests/test_gpu_configs.py).

Run: OMP_NUM_THREADS=3 python tests/golden/gen_unreflected_oracle.py [nt]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "3d-ose_amd"), os.path.join(ROOT, "oracle")]

from ose_amd import boundary as B  # noqa: E402
from ose_amd import host as H  # noqa: E402
import pyoracle  # noqa: E402

SEED, N, WINDOW = 2018, 64, 400


def grid():
    vi = H.model("train", N, N, N, seed=SEED)
    vm = float(vi.max())
    w = B.unreflected_pad(N, WINDOW, vmax=vm)
    g = B.Grid(vi, (w,) * 3, tuple(B.extent(N, w, w) for _ in range(3)), taper="none", vmax=vm)
    return g


def main():
    nt = int(sys.argv[1]) if len(sys.argv) > 1 else 100
    g = grid()
    nzb, nyb, nxb = g.shape
    s, h = B._wavelet(WINDOW, 20.0, 1e-3)
    src = (N // 2, N // 2, 2)
    bl, bf, bu = g.origin
    idx = ((bu + src[2]) * nyb + (bf + src[1])) * nxb + bl + src[0]
    om = pyoracle.Model(g.v.astype(np.float64), g.d, g.d, g.d, g.dt, g.R, g.coef,
                        *(t.astype(np.float64) for t in g.taper))
    t0 = time.time()
    tr, P, Q = om.run(nt, idx, s[:nt].astype(np.float64), h[:nt].astype(np.float64), bu, bl, bf, N, N)
    print(f"oracle {nt} steps on {nxb}x{nyb}x{nzb}, K={g.K}: {time.time() - t0:.0f} s")
    iy, ix = np.meshgrid(np.arange(0, N, 2), np.arange(0, N, 2), indexing="ij")
    sel = (iy * N + ix).ravel()
    np.savez(os.path.join(HERE, "unreflected_oracle.npz"), nt=nt, seed=SEED, padded=np.array([nxb, nyb, nzb]),
             K=g.K, idx=sel, traces=tr[sel].astype(np.float32), norm_full=np.linalg.norm(tr))


if __name__ == "__main__":
    main()
