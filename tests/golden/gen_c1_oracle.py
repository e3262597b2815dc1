"""FP64 oracle seismogram of configuration c1 (256^3 layered model with
anomalies, 40-cell taper, 336^3 padded, K = 12, surface receivers), the
first nt steps, for the GPU parity test at full size.

The model, source wavelets, order K, coefficients and taper are taken from
the C++ host library's preset (those are pinned against the oracle in
tests/test_host.py); the time loop is the FP64 oracle.  Every 8th receiver
in x and y is committed (c1_oracle.npz); the full array is written to the
path given as the second argument when present.

Run: OMP_NUM_THREADS=8 python tests/golden/gen_c1_oracle.py [nt] [full.npy]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "3d-ose_amd"), os.path.join(ROOT, "oracle")]

from ose_amd.host import Simulation  # noqa: E402
import pyoracle  # noqa: E402


def main():
    nt = int(sys.argv[1]) if len(sys.argv) > 1 else 150
    sim = Simulation(f"preset=c1 nt={nt}").build()
    m = sim.model(0)
    d = m["dims"]
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = d
    sh = sim.shot(0)
    src = (sh["src"][2] * nyb + sh["src"][1]) * nxb + sh["src"][0]
    om = pyoracle.Model(m["v"].astype(np.float64), 10.0, 10.0, 10.0, 1e-3, m["R"], m["coef"],
                        m["gx"].astype(np.float64), m["gy"].astype(np.float64), m["gz"].astype(np.float64))
    s = sh["s"].astype(np.float64)
    h = sh["h"].astype(np.float64)
    t0 = time.time()
    tr, P, Q = om.run(nt, src, s, h, bu + m["rec_z"], bl, bf, nx, ny)
    print(f"oracle {nt} steps on {nxb}x{nyb}x{nzb}, K={m['K']}: {time.time() - t0:.0f} s")
    iy, ix = np.meshgrid(np.arange(0, ny, 8), np.arange(0, nx, 8), indexing="ij")
    idx = (iy * nx + ix).ravel()
    np.savez(os.path.join(HERE, "c1_oracle.npz"), nt=nt, idx=idx, traces=tr[idx].astype(np.float32),
             K=m["K"], norm_full=np.linalg.norm(tr))
    if len(sys.argv) > 2:
        np.save(sys.argv[2], tr.astype(np.float32))
    sim.close()


if __name__ == "__main__":
    main()
