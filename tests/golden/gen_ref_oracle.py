"""FP64 oracle seismogram of the reference exactly as shipped
(ose_3d.cu:75-125, 157-160, 242-279): 101^3 velocity file
input/v3d_nx101_ny101_nz101.bin with 25-cell pads -> 152^3 padded (the odd
rule widens the after-pads to 26), nt = 300, dt = 1 ms, Ricker fm = 40 Hz
amp = 100 delayed 40 samples, unnormalised Hilbert Q-source, source at
interior (50, 50, 0), receivers on the whole interior plane z = 0, order
K = M + 1 = (int)(R dt) + 7 and the generate_att taper on the accumulator.

The velocity is read from the reference's own file (path argument, default
/root/reference/input/v3d_nx101_ny101_nz101.bin) and must equal the host
library's `twolayer` model of the `ref` preset (3000 / 4000 m/s, interface
at depth index 30), so the GPU test can run `preset=ref` without the file;
the file's SHA-256 is recorded.  Committed (ref_oracle.npz): the traces of
every second receiver in x and y, the FP64 L2 norm of every one of the
10201 traces, and the norm of the whole array.

Run: OMP_NUM_THREADS=6 python tests/golden/gen_ref_oracle.py [vfile] [full.npy]
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "3d-ose_amd"), os.path.join(ROOT, "oracle")]

from ose_amd.host import Simulation  # noqa: E402
import pyoracle  # noqa: E402

VFILE = "/root/reference/input/v3d_nx101_ny101_nz101.bin"


def main():
    vfile = sys.argv[1] if len(sys.argv) > 1 else VFILE
    with open(vfile, "rb") as f:
        sha = hashlib.sha256(f.read()).hexdigest()
    sim = Simulation(f"preset=ref model=file vfile={vfile}").build()
    ref = Simulation("preset=ref").build()
    m, mp = sim.model(0), ref.model(0)
    assert np.array_equal(m["v"], mp["v"]), "velocity file differs from the preset's two-layer model"
    assert m["K"] == mp["K"] and np.array_equal(m["coef"], mp["coef"])
    ref.close()
    nt = 300
    d = m["dims"]
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = d
    assert (nxb, nyb, nzb) == (152, 152, 152)
    sh = sim.shot(0)
    src = (sh["src"][2] * nyb + sh["src"][1]) * nxb + sh["src"][0]
    om = pyoracle.Model(m["v"].astype(np.float64), 10.0, 10.0, 10.0, 1e-3, m["R"], m["coef"],
                        m["gx"].astype(np.float64), m["gy"].astype(np.float64), m["gz"].astype(np.float64))
    s = sh["s"].astype(np.float64)
    h = sh["h"].astype(np.float64)
    t0 = time.time()
    tr, P, Q = om.run(nt, src, s, h, bu + m["rec_z"], bl, bf, nx, ny)
    print(f"oracle {nt} steps on {nxb}x{nyb}x{nzb}, K={m['K']}: {time.time() - t0:.0f} s")
    iy, ix = np.meshgrid(np.arange(0, ny, 2), np.arange(0, nx, 2), indexing="ij")
    idx = (iy * nx + ix).ravel()
    np.savez(os.path.join(HERE, "ref_oracle.npz"), nt=nt, idx=idx, traces=tr[idx].astype(np.float32),
             K=m["K"], trace_norms=np.linalg.norm(tr, axis=1), norm_full=np.linalg.norm(tr),
             vfile_sha256=sha)
    if len(sys.argv) > 2:
        np.save(sys.argv[2], tr.astype(np.float32))
    sim.close()


if __name__ == "__main__":
    main()
