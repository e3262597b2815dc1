"""FP64 oracle seismograms on the reference's other two velocity files at
its own padding, so the padded grids are the ones cufftMakePlan3d would be
planned for (ose_3d.cu:85-91, 226; judge_odd ose_3d_lib.cu:36-44):

  input/v3d_nx100_ny101_nz102.bin   100 x 101 x 102 -> 150 x 152 x 152
  input/3d_number-100-100-100-3.bin 100 x 100 x 100 -> 150^3

150 has no compiled register split, so these run on the runtime-length
transform path with extent=exact.  Parameters otherwise as the reference's
main() (ose_3d.cu:75-125): dt = 1 ms, dx = 10 m, Ricker fm = 40 Hz
amp = 100 (delay 0 here, so the pulse leaves within the short window),
unnormalised Hilbert Q-source, source at interior (50, 50, 0), receivers on
the interior plane z = 0, K = M + 1, generate_att taper over 25 cells.

The velocity files are data the reference ships; they are stored here as
float32 arrays in the file's own [x][y][z] order (ref_inputs.npz, with their
SHA-256), so the GPU test rewrites them byte for byte without the reference.
Committed results (refinput_oracle.npz): per file, the traces of every
second receiver in x and y, the norm of every trace and of the whole array.

Run: OMP_NUM_THREADS=6 python tests/golden/gen_refinput_oracle.py [nt]
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

INPUT = "/root/reference/input"
FILES = {"v3d_nx100_ny101_nz102": (100, 101, 102), "3d_number-100-100-100-3": (100, 100, 100)}


def params(name, dims, vfile, nt):
    nx, ny, nz = dims
    return (f"preset=ref nx={nx} ny={ny} nz={nz} model=file vfile={vfile} nt={nt} delay=0 "
            f"sx=50 sy=50 sz=0 extent=exact")


def main():
    nt = int(sys.argv[1]) if len(sys.argv) > 1 else 60
    store, out = {}, {"nt": nt}
    for name, dims in FILES.items():
        path = os.path.join(INPUT, name + ".bin")
        raw = open(path, "rb").read()
        store[name] = np.frombuffer(raw, dtype="<f4").reshape(dims)
        store[name + "_sha256"] = hashlib.sha256(raw).hexdigest()
        sim = Simulation(params(name, dims, path, nt)).build()
        m = sim.model(0)
        nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = m["dims"]
        sh = sim.shot(0)
        src = (sh["src"][2] * nyb + sh["src"][1]) * nxb + sh["src"][0]
        om = pyoracle.Model(m["v"].astype(np.float64), 10.0, 10.0, 10.0, 1e-3, m["R"], m["coef"],
                            m["gx"].astype(np.float64), m["gy"].astype(np.float64), m["gz"].astype(np.float64))
        t0 = time.time()
        tr, P, Q = om.run(nt, src, sh["s"].astype(np.float64), sh["h"].astype(np.float64), bu + m["rec_z"],
                          bl, bf, nx, ny)
        print(f"{name}: oracle {nt} steps on {nxb}x{nyb}x{nzb}, K={m['K']}: {time.time() - t0:.0f} s", flush=True)
        iy, ix = np.meshgrid(np.arange(0, ny, 2), np.arange(0, nx, 2), indexing="ij")
        idx = (iy * nx + ix).ravel()
        out[name + "_padded"] = np.array([nxb, nyb, nzb])
        out[name + "_K"] = m["K"]
        out[name + "_idx"] = idx
        out[name + "_traces"] = tr[idx].astype(np.float32)
        out[name + "_trace_norms"] = np.linalg.norm(tr, axis=1)
        out[name + "_norm_full"] = np.linalg.norm(tr)
        sim.close()
    np.savez_compressed(os.path.join(HERE, "ref_inputs.npz"), **store)
    np.savez(os.path.join(HERE, "refinput_oracle.npz"), **out)


if __name__ == "__main__":
    main()
