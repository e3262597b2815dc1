"""GPU runs of whole configurations against committed references:
the 100^3 two-layer configuration (1000 steps) against the FP64 oracle's
seismogram, and the trained thin pad against the plain thin taper on a
held-out model (reflections and seismogram error relative to a wide
sponge)."""
import os

import numpy as np
import pytest

from _cases import rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_c0_seismogram_matches_fp64_oracle():
    """100^3 two-layer, 20-cell Cerjan taper, receivers at depth 5, 1000
    steps: FP32 GPU traces against the FP64 oracle (every 5th receiver in x
    and y committed by tests/golden/gen_c0_oracle.py)."""
    from ose_amd.host import Simulation
    g = np.load(os.path.join(GOLD, "c0_oracle.npz"))
    nt = int(g["nt"])
    print(f"c0 {nt} steps", flush=True)
    sim = Simulation(f"preset=c0 nt={nt}").setup()
    try:
        assert sim.model(0)["K"] == int(g["K"])
        sim.run(nt)
        tr = sim.traces(0, 0, nt)
    finally:
        sim.close()
    sub = tr[g["idx"]]
    err = rel(sub, g["traces"])
    ratio = np.linalg.norm(tr.astype(np.float64)) / float(g["norm_full"])   # FP64 sum over 1e7 samples
    print(f"c0 rel-L2 (every 5th receiver) {err:.3g}; full-array norm ratio {ratio:.7f}", flush=True)
    assert err < 1e-4
    assert abs(ratio - 1.0) < 1e-4


def test_trained_thin_pad_beats_plain_taper():
    """Held-out seeded model (64^3 interior): the committed 8-cell learned
    pad against the 8-cell reference taper, both measured against a 64-cell
    sponge run; the network must cut the reflected interior energy at least
    in half and lower the seismogram error."""
    from ose_amd import boundary
    p = np.load(boundary.WEIGHTS)
    print("trained pad evaluation", flush=True)
    row = boundary.evaluate(p, [2003])[0]
    print(row, flush=True)
    assert row["learned_status"]["active"] and not row["learned_status"]["tripped"]
    assert row["learned_reflected"] < 0.5 * row["plain_reflected"]
    assert row["learned_rel_l2"] < row["plain_rel_l2"]
