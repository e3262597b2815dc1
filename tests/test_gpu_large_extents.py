"""The long-extent kernels (336 ... 1024 cells, the padded sizes of the
benchmark configurations) against the FP64 oracle, one compiled launch shape
at a time: a grid N cells long on one axis and 16 on the others runs one
step per variant of the pass that works along that axis (x pass for x,
forward and inverse strided passes for y, the z-symbol pass for z), each
from the same smooth state, and every result matches the oracle's step."""
import os
import sys

import numpy as np
import pytest

from _cases import Case, layered, rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "3d-ose_amd", "csrc", "gen"))
import gen_instances as G  # noqa: E402


def _smooth_field(shape, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    F = np.fft.fftn(f)
    kz, ky, kx = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    F *= np.exp(-((kx ** 2 + ky ** 2 + kz ** 2) / 0.02))
    return np.fft.ifftn(F)


def _codes(n):
    n1, n2 = G.SIZES[n]
    strided = [-t if split else t for t, split in G.strided_variants(n1, n2)]
    xs = [-(p + 1000 * twb) if split else p + 1000 * twb for p, split, twb in G.x_variants(n1, n2)]
    return strided, xs


@pytest.mark.parametrize("n", [336, 600, 784, 864, 1024])
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_every_long_extent_variant_matches_oracle(n, axis):
    import ose_amd
    # padded n on `axis`, 16 elsewhere: interior n - 8 with 4-cell pads
    dims = {a: (n - 8 if a == axis else 8) for a in "xyz"}
    pads = tuple(4 for _ in range(6))
    interior = (dims["x"], dims["y"], dims["z"])
    case = Case(interior, pads, layered(*interior, seed=n % 97))
    assert case.shape[{"z": 0, "y": 1, "x": 2}[axis]] == n
    u = _smooth_field(case.shape, 11)
    P, Q = case.oracle().step(u.real, u.imag)
    want = P + 1j * Q
    strided, xs = _codes(n)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0))
        c.set_wavefield(sid, u.real, u.imag)
        c.run(1)
        base = c.tuning(m)
        base = [base["T_fwd_y"], base["T_z"], base["T_inv_y"], base["P_x"]]
        variants = {"x": [(3, v) for v in xs], "y": [(0, v) for v in strided] + [(2, v) for v in strided],
                    "z": [(1, v) for v in strided]}[axis]
        print(f"{axis} = {n}: {len(variants)} launch shapes, tuned {base}", flush=True)
        assert len(variants) >= 3
        for slot, code in variants:
            shape = list(base)
            shape[slot] = code
            c.set_tuning(m, *shape)
            c.step = 0
            c.set_wavefield(sid, u.real, u.imag)
            c.run(1)
            re, im = c.wavefield(sid)
            err = rel(re + 1j * im, want)
            assert err < 1e-5, (axis, n, shape, err)
