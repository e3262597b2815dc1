"""GPU parity of the learned thin-pad network: BF16 MFMA forward and
gradients against the FP64 numpy reference, training, and its use inside the
time loop (a zero last layer is the plain taper; the guard falls back)."""
import numpy as np
import pytest

import bnet_ref as B
from _cases import Case, hilbert_unnorm, layered, rel, ricker

pytestmark = pytest.mark.gpu


def _geometry_case(seed):
    rng = np.random.default_rng(seed)
    shape = (32, 24, 16)                 # nzb, nyb, nxb
    origin = (4, 5, 6)                   # bl, bf, bu (x, y, z)
    interior = (9, 12, 20)               # nx, ny, nz
    geo = B.Geometry(shape, origin, interior)
    v = rng.uniform(1500.0, 4000.0, shape).astype(np.float32)
    u = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    t = (u + 0.3 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)
    p = (0.15 * rng.standard_normal(B.NPAR)).astype(np.float32)
    return geo, shape, origin, interior, v, u, t, p


def test_bnet_forward_matches_fp64_reference():
    import ose_amd
    geo, shape, origin, interior, v, u, t, p = _geometry_case(0)
    print("bnet forward 16x24x32", flush=True)
    net = ose_amd.BoundaryNet(v, origin, interior)
    net.set_params(p)
    got = net.forward(u)
    ref, _ = B.forward(geo, p, u.astype(np.complex128), v.astype(np.float64), float(v.max()))
    net.close()
    assert np.all(got[~geo.pad] == 0)
    # BF16 operands, FP32 accumulation
    assert rel(got, ref) < 2e-2


def test_bnet_gradient_matches_fp64_reference():
    import ose_amd
    geo, shape, origin, interior, v, u, t, p = _geometry_case(1)
    print("bnet gradient 16x24x32", flush=True)
    net = ose_amd.BoundaryNet(v, origin, interior)
    net.set_params(p)
    loss, g = net.gradient(u, t)
    net.close()
    Lref, gref = B.loss_and_grad(geo, p, u.astype(np.complex128), v.astype(np.float64), float(v.max()),
                                 t.astype(np.complex128))
    assert abs(loss - Lref) < 2e-2 * Lref
    # per layer: weights and biases
    o = 0
    for _ in range(3):
        for n in (B.NTAP * 256, 16):
            seg, ref = g[o:o + n], gref[o:o + n]
            if np.linalg.norm(ref) > 0:
                assert rel(seg, ref) < 5e-2, (o, n, rel(seg, ref))
            o += n


def test_bnet_training_reduces_loss():
    import ose_amd
    geo, shape, origin, interior, v, u, t, p = _geometry_case(2)
    print("bnet training", flush=True)
    net = ose_amd.BoundaryNet(v, origin, interior)
    rng = np.random.default_rng(3)
    for k in range(4):
        uk = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        tk = uk.copy()
        tk[geo.pad] *= 0.5          # target: damp the pad by half
        net.add_sample(uk, tk)
    losses = net.train(300, lr=3e-3, seed=1)
    net.close()
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.5 * losses[:5].mean()


def _learned_case():
    return Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=9))


def _run(case, params, nt=60, guard=2.0):
    import ose_amd
    s = ricker(nt, 30.0, case.dt, delay=10)
    h = hilbert_unnorm(s)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl + 18, case.bf + 10, case.bu + 2), s, h)
        if params is not None:
            c.set_boundary_net(m, params, guard)
        c.run(nt)
        tr = c.traces(sid)
        re, im = c.wavefield(sid)
        st = c.boundary_status(m) if params is not None else None
    return tr, re + 1j * im, st


def test_zero_last_layer_equals_plain_taper():
    print("learned pad with zero last layer", flush=True)
    case = _learned_case()
    p = (0.1 * np.random.default_rng(4).standard_normal(B.NPAR)).astype(np.float32)
    p[-(B.NTAP * 256 + 16):] = 0.0
    tr0, u0, _ = _run(case, None)
    tr1, u1, st = _run(case, p)
    assert st["active"] and not st["tripped"]
    assert rel(tr1, tr0) < 1e-6 and rel(u1, u0) < 1e-6


def test_learned_pad_changes_only_through_the_pad():
    print("learned pad nonzero", flush=True)
    case = _learned_case()
    p = (0.05 * np.random.default_rng(5).standard_normal(B.NPAR)).astype(np.float32)
    tr0, u0, _ = _run(case, None)
    tr1, u1, st = _run(case, p, guard=1e6)
    assert st["active"] and not st["tripped"]
    assert np.all(np.isfinite(u1))
    assert rel(u1, u0) > 1e-6          # the correction acts


def test_guard_falls_back_to_taper():
    print("learned pad guard", flush=True)
    case = _learned_case()
    p = np.zeros(B.NPAR, dtype=np.float32)
    p[-16] = 1e4                        # b3[0]: a huge constant correction on the pad
    tr0, u0, _ = _run(case, None)
    tr1, u1, st = _run(case, p, guard=2.0)
    assert st["tripped"] and not st["active"]
    assert rel(tr1, tr0) < 1e-6 and rel(u1, u0) < 1e-6


def test_nonfinite_parameters_rejected():
    import ose_amd
    case = _learned_case()
    p = np.zeros(B.NPAR, dtype=np.float32)
    p[7] = np.nan
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        with pytest.raises(ose_amd.OseError) as e:
            c.set_boundary_net(m, p)
        assert e.value.code == 5
