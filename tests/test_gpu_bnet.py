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


def test_bnet_forward_tiled_path_matches_fp64_reference():
    """A grid with an 8-cell pad on every side and a 64-cell x extent: the
    conv's LDS-staged tiles (full and partial, in all three bands) run next
    to the gathered edge tiles."""
    import ose_amd
    rng = np.random.default_rng(5)
    shape = (40, 44, 64)                 # nzb, nyb, nxb
    origin = (8, 8, 8)
    interior = (48, 28, 24)
    geo = B.Geometry(shape, origin, interior)
    v = rng.uniform(1500.0, 4000.0, shape).astype(np.float32)
    u = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    p = (0.15 * rng.standard_normal(B.NPAR)).astype(np.float32)
    print("bnet forward 64x44x40", flush=True)
    net = ose_amd.BoundaryNet(v, origin, interior)
    net.set_params(p)
    got = net.forward(u)
    ref, _ = B.forward(geo, p, u.astype(np.complex128), v.astype(np.float64), float(v.max()))
    net.close()
    assert np.all(got[~geo.pad] == 0)
    assert rel(got, ref) < 2e-2
    # every pad cell, not only the norm: no cell may be far off
    pad = geo.pad
    assert np.abs(got[pad] - ref[pad]).max() < 0.1 * np.abs(ref[pad]).max()


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


def _learned_case(taper="ref"):
    return Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=9), taper=taper)


def _run(case, params, nt=60, guard=2.0, replace_taper=True, chunks=1):
    import ose_amd
    s = ricker(nt, 30.0, case.dt, delay=10)
    h = hilbert_unnorm(s)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl + 18, case.bf + 10, case.bu + 2), s, h)
        if params is not None:
            c.set_boundary_net(m, params, guard, replace_taper=replace_taper)
        for k in range(chunks):
            c.run(nt // chunks)
        tr = c.traces(sid)
        re, im = c.wavefield(sid)
        st = c.boundary_status(m) if params is not None else None
    return tr, re + 1j * im, st


def test_zero_last_layer_is_identity_gain():
    """G = 1: on top of the taper the run is the plain taper; replacing the
    taper it is the untapered run."""
    print("learned pad with zero last layer", flush=True)
    p = (0.1 * np.random.default_rng(4).standard_normal(B.NPAR)).astype(np.float32)
    p[-(B.NTAP * 256 + 16):] = 0.0
    case = _learned_case()
    tr0, u0, _ = _run(case, None)
    tr1, u1, st = _run(case, p, replace_taper=False)
    assert st["active"] and not st["tripped"]
    assert rel(tr1, tr0) < 1e-6 and rel(u1, u0) < 1e-6
    tr2, u2, _ = _run(_learned_case("none"), None)
    tr3, u3, st = _run(case, p, replace_taper=True)
    assert st["active"] and not st["tripped"]
    assert rel(tr3, tr2) < 1e-6 and rel(u3, u2) < 1e-6
    assert rel(u3, u0) > 1e-4            # the taper does act on this case


def test_learned_pad_changes_only_through_the_pad():
    print("learned pad nonzero", flush=True)
    case = _learned_case()
    p = (0.05 * np.random.default_rng(5).standard_normal(B.NPAR)).astype(np.float32)
    tr0, u0, _ = _run(case, None)
    tr1, u1, st = _run(case, p, guard=1e6)
    assert st["active"] and not st["tripped"]
    assert np.all(np.isfinite(u1))
    assert rel(u1, u0) > 1e-6          # the correction acts


@pytest.mark.parametrize("replace_taper,chunks", [(False, 1), (True, 1), (True, 3)])
def test_guard_falls_back_to_taper(replace_taper, chunks):
    """A finite network whose output overflows (non-finite gain) trips the
    guard on its first step; from then on the run is the plain taper, also
    within one ose_run (graph replays) when the network replaced the taper.
    In that mode the tripping step itself ran untapered, so the reference is
    one untapered step followed by the plain taper."""
    print(f"learned pad guard replace={replace_taper} chunks={chunks}", flush=True)
    case = _learned_case()
    p = (0.1 * np.random.default_rng(6).standard_normal(B.NPAR)).astype(np.float32)
    w = B.NTAP * 256 + 16
    p[w:w + B.NTAP * 256] = 1e30        # W2
    p[2 * w:2 * w + B.NTAP * 256] = 1e30  # W3
    tr0, u0, _ = _run_switched(case) if replace_taper else _run(case, None)
    tr1, u1, st = _run(case, p, guard=2.0, replace_taper=replace_taper, chunks=chunks)
    assert st["tripped"] and not st["active"]
    assert rel(tr1, tr0) < 1e-6 and rel(u1, u0) < 1e-6


def _run_switched(case, nt=60):
    """One untapered step, then the plain taper (the run a guard trip on the
    first step of a taper-replacing network produces)."""
    import ose_amd
    s = ricker(nt, 30.0, case.dt, delay=10)
    h = hilbert_unnorm(s)
    src = (case.bl + 18, case.bf + 10, case.bu + 2)
    with ose_amd.Context(0) as c0:
        m0 = _learned_case("none").add_to(c0)
        s0 = c0.add_shot(m0, src, s, h)
        c0.run(1)
        re, im = c0.wavefield(s0)
        tr_first = c0.traces(s0)[:, :1].copy()
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, src, s, h)
        c.set_wavefield(sid, re, im)
        c.step = 1
        c.run(nt - 1)
        tr = c.traces(sid)
        tr[:, :1] = tr_first
        r2, i2 = c.wavefield(sid)
    return tr, r2 + 1j * i2, None


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
