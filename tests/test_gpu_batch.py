"""Batching on the GPU: several shots on one model and models of different
extents stepped together in one graph-captured step, against one FP64
oracle run per shot; the autotuner's launch-shape choice does not change
results; the recursion is right for odd and even series orders."""
import numpy as np
import pytest

from _cases import Case, hilbert_unnorm, layered, rel, ricker

pytestmark = pytest.mark.gpu


def _smooth_field(shape, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    F = np.fft.fftn(f)
    kz, ky, kx = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    F *= np.exp(-((kx ** 2 + ky ** 2 + kz ** 2) / 0.02))
    return np.fft.ifftn(F)


@pytest.mark.parametrize("K", [2, 3, 4, 5, 6, 7, 8, 9, 12, 13])
def test_one_step_odd_and_even_order(K):
    """The last order writes the new state over W_{K-2} (K even) or into
    the other physical buffer (K odd); K = 2..9 also cover every shape of
    the deferred accumulator schedule (which orders update it, from the
    first order on for K = 2 and 5, and where it starts)."""
    import ose_amd
    print(f"one step K={K}", flush=True)
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=7), K=K)
    u = _smooth_field(case.shape, 8)
    P, Q = case.oracle().step(u.real, u.imag)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0))
        c.set_wavefield(sid, u.real, u.imag)
        c.run(2)   # the second step starts from the buffers the first left behind
        P, Q = case.oracle().step(P, Q)
        re, im = c.wavefield(sid)
    assert rel(re + 1j * im, P + 1j * Q) < 1e-5


def _shots(case, specs, nt):
    out = []
    for (sx, sy, sz, fm, delay) in specs:
        s = ricker(nt, fm, case.dt, delay=delay)
        out.append(((sx, sy, sz), s, hilbert_unnorm(s)))
    return out


@pytest.mark.parametrize("autotune", [False, True])
def test_batched_shots_and_models_match_oracle(autotune):
    import ose_amd
    print(f"batched shots/models autotune={autotune}", flush=True)
    # nt and source depths are chosen so the direct waves reach the surface
    # receivers well inside the run: a trace set that holds only the
    # operator's early far tail is at the FP32 noise floor of the field
    nt = 100
    a = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=11))
    b = Case((52, 20, 20), (6, 6, 6, 6, 6, 6), layered(52, 20, 20, seed=12, v0=1800.0, v1=3000.0))
    shots_a = _shots(a, [(18, 10, 2, 30.0, 10), (5, 15, 8, 25.0, 14), (30, 3, 12, 35.0, 8)], nt)
    shots_b = _shots(b, [(14, 10, 3, 30.0, 10), (3, 17, 10, 28.0, 12)], nt)
    with ose_amd.Context(0) as c:
        c.set_autotune(autotune)
        ma = a.add_to(c)
        mb = b.add_to(c)
        ids = []
        # interleave the shots of the two models
        for k in range(3):
            (sx, sy, sz), s, h = shots_a[k]
            ids.append((a, shots_a[k], c.add_shot(ma, (a.bl + sx, a.bf + sy, a.bu + sz), s, h)))
            if k < 2:
                (sx, sy, sz), s, h = shots_b[k]
                ids.append((b, shots_b[k], c.add_shot(mb, (b.bl + sx, b.bf + sy, b.bu + sz), s, h)))
        c.run(nt // 2)
        c.run(nt - nt // 2)
        got = [(case, spec, c.traces(sid), c.wavefield(sid)) for case, spec, sid in ids]
    errs = []
    for case, ((sx, sy, sz), s, h), tr, (re, im) in got:
        tr_ref, P, Q = case.oracle().run(nt, case.src_index(sx, sy, sz), s, h, case.bu, case.bl, case.bf,
                                         case.nx, case.ny)
        errs.append((case.shape, (sx, sy, sz), rel(tr, tr_ref), rel(re + 1j * im, P + 1j * Q)))
    for e in errs:
        print("shape %s src %s: traces %.3g field %.3g" % e, flush=True)
    assert max(e[2] for e in errs) < 1e-4
    assert max(e[3] for e in errs) < 1e-4


@pytest.mark.parametrize("padded", [(48, 336, 336), (336, 48, 48), (600, 32, 24)])
def test_every_launch_shape_gives_the_same_step(padded):
    """Every compiled launch shape (T columns / P lines per workgroup, whole
    or split LDS exchange) runs the same arithmetic: one step from the same
    state agrees across shapes to rounding, and the default shape agrees
    with the FP64 oracle on the smaller grid."""
    import ose_amd
    nxb, nyb, nzb = padded
    print(f"launch shapes {padded}", flush=True)
    case = Case((nxb - 12, nyb - 12, nzb - 12), (6, 6, 6, 6, 6, 6), layered(nxb - 12, nyb - 12, nzb - 12, seed=13))
    u = _smooth_field(case.shape, 14).astype(np.complex64)
    results = {}
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0))
        strided, xs = c.launch_shapes(m)
        base = c.tuning(m)
        combos = [(t, t, t, base["P_x"]) for t in strided] + [(base["T_fwd_y"], base["T_z"], base["T_inv_y"], p)
                                                               for p in xs]
        for combo in combos:
            c.set_tuning(m, *combo)
            assert tuple(c.tuning(m).values()) == combo
            c.set_wavefield(sid, u.real, u.imag)
            c.step = 0
            c.run(1)
            re, im = c.wavefield(sid)
            results[combo] = re + 1j * im
    assert any(x < 0 for x in strided + xs)        # a split variant is among them
    ref = results[combos[0]]
    for combo, r in results.items():
        assert rel(r, ref) < 1e-6, combo
    if nxb * nyb * nzb < 1_000_000:
        P, Q = case.oracle().step(u.real.astype(np.float64), u.imag.astype(np.float64))
        assert rel(ref, P + 1j * Q) < 1e-5


@pytest.mark.parametrize("padded,nshot", [((48, 40, 336), 2), ((40, 22, 336), 3)])
def test_chunked_z_pass_equals_one_tile_per_workgroup(padded, nshot):
    """The chunked z-symbol pass (several consecutive tiles per workgroup,
    each tile's successor prefetched; codes 100 * tiles + 16) computes what
    the one-tile-per-workgroup pass computes, on every shot: chunks that
    cross rows and shots, a last chunk shorter than the rest (3 x 22 x 3 =
    198 tiles in chunks of 4) and a last column block with threads past the
    grid (40 = 2 x 16 + 8 columns, whose duplicate stores write the owner's
    values); the one-tile result agrees with the FP64 oracle."""
    import ose_amd
    nxb, nyb, nzb = padded
    print(f"chunked z pass {padded} x {nshot}", flush=True)
    case = Case((nxb - 12, nyb - 12, nzb - 12), (6, 6, 6, 6, 6, 6), layered(nxb - 12, nyb - 12, nzb - 12, seed=17))
    us = [_smooth_field(case.shape, 30 + k).astype(np.complex64) for k in range(nshot)]
    results = {}
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sids = [c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0)) for _ in range(nshot)]
        c.run(0)
        base = c.tuning(m)
        for zcode in (-16, 216, 316, 416, 616):
            c.set_tuning(m, base["T_fwd_y"], zcode, base["T_inv_y"], base["P_x"])
            assert c.tuning(m)["T_z"] == zcode
            for sid, u in zip(sids, us):
                c.set_wavefield(sid, u.real, u.imag)
            c.step = 0
            c.run(1)
            results[zcode] = [complex_field(c.wavefield(sid)) for sid in sids]
    for zcode, fields in results.items():
        for k in range(nshot):
            assert rel(fields[k], results[-16][k]) < 1e-6, (zcode, k)
    for k in range(nshot):
        P, Q = case.oracle().step(us[k].real.astype(np.float64), us[k].imag.astype(np.float64))
        assert rel(results[-16][k], P + 1j * Q) < 1e-5, k


def complex_field(reim):
    re, im = reim
    return re + 1j * im


@pytest.mark.parametrize("nshot", [1, 4])
def test_pattern_probe_leaves_the_state_intact(nshot):
    """ose_pattern_probe (the strided passes' reads and writes without the
    transforms) reports a rate for both axes, and a run that probes between
    its steps ends where the plain run ends, bit for bit: in place (one
    336^3 shot) and with the second spectral field (four shots)."""
    import ctypes
    import ose_amd
    print(f"pattern probe between steps, {nshot} shot(s) of 336^3", flush=True)
    nt = 4
    case = Case((324, 324, 324), (6, 6, 6, 6, 6, 6), layered(324, 324, 324, seed=41))
    shots = _shots(case, [(162, 162, 2, 30.0, 2), (40, 200, 10, 25.0, 3), (300, 20, 30, 35.0, 2),
                          (100, 100, 100, 28.0, 1)], nt)[:nshot]
    out = {}
    for probe in (False, True):
        with ose_amd.Context(0) as c:
            c.set_autotune(False)
            m = case.add_to(c)
            ids = [c.add_shot(m, (case.bl + sx, case.bf + sy, case.bu + sz), s, h) for (sx, sy, sz), s, h in shots]
            for _ in range(nt):
                c.run(1)
                if probe:
                    for axis in (1, 2):
                        g = ctypes.c_double()
                        assert c._L.ose_pattern_probe(c._h, m, axis, 1, ctypes.byref(g)) == 0
                        assert g.value > 0
            out[probe] = [(c.traces(i), c.wavefield(i)) for i in ids]
    for (t0, (r0, i0)), (t1, (r1, i1)) in zip(out[False], out[True]):
        assert np.abs(r0).max() > 0
        assert np.array_equal(t0, t1)
        assert np.array_equal(r0, r1) and np.array_equal(i0, i1)


@pytest.mark.parametrize("use_graph", [True, False])
def test_two_stream_lanes_equal_one_stream(use_graph):
    """The step's two concurrent lanes (shots dealt in halves per model,
    single-shot models whole) compute exactly what one stream computes: the
    same kernels on the same shots, bit for bit, with receivers, source
    injection and models of different extents."""
    import ose_amd
    print(f"two-stream lanes, graph={use_graph}", flush=True)
    nt = 24
    a = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=21))
    b = Case((52, 20, 20), (6, 6, 6, 6, 6, 6), layered(52, 20, 20, seed=22, v0=1800.0, v1=3000.0))
    shots_a = _shots(a, [(18, 10, 2, 30.0, 4), (5, 15, 8, 25.0, 6), (30, 3, 12, 35.0, 3)], nt)
    shots_b = _shots(b, [(14, 10, 3, 30.0, 4)], nt)
    out = {}
    for streams in (1, 2):
        with ose_amd.Context(0) as c:
            c.set_autotune(False)
            c.set_options(use_graph=use_graph, profile=False)
            c.set_streams(streams)
            ma = a.add_to(c)
            mb = b.add_to(c)
            ids = []
            for (sx, sy, sz), s, h in shots_a:
                ids.append(c.add_shot(ma, (a.bl + sx, a.bf + sy, a.bu + sz), s, h))
            for (sx, sy, sz), s, h in shots_b:
                ids.append(c.add_shot(mb, (b.bl + sx, b.bf + sy, b.bu + sz), s, h))
            c.run(nt // 3)
            c.run(nt - nt // 3)
            out[streams] = [(c.traces(i), c.wavefield(i)) for i in ids]
    for (t1, (r1, i1)), (t2, (r2, i2)) in zip(out[1], out[2]):
        assert np.abs(t1).max() > 0
        assert np.array_equal(t1, t2)
        assert np.array_equal(r1, r2) and np.array_equal(i1, i2)


def test_out_of_place_batch_equals_in_place_shot():
    """A batch of 1 GiB or more per field runs passes 1-3 out of place
    through a second spectral field; a single shot of the same model stays
    in place.  Same kernels and shapes: the shots agree bit for bit."""
    import ose_amd
    print("out-of-place batch (4 x 336^3) vs in-place shot", flush=True)
    nt = 6
    case = Case((324, 324, 324), (6, 6, 6, 6, 6, 6), layered(324, 324, 324, seed=31))
    assert 4 * case.nxb * case.nyb * case.nzb * 8 >= 1 << 30 > case.nxb * case.nyb * case.nzb * 8
    shots = _shots(case, [(162, 162, 2, 30.0, 2), (40, 200, 10, 25.0, 3), (300, 20, 30, 35.0, 2),
                          (100, 100, 100, 28.0, 1)], nt)
    out = {}
    for n in (4, 1):
        with ose_amd.Context(0) as c:
            c.set_autotune(False)
            m = case.add_to(c)
            ids = [c.add_shot(m, (case.bl + sx, case.bf + sy, case.bu + sz), s, h)
                   for (sx, sy, sz), s, h in shots[:n]]
            c.run(nt)
            out[n] = (c.traces(ids[0]), c.wavefield(ids[0]))
    (t4, (r4, i4)), (t1, (r1, i1)) = out[4], out[1]
    assert np.abs(r1).max() > 0
    assert np.array_equal(t4, t1)
    assert np.array_equal(r4, r1) and np.array_equal(i4, i1)
