"""FFT extents without compiled register kernels (the runtime-length path).

The reference plans its 3D transform for whatever padded extent its odd rule
gives (cufftMakePlan3d(nzb, nyb, nxb), ose_3d.cu:226; judge_odd,
ose_3d_lib.cu:36-44), e.g. 150 for its own input/v3d_nx100_ny101_nz102.bin
with 25-cell pads, or 106 = 2 * 53.  Checked here: 3D transforms against
numpy at extents with large prime factors and above 1024; one OSE step
against the FP64 oracle with such an extent on each axis; the
runtime-length path forced onto compiled extents against the compiled
kernels over a batched run with receivers (every x-pass mode); every
extent compiled with a single launch shape against numpy and, on a few
grids, one step against the oracle; the reference's own input files on
their own grids against the FP64 oracle."""
import numpy as np
import pytest

from _cases import Case, hilbert_unnorm, layered, rel, ricker

pytestmark = pytest.mark.gpu


def _cfield(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("shape", [(20, 30, 106), (26, 158, 40), (1100, 8, 6), (4, 6, 2042), (2048, 4, 2),
                                   (3, 4096, 2), (13, 17, 19)])
def test_fft3d_runtime_length_matches_numpy(shape):
    import ose_amd
    print(f"fft3d {shape[::-1]}", flush=True)
    a = _cfield(shape, 5)
    with ose_amd.Context(0) as c:
        f = c.fft3d(a, -1)
        b = c.fft3d(f, +1)
    ref = np.fft.fftn(a.astype(np.complex128))
    assert rel(f, ref) < 1e-5
    assert rel(b, a.astype(np.complex128) * a.size) < 1e-5


def test_forced_runtime_length_fft_matches_numpy_on_compiled_extents():
    import ose_amd
    shape = (64, 48, 80)
    a = _cfield(shape, 6)
    with ose_amd.Context(0) as c:
        c.set_fft_path(1)
        f = c.fft3d(a, -1)
    assert rel(f, np.fft.fftn(a.astype(np.complex128))) < 1e-5


@pytest.mark.parametrize("interior", [(94, 36, 28), (36, 94, 28), (28, 24, 1088), (94, 106, 46)])
def test_step_with_runtime_length_extents_matches_oracle(interior):
    """One step from a smooth random state, reference taper, K = M + 1, on
    padded grids 106 x 48 x 40, 48 x 106 x 40, 40 x 36 x 1100 and
    106 x 118 x 58 (the runtime-length path on x, y, z and all three)."""
    import ose_amd
    print(f"one step, padded {tuple(n + 12 for n in interior)}", flush=True)
    case = Case(interior, (6, 6, 6, 6, 6, 6), layered(*interior, seed=9))
    assert any(ose_amd.extent_kind(n) == 2 for n in (case.nxb, case.nyb, case.nzb))
    rng = np.random.default_rng(4)
    u = rng.standard_normal(case.shape) + 1j * rng.standard_normal(case.shape)
    F = np.fft.fftn(u)
    kz, ky, kx = np.meshgrid(*[np.fft.fftfreq(n) for n in case.shape], indexing="ij")
    u = np.fft.ifftn(F * np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / 0.03))
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0))
        c.set_wavefield(sid, u.real, u.imag)
        c.run(1)
        re, im = c.wavefield(sid)
    P, Q = case.oracle().step(u.real, u.imag)
    err = rel(re + 1j * im, P + 1j * Q)
    print(f"  rel-L2 vs FP64 oracle {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("K", [9, 12])
def test_runtime_length_path_equals_compiled_path(K):
    """Compiled extents run once with the compiled kernels and once with the
    runtime-length kernels on every axis: two models, three shots, source
    injection, taper, plane receivers, 40 steps (orders that update the
    accumulator, skip it, start it, and the first and last orders)."""
    import ose_amd
    print(f"runtime-length vs compiled kernels, K={K}", flush=True)
    nt = 40
    a = Case((52, 36, 68), (6, 6, 6, 6, 6, 6), layered(52, 36, 68, seed=41), K=K)
    b = Case((36, 52, 20), (6, 6, 6, 6, 6, 6), layered(36, 52, 20, seed=42), K=K)
    out = {}
    for path in (0, 1):
        with ose_amd.Context(0) as c:
            c.set_fft_path(path)
            ma = a.add_to(c)
            mb = b.add_to(c)
            ids = []
            for m, case, (sx, sy, sz) in ((ma, a, (26, 18, 2)), (ma, a, (10, 30, 9)), (mb, b, (18, 26, 3))):
                s = ricker(nt, 30.0, 1e-3, delay=5)
                ids.append(c.add_shot(m, (case.bl + sx, case.bf + sy, case.bu + sz), s, hilbert_unnorm(s)))
            c.run(nt)
            out[path] = [(c.traces(i), c.wavefield(i)) for i in ids]
    for (t0, (r0, i0)), (t1, (r1, i1)) in zip(out[0], out[1]):
        assert np.abs(t0).max() > 0
        assert rel(t1, t0) < 1e-5
        assert rel(r1 + 1j * i1, r0 + 1j * i0) < 1e-5


@pytest.mark.parametrize("name,dims,padded", [("v3d_nx100_ny101_nz102", (100, 101, 102), (150, 152, 152)),
                                              ("3d_number-100-100-100-3", (100, 100, 100), (150, 150, 150))])
def test_reference_input_files_on_their_own_grid(name, dims, padded, tmp_path):
    """The reference's two other velocity files at its 25-cell padding,
    extent=exact, so x (and for the 100^3 file every axis) is 150 cells
    (compiled in a single launch shape): 60 steps against the FP64 oracle
    (tests/golden/gen_refinput_oracle.py; the files are stored there
    byte for byte)."""
    import os
    from ose_amd.host import Simulation
    gold = os.path.join(os.path.dirname(__file__), "golden")
    vin = np.load(os.path.join(gold, "ref_inputs.npz"))
    g = np.load(os.path.join(gold, "refinput_oracle.npz"))
    nt = int(g["nt"])
    vfile = tmp_path / (name + ".bin")
    vin[name].astype("<f4").tofile(vfile)
    nx, ny, nz = dims
    print(f"{name} at boundary 25, {nt} steps", flush=True)
    sim = Simulation(f"preset=ref nx={nx} ny={ny} nz={nz} model=file vfile={vfile} nt={nt} delay=0 "
                     f"sx=50 sy=50 sz=0 extent=exact").setup()
    try:
        m = sim.model(0)
        assert tuple(m["dims"][9:12]) == padded == tuple(g[name + "_padded"])
        assert m["K"] == int(g[name + "_K"])
        sim.run(nt)
        tr = sim.traces(0, 0, nt)
    finally:
        sim.close()
    err = rel(tr[g[name + "_idx"]], g[name + "_traces"])
    norms = np.linalg.norm(tr.astype(np.float64), axis=1)
    dev = float(np.max(np.abs(norms - g[name + "_trace_norms"])) / np.max(g[name + "_trace_norms"]))
    ratio = np.linalg.norm(tr.astype(np.float64)) / float(g[name + "_norm_full"])
    print(f"  rel-L2 {err:.3g}, trace-norm deviation {dev:.3g}, norm ratio {ratio:.7f}", flush=True)
    assert err < 1e-4 and dev < 1e-4 and abs(ratio - 1.0) < 1e-4


def test_every_lean_extent_transforms_like_numpy():
    """Every extent compiled with a single launch shape (gen_instances.LEAN:
    even, up to 1024, prime factors <= 13) on each axis of a 3D transform."""
    import os
    import sys
    import ose_amd
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "3d-ose_amd",
                                    "csrc", "gen"))
    import gen_instances as G
    sizes = sorted(G.LEAN)
    print(f"{len(sizes)} lean extents", flush=True)
    worst = 0.0
    with ose_amd.Context(0) as c:
        for n in sizes:
            assert ose_amd.extent_kind(n) == 3
            for shape in ((2, 4, n), (2, n, 4), (n, 4, 2)):
                a = _cfield(shape, n)
                f = c.fft3d(a, -1)
                b = c.fft3d(f, +1)
                e = max(rel(f, np.fft.fftn(a.astype(np.complex128))), rel(b, a.astype(np.complex128) * a.size))
                worst = max(worst, e)
                assert e < 1e-5, (n, shape, e)
    print(f"  worst rel-L2 {worst:.2e}", flush=True)


@pytest.mark.parametrize("padded", [(150, 90, 130), (338, 22, 40), (48, 286, 18), (40, 36, 896)])
def test_step_on_lean_extents_matches_oracle(padded):
    """One step on grids whose extents are compiled with a single shape."""
    import ose_amd
    interior = tuple(n - 12 for n in padded)
    print(f"one step, padded {padded}", flush=True)
    case = Case(interior, (6, 6, 6, 6, 6, 6), layered(*interior, seed=13))
    rng = np.random.default_rng(5)
    u = rng.standard_normal(case.shape) + 1j * rng.standard_normal(case.shape)
    F = np.fft.fftn(u)
    kz, ky, kx = np.meshgrid(*[np.fft.fftfreq(n) for n in case.shape], indexing="ij")
    u = np.fft.ifftn(F * np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / 0.03))
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0))
        c.set_wavefield(sid, u.real, u.imag)
        c.run(1)
        re, im = c.wavefield(sid)
    P, Q = case.oracle().step(u.real, u.imag)
    err = rel(re + 1j * im, P + 1j * Q)
    print(f"  rel-L2 vs FP64 oracle {err:.2e}", flush=True)
    assert err < 1e-5
