"""GPU parity: the HIP path (through the C-ABI) against the FP64 oracle and
the numpy fixtures, at sizes the oracle finishes in seconds."""
import os

import numpy as np
import pytest

from _cases import Case, hilbert_unnorm, layered, rel, ricker

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ose():
    import ose_amd
    return ose_amd


@pytest.mark.parametrize("shape", [(16, 24, 32), (32, 48, 64), (24, 16, 140), (16, 152, 24),
                                   (336, 16, 24), (24, 16, 864)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft3d_matches_numpy(ose, shape, sign):
    """Pass kernels as a plain 3D DFT (cufftXtExecDescriptorC2C, ose_3d_lib.cu:668)."""
    print(f"fft3d shape={shape} sign={sign}", flush=True)
    rng = np.random.default_rng(1)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ref = np.fft.fftn(a.astype(np.complex128)) if sign < 0 else np.fft.ifftn(a.astype(np.complex128)) * a.size
    with ose.Context(0) as c:
        got = c.fft3d(a, sign)
    assert rel(got, ref) < 2e-6


def test_apply_operator_matches_oracle(ose):
    """L u = (v/R) IFFT(|k| FFT u)/N (ose_3d_lib.cu:649-687)."""
    print("apply_operator 48x32x48", flush=True)
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=2))
    rng = np.random.default_rng(3)
    u = rng.standard_normal(case.shape) + 1j * rng.standard_normal(case.shape)
    with ose.Context(0) as c:
        m = case.add_to(c)
        got = c.apply_operator(m, u.astype(np.complex64))
    ref = case.oracle().apply_L(u)
    assert rel(got, ref) < 1e-5


def _smooth_field(shape, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    F = np.fft.fftn(f)
    kz, ky, kx = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    F *= np.exp(-((kx ** 2 + ky ** 2 + kz ** 2) / 0.02))
    return np.fft.ifftn(F)


@pytest.mark.parametrize("taper", ["ref", "none"])
def test_one_step_matches_oracle(ose, taper):
    """One wavefield_extro_ori2 step (ose_3d_lib.cu:639-874) from a set state."""
    print(f"one step taper={taper}", flush=True)
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=4), taper=taper)
    u = _smooth_field(case.shape, 5)
    P, Q = case.oracle().step(u.real, u.imag)
    with ose.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl, case.bf, case.bu), np.zeros(0), np.zeros(0))
        c.set_wavefield(sid, u.real, u.imag)
        c.run(1)
        re, im = c.wavefield(sid)
    assert rel(re + 1j * im, P + 1j * Q) < 1e-5


def test_tiny_run_matches_numpy_fixture(ose):
    """12 steps with source and receivers vs the committed numpy restatement."""
    print("tiny fixture run", flush=True)
    t = np.load(os.path.join(GOLD, "step_tiny.npz"))
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = (int(x) for x in t["dims"])
    dx, dy, dz, dt = t["d"]
    src = int(t["src"])
    sx, sy, sz = src % nxb, (src // nxb) % nyb, src // (nxb * nyb)
    with ose.Context(0) as c:
        m = c.add_model(t["v"], interior=(nx, ny, nz), origin=(bl, bf, bu), d=(dx, dy, dz), dt=dt,
                        R=float(t["R"]), coef=t["coef"], taper=(t["gx"], t["gy"], t["gz"]),
                        rec_z=int(t["rz"]) - bu)
        sid = c.add_shot(m, (sx, sy, sz), t["s"], t["h"])
        c.run(t["s"].size)
        tr = c.traces(sid)
        re, im = c.wavefield(sid)
    assert rel(tr, t["traces"]) < 1e-5
    assert rel(re + 1j * im, t["P"] + 1j * t["Q"]) < 1e-5


@pytest.mark.parametrize("graph", [True, False])
def test_run_matches_oracle(ose, graph):
    """150 steps, Ricker + unnormalised Hilbert source, surface receivers."""
    print(f"run 150 steps graph={graph}", flush=True)
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=6))
    nt = 150
    s = ricker(nt, 30.0, case.dt, delay=10)
    h = hilbert_unnorm(s)
    sx, sy, sz = 18, 10, 2
    tr_ref, P, Q = case.oracle().run(nt, case.src_index(sx, sy, sz), s, h, case.bu, case.bl,
                                      case.bf, case.nx, case.ny)
    with ose.Context(0) as c:
        c.set_options(use_graph=graph)
        m = case.add_to(c)
        sid = c.add_shot(m, (case.bl + sx, case.bf + sy, case.bu + sz), s, h)
        c.run(nt)
        tr = c.traces(sid)
        re, im = c.wavefield(sid)
        assert c.step == nt
    print(f"traces {rel(tr, tr_ref):.3g} field {rel(re + 1j * im, P + 1j * Q):.3g}", flush=True)
    assert rel(tr, tr_ref) < 1e-4
    assert rel(re + 1j * im, P + 1j * Q) < 1e-4
