"""Pins of the FP64 CPU oracle (CPU only).

Every oracle function is checked against an independent source: the numpy /
scipy fixtures of tests/golden (a second restatement in another library),
brute force on tiny inputs, or a closed form of the mathematics.  The
planted-mistake tests show that the closed-form checks reject plausible
errors (a flipped sign, a dropped factor 2) made in the oracle's inputs.
"""
import os

import numpy as np
import pytest

import pyoracle as orc
from _cases import Case, layered, rel, taper_ref

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_INPUT = "/root/reference/input"


def test_fft_matches_numpy_fixtures():
    g = np.load(os.path.join(GOLD, "fft.npz"))
    assert rel(orc.fft3d(g["a3"]), g["f3"]) < 1e-14
    assert rel(orc.fft3d(g["a3"], +1), g["i3"]) < 1e-14
    for n in (19, 53, 140, 152, 336, 864):
        assert rel(orc.fft1d(g[f"a{n}"]), g[f"f{n}"]) < 1e-14


def test_fft_matches_brute_force():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((3, 5, 7)) + 1j * rng.standard_normal((3, 5, 7))
    assert rel(orc.fft3d(a), orc.dft3d_brute(a)) < 1e-13
    assert rel(orc.fft3d(a, +1), orc.dft3d_brute(a, +1)) < 1e-13


def test_bessel_matches_scipy_fixture_and_identity():
    B = np.load(os.path.join(GOLD, "bessel.npz"))
    for i, n in enumerate(B["n"]):
        for j, x in enumerate(B["x"]):
            assert abs(orc.bessel_jn(int(n), float(x)) - B["J"][i, j]) < 1e-14
    for x in (0.3, 1.904, 2.449, 7.0):
        s = orc.bessel_jn(0, x) + 2 * sum(orc.bessel_jn(2 * k, x) for k in range(1, 40))
        assert abs(s - 1.0) < 1e-14


def test_ricker_and_hilbert():
    W = np.load(os.path.join(GOLD, "wavelet.npz"))
    for nt in (300, 1000, 301):
        nt_, fm, dt = W[f"meta_{nt}"]
        s = orc.ricker(nt, fm, dt)
        assert np.abs(s - W[f"s_{nt}"]).max() < 1e-12
        assert rel(orc.hilbert(s), W[f"h_{nt}"]) < 1e-12
    # closed form: peak amp at i = delay + 1/(fm dt)
    s = orc.ricker(300, 40.0, 1e-3)
    assert np.argmax(s) == 40 + 25 and abs(s.max() - 100.0) < 1e-12
    # the discrete Hilbert transform of cos(2 pi m n/N) is sin(2 pi m n/N); the
    # reference's per-shot copy is not divided by N (quirk Q4)
    N, m = 64, 5
    n = np.arange(N)
    assert np.abs(orc.hilbert(np.cos(2 * np.pi * m * n / N)) - N * np.sin(2 * np.pi * m * n / N)).max() < 1e-10


def test_taper_closed_form():
    a = orc.taper_ref(25, 101, 26)
    assert a[24] == 1.0 and abs(a[0] - np.exp(-(24 / 25) ** 2)) < 1e-15
    assert abs(a[126] - np.exp(-(1 / 26) ** 2)) < 1e-15 and abs(a[-1] - np.exp(-1.0)) < 1e-15
    assert np.all(a[25:126] == 1.0)
    c = orc.taper_cerjan(20, 100, 20, 0.015)
    assert abs(c[0] - np.exp(-(0.3) ** 2)) < 1e-15 and abs(c[19] - np.exp(-(0.015) ** 2)) < 1e-15


def test_judge_odd():
    assert orc.judge_odd(25, 151) == (26, 152)
    assert orc.judge_odd(20, 140) == (20, 140)


def test_velocity_padding_is_edge_mode():
    rng = np.random.default_rng(1)
    vi = rng.uniform(1500, 4500, (5, 6, 7)).astype(np.float32)
    v = orc.velocity_pad_interior(vi, 3, 2, 4, 7 + 3 + 5, 6 + 2 + 2, 5 + 4 + 1)
    ref = np.pad(vi.astype(np.float64), ((4, 1), (2, 2), (3, 5)), mode="edge")
    assert np.array_equal(v, ref)


@pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="reference inputs not mounted")
def test_velocity_file_order_on_reference_input():
    # input/v3d_nx101_ny101_nz101.bin: [x][y][z] file order (quirk Q13)
    f = np.fromfile(os.path.join(REF_INPUT, "v3d_nx101_ny101_nz101.bin"), dtype="<f4")
    v = orc.velocity_from_file_order(f, 101, 101, 101, 25, 25, 25, 152, 152, 152)
    assert v.shape == (152, 152, 152)
    assert np.all(v[: 25 + 30] == 3000.0) and np.all(v[25 + 30:] == 4000.0)


def _plane_wave_case(K=None, coef_hook=None):
    n = 32
    case = Case((n - 8, n - 8, n - 8), (4, 4, 4, 4, 4, 4), np.full((n - 8,) * 3, 3000.0),
                taper="none", K=K, d=(10.0, 12.0, 8.0))
    if coef_hook:
        case.coef = coef_hook(case.coef.copy())
    z, y, x = np.meshgrid(*[np.arange(s) for s in case.shape], indexing="ij")
    kx, ky, kz = 3, 2, 5
    phase = 2 * np.pi * (kx * x / case.nxb + ky * y / case.nyb + kz * z / case.nzb)
    u = np.exp(1j * phase)
    kabs = 2 * np.pi * np.sqrt((kx / (case.d[0] * case.nxb)) ** 2 + (ky / (case.d[1] * case.nyb)) ** 2
                               + (kz / (case.d[2] * case.nzb)) ** 2)
    exact = np.exp(-1j * case.dt * 3000.0 * kabs) * u
    P, Q = case.oracle().step(u.real, u.imag)
    zz = case.R * case.dt
    tail = 2 * sum(abs(orc.bessel_jn(k, zz)) for k in range(case.K + 1, case.K + 60))
    return np.abs(P + 1j * Q - exact).max(), tail


def test_jacobi_anger_plane_wave():
    """Constant v, no taper: one step of a Fourier mode is exp(-i dt v|k|)
    times the mode, to within the series truncation tail."""
    err, tail = _plane_wave_case()
    assert err <= tail + 1e-12
    err20, tail20 = _plane_wave_case(K=20)
    assert err20 < 1e-12 and tail20 < 1e-15


def test_planted_mistakes_are_caught():
    # a sign error in the odd coefficients (exp(+i...) instead of exp(-i...))
    err, tail = _plane_wave_case(coef_hook=lambda c: c * (-1.0) ** np.arange(c.size))
    assert err > 100 * (tail + 1e-12)
    # a dropped factor 2 in c_k, k >= 1
    def halve(c):
        c[1:] *= 0.5
        return c
    err, tail = _plane_wave_case(coef_hook=halve)
    assert err > 100 * (tail + 1e-12)


def test_norm_conserved_without_taper():
    """exp(-i dt v|k|) is unitary for constant v."""
    n = 24
    case = Case((n - 8,) * 3, (4,) * 6, np.full((n - 8,) * 3, 2500.0), taper="none", K=18)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(case.shape) + 1j * rng.standard_normal(case.shape)
    P, Q = case.oracle().step(u.real, u.imag)
    assert abs(np.linalg.norm(P + 1j * Q) / np.linalg.norm(u) - 1.0) < 1e-12


def test_taper_power_law_on_constant_field():
    """A constant field has L u = 0, so W_even = u, W_odd = 0 and the step is
    sum_k c_k g^{e_k} u with e_k = M for k <= 1 and M - k + 2 for k >= 2
    (the taper hits the accumulator M times, quirk Q3)."""
    case = Case((12, 10, 8), (3, 3, 3, 3, 3, 3), layered(12, 10, 8, seed=1))
    M = case.K - 1
    g = case.gz[:, None, None] * case.gy[None, :, None] * case.gx[None, None, :]
    u = np.full(case.shape, 1.5)
    P, Q = case.oracle().step(u, 0 * u)
    expect = case.coef[0] * g ** M
    for k in range(2, case.K + 1, 2):
        expect = expect + case.coef[k] * g ** (M - k + 2)
    assert np.abs(P - 1.5 * expect).max() < 1e-12 and np.abs(Q).max() < 1e-12


def test_tiny_run_matches_numpy_restatement():
    t = np.load(os.path.join(GOLD, "step_tiny.npz"))
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = (int(x) for x in t["dims"])
    m = orc.Model(t["v"], *t["d"], float(t["R"]), t["coef"], t["gx"], t["gy"], t["gz"])
    tr, P, Q = m.run(t["s"].size, int(t["src"]), t["s"], t["h"], int(t["rz"]), bl, bf, nx, ny)
    assert rel(tr, t["traces"]) < 1e-12
    assert rel(P + 1j * Q, t["P"] + 1j * t["Q"]) < 1e-12


def test_apply_L_symbol_on_plane_wave():
    case = Case((16, 16, 16), (4,) * 6, np.full((16,) * 3, 2000.0), taper="none")
    z, y, x = np.meshgrid(*[np.arange(s) for s in case.shape], indexing="ij")
    u = np.exp(2j * np.pi * (-3 * x / 24 + 5 * z / 24))
    k = 2 * np.pi * np.hypot(3 / (10 * 24), 5 / (10 * 24))
    assert rel(case.oracle().apply_L(u), 2000.0 / case.R * k * u) < 1e-13


def test_resume_from_state_is_bit_identical():
    """The per-shot state is (P, Q, it): N steps equal N/2 steps, then N/2
    more started from the saved (P, Q) with the wavelet from step N/2 on."""
    t = np.load(os.path.join(GOLD, "step_tiny.npz"))
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = (int(x) for x in t["dims"])
    m = orc.Model(t["v"], *t["d"], float(t["R"]), t["coef"], t["gx"], t["gy"], t["gz"])
    n = t["s"].size
    h = n // 2
    args = (int(t["src"]),)
    tr, P, Q = m.run(n, *args, t["s"], t["h"], int(t["rz"]), bl, bf, nx, ny)
    tr1, P1, Q1 = m.run(h, *args, t["s"][:h], t["h"][:h], int(t["rz"]), bl, bf, nx, ny)
    tr2, P2, Q2 = m.run(n - h, *args, t["s"][h:], t["h"][h:], int(t["rz"]), bl, bf, nx, ny, P=P1, Q=Q1)
    assert np.array_equal(np.concatenate([tr1, tr2], axis=1), tr)
    assert np.array_equal(P2, P) and np.array_equal(Q2, Q)


def test_profile_source_pins_to_point_runs():
    """A unit-vector profile is the point source (bit for bit); a profile
    with two nonzero x entries equals the weighted sum of two point runs
    (linearity of the step)."""
    t = np.load(os.path.join(GOLD, "step_tiny.npz"))
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = (int(x) for x in t["dims"])
    m = orc.Model(t["v"], *t["d"], float(t["R"]), t["coef"], t["gx"], t["gy"], t["gz"])
    n = min(t["s"].size, 6)
    s, h = t["s"][:n], t["h"][:n]
    src = int(t["src"])
    sz, sy, sx = src // (nxb * nyb), (src // nxb) % nyb, src % nxb
    e = [np.zeros(k) for k in (nxb, nyb, nzb)]
    e[0][sx], e[1][sy], e[2][sz] = 1.0, 1.0, 1.0
    tr, P, Q = m.run(n, src, s, h, int(t["rz"]), bl, bf, nx, ny)
    tp, Pp, Qp = orc.run_profile(m, n, e[0], e[1], e[2], s, h, int(t["rz"]), bl, bf, nx, ny)
    assert np.array_equal(tp, tr) and np.array_equal(Pp, P) and np.array_equal(Qp, Q)
    px = np.zeros(nxb)
    px[sx], px[sx - 2] = 0.75, -0.5
    tp, Pp, Qp = orc.run_profile(m, n, px, e[1], e[2], s, h, int(t["rz"]), bl, bf, nx, ny)
    _, P2, Q2 = m.run(n, src - 2, s, h, int(t["rz"]), bl, bf, nx, ny)
    assert rel(Pp + 1j * Qp, 0.75 * (P + 1j * Q) - 0.5 * (P2 + 1j * Q2)) < 1e-12
