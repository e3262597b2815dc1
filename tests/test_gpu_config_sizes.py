"""Correctness at the benchmarked sizes.

The bench quotes two batched workloads: 768^3 x 8 shots (864^3 padded, five
out-of-place fields, K = 11; the learned leg on 784^3) and 16 shots on
512^2 x 384 (600^2 x 480 padded) plus four uneven models in two lanes.  Here
each runs at that size:

* closed form: constant velocity, no taper, every shot a different sum of
  Fourier modes (low, high and Nyquist wavenumbers); after n steps each mode
  must carry exp(-i n dt v |k|) to within the series tail and FP32 rounding,
  and nothing may leak into other modes (a wrong batch stride, plane stride
  or pencil index gives O(1) there);
* batch against single shot: with the real preset (salt or layered model,
  taper, Ricker sources, receivers) every shot of the batch gives the same
  traces and wavefield, bit for bit, as the same shot run alone with the
  same launch shapes;
* learned leg: the 8-cell network at 784^3 x 8 keeps the energy
  sum |u|^2 / v from growing (the propagator is unitary in that norm and
  the network's gain passive), stays finite and does not trip its guard.

Reference: the time loop ose_3d.cu:242-279 around wavefield_extro_ori2,
ose_3d_lib.cu:639-874.
"""
import os

import numpy as np
import pytest

import _planewave as pw

# each test moves tens of GB through the host for its checks: minutes, not seconds
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
D, DT = pw.D, pw.DT
_add_sim_model = pw.add_sim_model


def _digest(a):
    try:
        import xxhash
        return xxhash.xxh3_128_hexdigest(memoryview(np.ascontiguousarray(a)).cast("B"))
    except ImportError:
        import hashlib
        return hashlib.blake2b(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()


def _const_batch(preset, v, nsteps, lanes=None, extra=""):
    return pw.const_batch(preset, v, nsteps, lanes=lanes, extra=extra, log=lambda m: print(m, flush=True))


def _assert_modes(results):
    for r in results:
        assert pw.within(r), r


def test_c4_batch_plane_waves_closed_form():
    """864^3 x 8 shots, K = 11 (v = 4500 m/s as the salt body), 20 steps."""
    print("c4 geometry, 8 x 864^3, constant v, 20 steps", flush=True)
    res = _const_batch("c4", 4500.0, 20)
    assert len(res) == 8
    _assert_modes(res)


def test_c2_batch_plane_waves_closed_form():
    """16 shots of 600^2 x 480 and four uneven models (272^3 ...
    400 x 320 x 240 padded) in two lanes, 20 steps."""
    print("c2 geometry, 16 + 4 shots, constant v, 20 steps, two lanes", flush=True)
    res = _const_batch("c2", 4000.0, 20, lanes=2)
    assert len(res) == 20
    _assert_modes(res)


@pytest.mark.parametrize("preset,nsteps", [("c1", 2000), ("c4", 300)])
def test_long_horizon_closed_form(preset, nsteps):
    """FP32 drift over the configurations' horizons: the 256^3 workload's
    336^3 grid (K = 11 at v = 4500 m/s) for its 2000 steps, and one 864^3
    shot of the 768^3 workload for 300 steps."""
    print(f"{preset} geometry, one shot, constant v, {nsteps} steps", flush=True)
    res = _const_batch(preset, 4500.0, nsteps, extra="max_shots=1")
    assert len(res) == 1
    _assert_modes(res)


def _batch_vs_single(preset, nsteps, extra=""):
    """Every shot of a preset's batch against the same shot run alone with
    the batch's launch shapes: traces and wavefield bit for bit."""
    import ose_amd
    from ose_amd.host import Simulation
    sim = Simulation(f"preset={preset} nt={nsteps} {extra}").build()
    try:
        batch = {}
        tuning = {}
        with ose_amd.Context(0) as c:
            mids = [_add_sim_model(c, sim, mi)[0] for mi in range(sim.nmodels())]
            ids = []
            for k in range(sim.nshots()):
                sh = sim.shot(k)
                ids.append(c.add_shot(mids[sh["model"]], tuple(sh["src"]), sh["s"], sh["h"]))
            c.run(nsteps)
            c.sync()
            for mi, mid in enumerate(mids):
                tuning[mi] = c.tuning_modes(mid)
            for k, sid in enumerate(ids):
                tr = c.traces(sid)
                re, im = c.wavefield(sid)
                assert np.isfinite(re).all()
                batch[k] = (tr, _digest(re), _digest(im), float(np.abs(re).max()))
                del re, im
        print(f"  batch of {len(ids)} shots done; launch shapes {tuning}", flush=True)
        for k in range(sim.nshots()):
            sh = sim.shot(k)
            with ose_amd.Context(0) as c:
                c.set_autotune(False)
                mid = _add_sim_model(c, sim, sh["model"])[0]
                c.set_tuning_modes(mid, tuning[sh["model"]])
                sid = c.add_shot(mid, tuple(sh["src"]), sh["s"], sh["h"])
                c.run(nsteps)
                tr = c.traces(sid)
                re, im = c.wavefield(sid)
                got = (_digest(re), _digest(im))
                del re, im
            tb, hr, hi, amax = batch[k]
            print(f"  shot {k}: traces equal {np.array_equal(tr, tb)}, field equal {got == (hr, hi)}, "
                  f"max |Re u| {amax:.3g}", flush=True)
            assert amax > 0
            assert np.array_equal(tr, tb)
            assert got == (hr, hi)
    finally:
        sim.close()


def test_c4_batched_shots_equal_single_shots():
    """768^3 x 8 shots on the salt model, 40-cell taper, 6 steps."""
    print("c4 batch vs single shots", flush=True)
    _batch_vs_single("c4", 6, "delay=-40")


def test_c2_batched_shots_equal_single_shots():
    """16 + 4 shots on five models in two lanes, 6 steps."""
    print("c2 batch vs single shots", flush=True)
    _batch_vs_single("c2", 6, "delay=-40")


def test_c4_learned_leg_is_passive_and_stays_on():
    """784^3 x 8 shots with the committed 8-cell network: from a wave packet
    aimed at a different face per shot, 12 steps; the energy sum |u|^2 / v
    (the norm the propagator conserves on a variable velocity) never grows
    beyond FP32 rounding, the field stays finite, the guard holds."""
    from ose_amd import boundary
    from ose_amd.host import Simulation
    print("c4 learned leg, 8 x 784^3, 12 steps", flush=True)
    sim = Simulation(f"preset=c4 boundary=8 amp=0 nt=12 bnet={boundary.WEIGHTS}").setup()
    try:
        d = sim.model_dims(0)
        nxb, nyb, nzb = d[9], d[10], d[11]
        assert (nxb, nyb, nzb) == (784, 784, 784)
        v = sim.model(0)["v"]
        shape = (nzb, nyb, nxb)
        e0 = []
        for k in range(sim.nshots()):
            # a Gaussian packet 24 cells from face k % 6 (16 inside the pad's
            # edge), moving towards it
            ax, side = (k % 6) // 2, (k % 6) % 2
            c = [nxb // 2, nyb // 2, nzb // 2]
            n = (nxb, nyb, nzb)[ax]
            c[ax] = 24 if side == 0 else n - 25
            m = [0, 0, 0]
            m[ax] = -40 if side == 0 else 40
            g = [np.exp(-0.5 * ((np.arange(nn) - cc) / 12.0) ** 2) for nn, cc in zip((nxb, nyb, nzb), c)]
            ph = [np.exp(2j * np.pi * mm * np.arange(nn) / nn) for nn, mm in zip((nxb, nyb, nzb), m)]
            fx, fy, fz = (gg * pp for gg, pp in zip(g, ph))
            u = (fz[:, None, None] * (fy[:, None] * fx[None, :])[None]).astype(np.complex64)
            re, im = np.ascontiguousarray(u.real), np.ascontiguousarray(u.imag)
            del u
            e0.append(pw.weighted_energy(re, im, v))
            sim.set_wavefield(k, re, im)
            del re, im
        sim.run(12)
        st = sim.boundary_status(0)
        for k in range(sim.nshots()):
            re, im = sim.wavefield(k)
            assert np.isfinite(re).all() and np.isfinite(im).all()
            e = pw.weighted_energy(re, im, v)
            del re, im
            print(f"  shot {k}: energy ratio after 12 steps {e / e0[k]:.6f}", flush=True)
            assert e <= e0[k] * (1 + 12 * 2e-6)
        assert st["active"] and not st["tripped"]
    finally:
        sim.close()
