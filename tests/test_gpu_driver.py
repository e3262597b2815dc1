"""The C++ host driver (bin/ose3d) end to end on the GPU: the reference's
main() as key=value presets, writing the raw trace-major seismogram, the .su
file and [y][x][z] snapshots; its traces equal those of the same preset
driven through the Python bindings, and the learned pad loads from a
parameter file."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OSE3D = os.path.join(ROOT, "3d-ose_amd", "bin", "ose3d")


def _run(args, tmp):
    print("ose3d " + " ".join(args), flush=True)
    p = subprocess.run([OSE3D] + args + [f"out={tmp}"], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_reference_preset_files_match_bindings(tmp_path):
    from ose_amd import host as H
    nt = 60
    out = _run(["preset=ref", f"nt={nt}", "snap=1"], str(tmp_path))
    assert out["steps"] == nt and out["gcell_steps_per_s"] > 0
    raw = np.fromfile(tmp_path / "shot0.bin", dtype=np.float32).reshape(101 * 101, nt)
    su, dt = H.read_su(str(tmp_path / "shot0.su"))
    assert abs(dt - 1e-3) < 1e-12 and np.array_equal(su, raw)
    p = np.fromfile(tmp_path / f"p-{nt}-shot0.bin", dtype=np.float32).reshape(152, 152, 152)   # [y][x][z]
    sim = H.Simulation(f"preset=ref nt={nt}").setup()
    try:
        sim.run(nt)
        tr = sim.traces(0, 0, nt)
        re, im = sim.wavefield(0)
    finally:
        sim.close()
    assert np.abs(raw - tr).max() <= 1e-6 * np.abs(tr).max()
    assert np.abs(p - re.transpose(1, 2, 0)).max() <= 1e-6 * np.abs(re).max()
    assert np.all(np.isfinite(raw)) and np.abs(raw).max() > 0


def test_learned_pad_from_parameter_file(tmp_path):
    from ose_amd import boundary
    out = _run(["preset=c3", "nt=20", f"bnet={boundary.WEIGHTS}", "guard=4"], str(tmp_path))
    assert out["setup"]["bnet"] == boundary.WEIGHTS
    assert out["setup"]["models"][0]["padded"] == [272, 272, 272]
    raw = np.fromfile(tmp_path / "shot0.bin", dtype=np.float32)
    assert raw.size == 256 * 256 * 20 and np.all(np.isfinite(raw))


def test_checkpoint_resume_equals_straight_run(tmp_path):
    """20 steps, checkpoint, then a fresh simulation restores and runs the
    remaining 20: traces and wavefield equal the straight 40-step run."""
    from ose_amd import host as H
    print("checkpoint/resume", flush=True)
    nt = 40
    a = H.Simulation(f"preset=ref nt={nt}").setup()
    b = H.Simulation(f"preset=ref nt={nt}").setup()
    try:
        a.run(nt // 2)
        a.checkpoint(tmp_path)
        a.run(nt - nt // 2)
        tr_a, (re_a, im_a) = a.traces(0, 0, nt), a.wavefield(0)
        b.restore(tmp_path)
        b.run(nt - nt // 2)
        tr_b, (re_b, im_b) = b.traces(0, 0, nt), b.wavefield(0)
        bad = H.Simulation("preset=ref nt=40 nshot=2").setup()
        with pytest.raises(Exception):
            bad.restore(tmp_path)                  # other shot set: refused
        bad.close()
    finally:
        a.close()
        b.close()
    assert np.abs(tr_b - tr_a).max() <= 1e-6 * np.abs(tr_a).max()
    assert np.abs(re_b - re_a).max() <= 1e-6 * np.abs(re_a).max()
    assert np.abs(im_b - im_a).max() <= 1e-6 * np.abs(im_a).max()


def test_checkpoint_refuses_another_configuration(tmp_path):
    """A checkpoint records a fingerprint of the run (velocity, order,
    coefficients, tapers, sources, receivers, wavelets, spacing): a
    simulation with the same grid and shots but another velocity or wavelet
    refuses it; checkpoints replace their files (only the latest remain)."""
    from ose_amd import host as H
    print("checkpoint fingerprint", flush=True)
    a = H.Simulation("preset=ref nt=20").setup()
    try:
        a.run(5)
        a.checkpoint(tmp_path)
        a.run(5)
        a.checkpoint(tmp_path)
    finally:
        a.close()
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["ckpt-s10-shot0-im.bin", "ckpt-s10-shot0-re.bin", "ckpt-s10-shot0-traces.bin", "ckpt.txt"]
    for other in ("v1=4200", "fm=30", "korder=12"):
        b = H.Simulation(f"preset=ref nt=20 {other}").setup()
        try:
            with pytest.raises(Exception, match="another configuration"):
                b.restore(tmp_path)
        finally:
            b.close()
    c = H.Simulation("preset=ref nt=20").setup()
    try:
        c.restore(tmp_path)
        assert c.info()["nt"] == 20
    finally:
        c.close()


def test_resume_after_a_tripped_guard_equals_straight_run(tmp_path):
    """The learned pad's guard trips (forced from the host) at step 8; a
    checkpoint at step 12 records it, and the resumed run continues on the
    plain taper exactly as the uninterrupted run does."""
    from ose_amd import boundary
    from ose_amd import host as H
    print("resume after a tripped guard", flush=True)
    cfg = f"nx=64 ny=64 nz=64 boundary=8 model=layered8 seed=3 nt=24 sx=32 sy=32 sz=2 bnet={boundary.WEIGHTS}"
    a = H.Simulation(cfg).setup()
    b = H.Simulation(cfg).setup()
    try:
        a.run(8)
        assert a.boundary_status(0) == {"active": True, "tripped": False}
        a.trip_boundary_net(0)
        a.run(4)
        a.checkpoint(tmp_path)
        a.run(12)
        tr_a, (re_a, im_a) = a.traces(0, 0, 24), a.wavefield(0)
        assert a.boundary_status(0) == {"active": False, "tripped": True}
        b.restore(tmp_path)
        assert b.boundary_status(0) == {"active": False, "tripped": True}
        b.run(12)
        tr_b, (re_b, im_b) = b.traces(0, 0, 24), b.wavefield(0)
    finally:
        a.close()
        b.close()
    assert np.abs(tr_a).max() > 0
    assert np.array_equal(tr_b, tr_a)
    assert np.array_equal(re_b, re_a) and np.array_equal(im_b, im_a)


def test_driver_checkpoint_and_restore(tmp_path):
    nt = 40
    ck = tmp_path / "ck"
    full = tmp_path / "full"
    part = tmp_path / "part"
    full.mkdir()
    part.mkdir()
    out = _run(["preset=ref", f"nt={nt}", f"checkpoint={ck}", "checkpoint_every=20"], str(full))
    assert out["steps"] == nt and (ck / "ckpt.txt").exists()
    out = _run(["preset=ref", f"nt={nt}", f"restore={ck}", "segy=1"], str(part))
    assert out["first_step"] == 20 and out["steps"] == nt
    a = np.fromfile(full / "shot0.bin", dtype=np.float32)
    b = np.fromfile(part / "shot0.bin", dtype=np.float32)
    assert a.size == b.size == 101 * 101 * nt
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
    from ose_amd import host as H
    sg, dt = H.read_segy(part / "shot0.segy")
    assert np.array_equal(sg.ravel(), b)
