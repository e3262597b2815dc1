"""Snapshot streaming on the GPU (ose_set_snapshots): every encoding, box
and lane choice equals the numpy restatement of the encoding applied to the
state read back at that step; a run that streams a snapshot of every shot
after every step without intermediate syncs delivers the same bytes as a
step-by-step run; the host driver writes decodable .osnap files and float32
files in the reference's snapshot order."""
import os
import subprocess
import threading

import numpy as np
import pytest

import _snapcodec as sc
from _cases import Case, hilbert_unnorm, layered, ricker

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OSE3D = os.path.join(ROOT, "3d-ose_amd", "bin", "ose3d")


def _context(ose_amd, nt):
    a = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=31))
    b = Case((52, 20, 20), (6, 6, 6, 6, 6, 6), layered(52, 20, 20, seed=32, v0=1800.0, v1=3000.0))
    s = ricker(nt, 30.0, a.dt, delay=6)
    h = hilbert_unnorm(s)
    c = ose_amd.Context(0)
    c.set_autotune(False)
    ma, mb = a.add_to(c), b.add_to(c)
    c.add_shot(ma, (a.bl + 18, a.bf + 10, a.bu + 4), s, h)
    c.add_shot(ma, (a.bl + 6, a.bf + 5, a.bu + 12), s, h)
    c.add_shot(mb, (b.bl + 26, b.bf + 10, b.bu + 3), s, h)
    return c, (a, a, b)


class _Sink:
    def __init__(self):
        self.got = {}
        self.lock = threading.Lock()

    def __call__(self, info, data):
        with self.lock:
            self.got[(info["shot_id"], info["step"])] = (info, data)


def _expected(case, re, im, lanes, fmt, interior):
    if interior:
        box = (slice(case.bu, case.bu + case.nz), slice(case.bf, case.bf + case.ny), slice(case.bl, case.bl + case.nx))
        re, im = re[box], im[box]
    data = [v for bit, v in ((1, re), (2, im)) if lanes & bit]
    return sc.encode(fmt, data), re.shape


@pytest.mark.parametrize("fmt,interior,lanes", [(0, False, 3), (1, True, 3), (2, False, 2), (3, True, 1),
                                               (3, False, 3), (2, True, 3)])
def test_streamed_snapshots_equal_encoded_state(fmt, interior, lanes):
    import ose_amd
    print(f"snapshots fmt={fmt} interior={interior} lanes={lanes}", flush=True)
    every, nt = 3, 12
    c, cases = _context(ose_amd, nt)
    sink = _Sink()
    try:
        c.set_snapshots(every, sink, lanes=lanes, fmt=["f32", "bf16", "q16", "q8"][fmt], interior=interior)
        for k in range(nt // every):
            c.run(every)
            states = [c.wavefield(sid) for sid in range(3)]   # syncs: snapshots of this step delivered
            step = (k + 1) * every
            for sid, (re, im) in enumerate(states):
                info, data = sink.got[(sid, step)]
                want, shape = _expected(cases[sid], re, im, lanes, fmt, interior)
                assert (info["nz"], info["ny"], info["nx"]) == shape
                assert info["lanes"] == lanes and info["format"] == fmt and info["bytes"] == len(want)
                assert data == want, (sid, step)
        assert len(sink.got) == 3 * nt // every
    finally:
        c.close()


def test_streaming_without_syncs_matches_step_by_step():
    """Every step streams all three shots through three device and three
    pinned host slots: the enqueuing side and the copier wait for free
    slots, nothing is lost or reordered."""
    import ose_amd
    print("snapshots every step, one run call", flush=True)
    nt = 12
    runs = []
    for chunk in (nt, 1):
        c, _ = _context(ose_amd, nt)
        sink = _Sink()
        try:
            c.set_snapshots(1, sink, lanes=3, fmt="q8", interior=False)
            for _ in range(nt // chunk):
                c.run(chunk)
            c.snapshot_drain()
            final = c.wavefield(2)
            c.set_snapshots(0)
            c.run(1)                                   # streaming off: nothing more arrives
            c.sync()
        finally:
            c.close()
        assert sorted(sink.got) == [(sid, it) for sid in range(3) for it in range(1, nt + 1)]
        runs.append((sink.got, final))
    (a, fa), (b, fb) = runs
    for key in a:
        assert a[key][1] == b[key][1], key
    assert np.array_equal(fa[1], fb[1])


def test_simulation_streams_float32_in_reference_order(tmp_path):
    from ose_amd import host as H
    print("simulation f32 snapshot files", flush=True)
    sim = H.Simulation("preset=ref nt=10").setup()
    try:
        sim.stream_snapshots(tmp_path, 5)
        sim.run(5)
        re, im = sim.wavefield(0)
        sim.run(5)
        sim.write(str(tmp_path))                       # drains; reports a failed write
    finally:
        sim.close()
    q = np.fromfile(tmp_path / "q-5-shot0.bin", dtype=np.float32).reshape(152, 152, 152)   # [y][x][z]
    p = np.fromfile(tmp_path / "p-5-shot0.bin", dtype=np.float32).reshape(152, 152, 152)
    assert np.array_equal(q, im.transpose(1, 2, 0)) and np.array_equal(p, re.transpose(1, 2, 0))
    assert (tmp_path / "q-10-shot0.bin").exists()


def test_driver_streams_q8_interior(tmp_path):
    from ose_amd import host as H
    print("ose3d snap_every q8 interior", flush=True)
    args = [OSE3D, "preset=ref", "nt=20", "snap=1", "snap_every=10", "snap_format=q8", "snap_region=interior",
            "snap_lanes=q", f"out={tmp_path}"]
    p = subprocess.run(args, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    got, meta = H.read_osnap(tmp_path / "q-20-shot0.osnap")
    assert meta["lane"] == 1 and meta["step"] == 20 and (meta["nx"], meta["ny"], meta["nz"]) == (101, 101, 101)
    assert (meta["x0"], meta["y0"], meta["z0"]) == (25, 25, 25)
    full = np.fromfile(tmp_path / "q-20-shot0.bin", dtype=np.float32).reshape(152, 152, 152).transpose(2, 0, 1)
    ref = full[25:126, 25:126, 25:126]
    assert (np.abs(got - ref) <= sc.bound(sc.Q8, ref)).all()
    assert np.abs(ref).max() > 0
    assert (tmp_path / "q-10-shot0.osnap").exists() and not (tmp_path / "p-10-shot0.osnap").exists()


def test_failed_snapshot_write_is_reported(tmp_path):
    """A snapshot the writer thread cannot store is reported (by the next
    run call, or at the latest by write, which drains the stream) with the
    file named, instead of being dropped."""
    from ose_amd import host as H
    print("snapshot write failure", flush=True)
    sim = H.Simulation("preset=ref nt=10").setup()
    try:
        sim.stream_snapshots(tmp_path / "missing" / "dir", 5)
        with pytest.raises(H.OseError, match="cannot write snapshot"):
            sim.run(10)
            sim.write(str(tmp_path))
    finally:
        sim.close()
