"""Receivers at listed cells (ose_set_receivers): the geometry the reference
declares (gxbeg..jgz, ng, csdgather; ose_3d.cu:100-114, sg_init_device
ose_3d_kernel.cu:202-212) recorded on the GPU after every step.

Checks: a listed receiver equals Im(u) of the state read back after each
step (exactly: the gather only copies), listed receivers on the plane equal
the plane traces of an identical shot (exactly), receivers at depth match
the FP64 oracle's traces on that depth's plane, a model whose shots all list
receivers records no plane, and the host driver writes the line geometry
with receiver depths into the SU headers."""
import os
import struct
import subprocess

import numpy as np
import pytest

from _cases import Case, hilbert_unnorm, layered, rel, ricker

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OSE3D = os.path.join(ROOT, "3d-ose_amd", "bin", "ose3d")


def test_listed_receivers_match_state_plane_and_oracle():
    import ose_amd
    print("listed receivers: state, plane, oracle", flush=True)
    nt = 60
    a = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=21), rec_z=0)
    b = Case((52, 20, 20), (6, 6, 6, 6, 6, 6), layered(52, 20, 20, seed=22, v0=1800.0, v1=3000.0))
    s = ricker(nt, 30.0, a.dt, delay=10)
    h = hilbert_unnorm(s)
    src_a = (a.bl + 18, a.bf + 10, a.bu + 4)
    src_b = (b.bl + 26, b.bf + 10, b.bu + 3)
    # interior cells: on the plane, at depth 9 and 17, one in the pad
    on_plane = [(3, 4, 0), (18, 10, 0), (30, 15, 0)]
    deep = [(18, 10, 9), (7, 3, 9), (25, 12, 17)]
    rec_a = np.array([(x + a.bl, y + a.bf, z + a.bu) for (x, y, z) in on_plane + deep] + [(1, 2, 3)])
    rec_b = np.array([(b.bl + x, b.bf + 10, b.bu + 9) for x in (5, 26, 40)])
    with ose_amd.Context(0) as c:
        ma, mb = a.add_to(c), b.add_to(c)
        listed = c.add_shot(ma, src_a, s, h)
        plane = c.add_shot(ma, src_a, s, h)          # same source: recorded on the plane
        only = c.add_shot(mb, src_b, s, h)
        c.set_receivers(listed, rec_a)
        c.set_receivers(only, rec_b)
        assert c.receiver_count(listed) == len(rec_a) and c.receiver_count(plane) == 36 * 20
        assert c.receiver_count(only) == 3
        state = np.zeros((len(rec_a), nt), dtype=np.float32)
        for it in range(nt):
            c.run(1)
            _, im = c.wavefield(listed)
            state[:, it] = im[rec_a[:, 2], rec_a[:, 1], rec_a[:, 0]]
        tr = c.traces(listed)
        tp = c.traces(plane)
        tb = c.traces(only)
        with pytest.raises(ose_amd.OseError):
            c.set_receivers(listed, rec_a[:2])       # the context is running
    assert tr.shape == (len(rec_a), nt) and tb.shape == (3, nt)
    assert np.array_equal(tr, state)
    for r, (x, y, _) in enumerate(on_plane):
        assert np.array_equal(tr[r], tp[y * 36 + x])
    # the oracle records one plane per run (its depth argument is a padded index)
    orc = a.oracle()
    planes = {z: orc.run(nt, a.src_index(18, 10, 4), s, h, a.bu + z, a.bl, a.bf, a.nx, a.ny)[0] for z in (9, 17)}
    want = np.stack([planes[z][y * a.nx + x] for (x, y, z) in deep])
    assert rel(tr[len(on_plane):len(on_plane) + len(deep)], want) < 1e-5
    otr, _, _ = b.oracle().run(nt, b.src_index(26, 10, 3), s, h, b.bu + 9, b.bl, b.bf, b.nx, b.ny)
    assert rel(tb, otr[[10 * b.nx + x for x in (5, 26, 40)]]) < 1e-5


def test_receivers_outside_the_grid_are_rejected():
    import ose_amd
    case = Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=23))
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        sid = c.add_shot(m, (20, 15, 8), np.zeros(4), np.zeros(4))
        for bad in ([(48, 0, 0)], [(0, -1, 0)], [(0, 0, 48)]):
            with pytest.raises(ose_amd.OseError):
                c.set_receivers(sid, bad)
        c.set_receivers(sid, [(0, 0, 47)])
        c.set_receivers(sid, np.zeros((0, 3)))       # back to the plane
        assert c.receiver_count(sid) == 36 * 20


def test_driver_receiver_line_with_moving_spread(tmp_path):
    """ose3d with a receiver line at depth that moves with the source
    (csdgather=1): traces equal the plane traces of that depth, and the SU
    headers carry the receiver coordinates and depth."""
    from ose_amd import host as H
    nt = 30
    geo = ["nshot=2", "sx=40", "sy=50", "sz=0", "jsx=10", "jsy=0", "jsz=0", "ng=20", "gxbeg=30", "gybeg=50",
           "gzbeg=10", "jgx=1", "csdgather=1"]
    print("ose3d receiver line", flush=True)
    p = subprocess.run([OSE3D, "preset=ref", f"nt={nt}"] + geo + [f"out={tmp_path}"], capture_output=True,
                       text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    sim = H.Simulation(f"preset=ref nt={nt} nshot=2 sx=40 sy=50 sz=0 jsx=10 jsy=0 jsz=0 rec_z=10").setup()
    try:
        sim.run(nt)
        planes = [sim.traces(k, 0, nt) for k in range(2)]
    finally:
        sim.close()
    for k in range(2):
        raw = np.fromfile(tmp_path / f"shot{k}.bin", dtype=np.float32).reshape(20, nt)
        want = planes[k][[50 * 101 + 30 + 10 * k + r for r in range(20)]]
        assert np.abs(raw - want).max() <= 1e-5 * np.abs(want).max()
        su, _ = H.read_su(str(tmp_path / f"shot{k}.su"))
        assert np.array_equal(su, raw)
        hdr = open(tmp_path / f"shot{k}.su", "rb").read()[240 + 4 * nt:][:240]   # second receiver
        assert struct.unpack_from("<i", hdr, 80)[0] == (31 + 10 * k) * 10 * 100   # gx (m * 100)
        assert struct.unpack_from("<i", hdr, 40)[0] == -10 * 10 * 100             # gelev = -depth
