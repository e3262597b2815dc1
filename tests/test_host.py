"""The C++ host library (libose_host.so) against the oracle and fixtures.
CPU only: nothing here touches the device."""
import ctypes
import os
import struct

import numpy as np
import pytest

import pyoracle as orc
from _cases import rel

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_INPUT = "/root/reference/input"


@pytest.fixture(scope="module")
def H():
    from ose_amd import host
    return host


def test_bessel_miller_matches_scipy(H):
    B = np.load(os.path.join(GOLD, "bessel.npz"))
    for j, x in enumerate(B["x"]):
        J = H.bessel_all(float(x), 24)
        assert np.abs(J - B["J"][:, j]).max() < 1e-14


def test_bessel_matches_glibc_jn(H):
    for x in (1e-3, 0.7, 2.177, 2.449, 9.5, 20.0):
        J = H.bessel_all(x, 30)
        ref = np.array([orc.bessel_jn(n, x) for n in range(31)])
        assert np.abs(J - ref).max() < 1e-14


def test_coefficients_and_orders(H):
    z = 2.449
    c = H.coefficients(z, 11)
    assert abs(c[0] - orc.bessel_jn(0, z)) < 1e-15
    assert all(abs(c[k] - 2 * orc.bessel_jn(k, z)) < 1e-15 for k in range(1, 12))
    # reference rule M = (int)z + 6, K = M + 1 (ose_3d.cu:197)
    assert H.order_reference(2.177) == 9
    assert H.order_reference(1.904) == 8
    # tail rule 2 sum_{k>K}|J_k(z)| < 1e-7
    for z in (1.904, 2.177, 2.449, 2.816):
        K = H.order_tail(z, 1e-7)
        tail = lambda K: 2 * sum(abs(orc.bessel_jn(k, z)) for k in range(K + 1, K + 60))
        assert tail(K) < 1e-7 <= tail(K - 1)
    R = H.spectral_radius(4000.0, 10, 10, 10)
    assert abs(R - 4000 * np.pi * np.sqrt(0.03)) < 1e-9


def test_wavelet_matches_oracle(H):
    for nt, fm in ((300, 40.0), (1000, 25.0), (301, 20.0)):
        s = H.ricker(nt, fm, 1e-3)
        assert np.abs(s - orc.ricker(nt, fm, 1e-3)).max() < 1e-12
        h = H.hilbert(s)
        assert rel(h, orc.hilbert(s)) < 1e-12
        assert rel(H.hilbert(s, normalize=True), h / nt) < 1e-14


def test_taper_matches_oracle(H):
    assert np.abs(H.taper(25, 101, 26) - orc.taper_ref(25, 101, 26)).max() < 1e-7
    assert np.abs(H.taper(20, 100, 20, "cerjan", 0.015) - orc.taper_cerjan(20, 100, 20, 0.015)).max() < 1e-7
    assert np.all(H.taper(3, 4, 5, "none") == 1.0)
    assert H.judge_odd(25, 151) == (26, 152)


def test_pad_velocity_edge_mode(H):
    rng = np.random.default_rng(3)
    vi = rng.uniform(1500, 4500, (7, 5, 6)).astype(np.float32)
    pads = (2, 3, 4, 1, 3, 2)
    got = H.pad_velocity(vi, pads)
    ref = np.pad(vi, ((3, 2), (4, 1), (2, 3)), mode="edge")
    assert np.array_equal(got, ref)
    o = orc.velocity_pad_interior(vi, 2, 4, 3, 6 + 5, 5 + 5, 7 + 5)
    assert np.array_equal(got.astype(np.float64), o)


@pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="reference inputs not mounted")
def test_reference_velocity_file_equals_twolayer_generator(H):
    v = H.read_velocity_xyz(os.path.join(REF_INPUT, "v3d_nx101_ny101_nz101.bin"), 101, 101, 101)
    g = H.model("twolayer", 101, 101, 101, v0=3000.0, v1=4000.0, iface=30)
    assert np.array_equal(v, g)
    # the other two reference inputs parse with their own dimensions
    v2 = H.read_velocity_xyz(os.path.join(REF_INPUT, "v3d_nx100_ny101_nz102.bin"), 100, 101, 102)
    assert set(np.unique(v2)) == {3000.0, 4000.0}
    v3 = H.read_velocity_xyz(os.path.join(REF_INPUT, "3d_number-100-100-100-3.bin"), 100, 100, 100)
    assert v3.min() == 6000.0 and v3.max() == 8500.0


def test_read_velocity_rejects_short_file(H, tmp_path):
    p = tmp_path / "short.bin"
    np.zeros(10, dtype=np.float32).tofile(p)
    with pytest.raises(Exception):
        H.read_velocity_xyz(str(p), 4, 4, 4)


def test_generators_are_seeded_and_in_range(H):
    a = H.model("layered8", 48, 40, 32, seed=2)
    b = H.model("layered8", 48, 40, 32, seed=2)
    c = H.model("layered8", 48, 40, 32, seed=3)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert 1500 * 0.8 < a.min() and a.max() < 4500 * 1.2
    # velocity increases with depth on average
    assert a[:4].mean() < a[-4:].mean()
    s = H.model("salt12", 64, 64, 64, seed=5)
    assert (s == 4500.0).sum() > 0 and s.min() >= 1500.0


def test_su_round_trip_and_header(H, tmp_path):
    rng = np.random.default_rng(4)
    tr = rng.standard_normal((6, 37)).astype(np.float32)
    p = str(tmp_path / "x.su")
    H.write_su(p, tr, 1e-3, rec_x=np.arange(6) * 10.0, rec_y=np.zeros(6), src=(25.0, 0.0), shot=2)
    got, dt = H.read_su(p)
    assert np.array_equal(got, tr) and abs(dt - 1e-3) < 1e-12
    raw = open(p, "rb").read()
    assert len(raw) == 6 * (240 + 4 * 37)
    h = raw[240 + 4 * 37:][:240]   # second trace
    assert struct.unpack_from("<i", h, 0)[0] == 2          # tracl
    assert struct.unpack_from("<i", h, 8)[0] == 3          # fldr = shot + 1
    assert struct.unpack_from("<h", h, 70)[0] == -100      # scalco
    assert struct.unpack_from("<i", h, 80)[0] == 1000      # gx = 10 m * 100
    assert struct.unpack_from("<H", h, 114)[0] == 37       # ns
    assert struct.unpack_from("<H", h, 116)[0] == 1000     # dt (us)
    assert struct.unpack_from("<i", h, 40)[0] == 0         # gelev: no depth given
    # receiver depths and the source depth: gelev = -depth, sdepth, scalel
    H.write_su(p, tr, 1e-3, rec_x=np.arange(6) * 10.0, rec_y=np.zeros(6), rec_z=np.arange(6) * 5.0,
               src=(25.0, 0.0, 30.0), shot=2)
    raw = open(p, "rb").read()
    h = raw[3 * (240 + 4 * 37):][:240]   # fourth trace
    assert struct.unpack_from("<i", h, 40)[0] == -1500     # gelev = -(15 m) * 100
    assert struct.unpack_from("<i", h, 48)[0] == 3000      # sdepth = 30 m * 100
    assert struct.unpack_from("<h", h, 68)[0] == -100      # scalel


def test_snapshot_order(H, tmp_path):
    f = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)   # [z][y][x]
    p = str(tmp_path / "s.bin")
    H.write_snapshot(p, f)
    got = np.fromfile(p, dtype=np.float32).reshape(3, 4, 2)       # [y][x][z]
    assert np.array_equal(got, f.transpose(1, 2, 0))


@pytest.mark.parametrize("preset,padded,K", [
    ("preset=ref", (152, 152, 152), 9),
    ("preset=c0", (140, 140, 140), None),
    ("preset=c1", (336, 336, 336), None),
    ("preset=c3", (272, 272, 272), None),
])
def test_presets_build(H, preset, padded, K):
    sim = H.Simulation(preset).build()
    d = sim.model_dims(0)
    assert tuple(d[9:12]) == padded
    m = sim.model(0)
    if K is not None:
        assert m["K"] == K
    # the padded velocity is the edge-replicated interior
    vi = m["v"][d[7]:d[7] + d[2], d[5]:d[5] + d[1], d[3]:d[3] + d[0]]
    assert np.array_equal(m["v"], np.pad(vi, ((d[7], d[8]), (d[5], d[6]), (d[3], d[4])), mode="edge"))
    assert abs(m["coef"][0] - orc.bessel_jn(0, m["R"] * 1e-3)) < 1e-14
    sim.close()


def test_reference_preset_parameters(H):
    sim = H.Simulation("preset=ref").build()
    d = sim.model_dims(0)
    assert d[:9] == [101, 101, 101, 25, 26, 25, 26, 25, 26]
    m = sim.model(0)
    assert np.abs(m["gx"] - orc.taper_ref(25, 101, 26)).max() < 1e-7
    assert abs(m["R"] - 4000 * np.pi * np.sqrt(0.03)) < 1e-9
    sh = sim.shot(0)
    assert sh["src"] == [75, 75, 25] and sh["nt"] == 300
    assert np.abs(sh["s"] - orc.ricker(300, 40.0, 1e-3)).max() < 1e-4
    sim.close()


def test_extent_rule_exact_and_smooth(H):
    """extent=exact (the default) keeps the odd rule's padded extent, any
    size up to 4096, e.g. the 150 of the reference's 100-cell file at 25-cell
    pads or 106 = 2 * 53 (no compiled split); extent=smooth widens an extent
    without compiled kernels to the next compiled one and records it."""
    import ose_amd
    sim = H.Simulation("nx=100 ny=101 nz=102 boundary=25 model=const nt=4").build()
    assert sim.model_dims(0)[9:12] == [150, 152, 152]
    sim.close()
    sim = H.Simulation("nx=56 ny=101 nz=102 boundary=25 model=const nt=4").build()
    assert sim.model_dims(0)[9:12] == [106, 152, 152]
    assert sim.info()["models"][0]["extent_adjusted"] == []
    sim.close()
    sim = H.Simulation("nx=56 ny=101 nz=102 boundary=25 model=const nt=4 extent=smooth").build()
    d = sim.model_dims(0)
    assert ose_amd.extent_kind(106) == 2
    assert d[9] == ose_amd.supported_extent(106) > 106 and d[4] == 25 + d[9] - 106 and d[10:12] == [152, 152]
    assert sim.info()["models"][0]["extent_adjusted"] == [{"axis": "x", "from": 106, "to": d[9]}]
    sim.close()
    with pytest.raises(Exception):
        H.Simulation("nx=4090 ny=8 nz=8 boundary=4 model=const nt=4").build()


def test_reference_input_fixture_is_the_file():
    """tests/golden/ref_inputs.npz holds the reference's velocity files byte
    for byte (their SHA-256 recorded when the fixture was made), and matches
    the files themselves where the reference is present."""
    import hashlib
    g = np.load(os.path.join(GOLD, "ref_inputs.npz"))
    for name in ("v3d_nx100_ny101_nz102", "3d_number-100-100-100-3"):
        raw = g[name].astype("<f4").tobytes()
        assert hashlib.sha256(raw).hexdigest() == str(g[name + "_sha256"])
        path = os.path.join(REF_INPUT, name + ".bin")
        if os.path.exists(path):
            assert open(path, "rb").read() == raw


def test_c2_batch_layout(H):
    sim = H.Simulation("preset=c2").build()
    assert sim.nmodels() == 5
    assert sim.nshots() == 20
    pads = [sim.model_dims(i)[9:12] for i in range(5)]
    assert pads[0] == [600, 600, 480]
    sim.close()


def test_boundary_parameter_file_formats(H, tmp_path):
    """bnet=<file>: .npy (float32, C order) and raw float32 parameter files."""
    rng = np.random.default_rng(9)
    p = rng.standard_normal(20784).astype(np.float32)
    np.save(tmp_path / "w.npy", p)
    p.tofile(tmp_path / "w.bin")
    assert np.array_equal(H.read_params(tmp_path / "w.npy"), p)
    assert np.array_equal(H.read_params(tmp_path / "w.bin"), p)
    np.save(tmp_path / "d.npy", p.astype(np.float64))
    with pytest.raises(Exception):
        H.read_params(tmp_path / "d.npy")
    from ose_amd import boundary
    w = H.read_params(boundary.WEIGHTS)
    assert w.size == 20784 and np.all(np.isfinite(w))


def test_segy_round_trip_and_headers(H, tmp_path):
    """SEG-Y rev 1: EBCDIC textual header, big-endian binary and trace
    headers, big-endian IEEE samples."""
    rng = np.random.default_rng(5)
    tr = rng.standard_normal((5, 43)).astype(np.float32)
    p = tmp_path / "x.segy"
    H.write_segy(p, tr, 2e-3, rec_x=np.arange(5) * 10.0, rec_y=np.full(5, 20.0), src=(15.0, 20.0), shot=1)
    got, dt = H.read_segy(p)
    assert np.array_equal(got, tr) and abs(dt - 2e-3) < 1e-12
    raw = open(p, "rb").read()
    assert len(raw) == 3600 + 5 * (240 + 4 * 43)
    assert raw[:3] == bytes([0xC3, 0x40, 0xF1])                       # "C 1" in EBCDIC
    assert struct.unpack_from(">h", raw, 3216)[0] == 2000             # sample interval (us)
    assert struct.unpack_from(">h", raw, 3220)[0] == 43               # samples per trace
    assert struct.unpack_from(">h", raw, 3224)[0] == 5                # IEEE float
    assert struct.unpack_from(">H", raw, 3500)[0] == 0x0100           # revision 1
    h = raw[3600 + 2 * (240 + 4 * 43):][:240]                          # third trace
    assert struct.unpack_from(">i", h, 0)[0] == 3                     # tracl
    assert struct.unpack_from(">i", h, 8)[0] == 2                     # fldr = shot + 1
    assert struct.unpack_from(">h", h, 70)[0] == -100                 # scalco
    assert struct.unpack_from(">i", h, 72)[0] == 1500                 # sx = 15 m * 100
    assert struct.unpack_from(">i", h, 80)[0] == 2000                 # gx = 20 m * 100
    assert struct.unpack_from(">i", h, 36)[0] == 5                    # offset (m)
    assert struct.unpack_from(">H", h, 114)[0] == 43
    samples = np.frombuffer(raw[3600 + 240:3600 + 240 + 4 * 43], dtype=">f4")
    assert np.array_equal(samples, tr[0])


def test_segy_reads_ibm_floats(H, tmp_path):
    """Format code 1 (IBM System/360 floats): known encodings."""
    vals = {0x41100000: 1.0, 0xC276A000: -118.625, 0x00000000: 0.0, 0x42640000: 100.0, 0x40800000: 0.5}
    head = bytearray(3600)
    struct.pack_into(">h", head, 3216, 1000)
    struct.pack_into(">h", head, 3220, len(vals))
    struct.pack_into(">h", head, 3224, 1)
    th = bytearray(240)
    struct.pack_into(">H", th, 114, len(vals))
    body = b"".join(struct.pack(">I", k) for k in vals)
    p = tmp_path / "ibm.segy"
    p.write_bytes(bytes(head) + bytes(th) + body)
    got, dt = H.read_segy(p)
    assert got.shape == (1, len(vals)) and abs(dt - 1e-3) < 1e-12
    assert np.allclose(got[0], list(vals.values()), rtol=0, atol=1e-6)


def _receivers(H, params):
    sim = H.Simulation(params).build()
    out = [sim.receivers(k) for k in range(sim.nshots())]
    sim.close()
    return out


def test_receiver_geometry(H):
    """Receiver line and grid from gxbeg..jgz / ng / ngx, ngy, csdgather
    (ose_3d.cu:100-114 and sg_init_device ose_3d_kernel.cu:202-212: receiver
    r at (gxbeg + r*jgx, gybeg + r*jgy, gzbeg + r*jgz))."""
    base = "nx=20 ny=16 nz=12 boundary=4 nt=4 model=const extent=smooth "
    # default: the interior plane rec_z, x fastest
    (plane,) = _receivers(H, base + "rec_z=3")
    assert plane.shape == (20 * 16, 3)
    assert np.array_equal(plane[21], [1, 1, 3]) and (plane[:, 2] == 3).all()
    # a line exactly as sg_init_device lays it out
    (line,) = _receivers(H, base + "ng=5 gxbeg=2 gybeg=3 gzbeg=1 jgx=3 jgy=2 jgz=2")
    r = np.arange(5)
    assert np.array_equal(line, np.stack([2 + 3 * r, 3 + 2 * r, 1 + 2 * r], axis=1))
    # defaults of a line: along x at the surface
    (line,) = _receivers(H, base + "ng=4 gxbeg=1")
    assert np.array_equal(line, [[1, 0, 0], [2, 0, 0], [3, 0, 0], [4, 0, 0]])
    # a grid at depth gzbeg, x fastest
    (grid,) = _receivers(H, base + "ngx=3 ngy=2 gxbeg=1 gybeg=2 gzbeg=4 jgx=2 jgy=5")
    assert np.array_equal(grid, [[1, 2, 4], [3, 2, 4], [5, 2, 4], [1, 7, 4], [3, 7, 4], [5, 7, 4]])
    # three shots along x; fixed spread, then the spread moving with the source
    shots = base + "nshot=3 sx=4 sy=8 jsx=3 jsy=0 jsz=0 ng=4 gxbeg=2 gybeg=8 "
    fixed = _receivers(H, shots)
    assert all(np.array_equal(f, fixed[0]) for f in fixed)
    moving = _receivers(H, shots + "csdgather=1")
    for k, m in enumerate(moving):
        assert np.array_equal(m, fixed[0] + [3 * k, 0, 0])


@pytest.mark.parametrize("bad", ["ng=30", "ng=3 gzbeg=12", "ngx=2 gxbeg=-1", "ng=2 ngx=2", "ng=-1",
                                 "nshot=3 sx=4 jsx=6 ng=4 gxbeg=10 csdgather=1"])
def test_receiver_geometry_rejects_cells_outside(H, bad):
    with pytest.raises(H.OseError):
        H.Simulation("nx=20 ny=16 nz=12 boundary=4 nt=4 model=const extent=smooth " + bad).build()


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_snapshot_codec_and_osnap_files(H, tmp_path, fmt):
    """The C decoder and lane extraction against the numpy restatement of
    the encodings; .osnap round trip; decode error within the format's
    bound (bf16: 2^-8 relative; Q16/Q8: half a step of the line's scale)."""
    import _snapcodec as sc
    rng = np.random.default_rng(30 + fmt)
    nz, ny, nx = 5, 4, 37
    p = (rng.standard_normal((nz, ny, nx)) * np.logspace(-3, 3, nx)).astype(np.float32)
    q = (rng.standard_normal((nz, ny, nx)) * 50).astype(np.float32)
    q[1, 2] = 0.0                                   # an all-zero line: scale 0
    data = sc.encode(fmt, [p, q])
    lb = H._L().oseh_snapshot_lane_bytes(fmt, nx, ny, nz)
    assert len(data) == 2 * lb
    for lane, v in enumerate((p, q)):
        raw = H.snapshot_lane(fmt, 2, lane, nx, ny, nz, data)
        assert raw == sc.encode(fmt, [v])
        dec = H.decode_snapshot(fmt, nx, ny, nz, raw)
        s, vals = sc.encode_lane(fmt, v)
        assert np.array_equal(dec, sc.decode_lane(fmt, (nz, ny, nx), s, vals))
        assert (np.abs(dec - v) <= sc.bound(fmt, v)).all()
        path = tmp_path / f"l{lane}.osnap"
        box = (ctypes.c_int * 6)(2, 3, 4, nx, ny, nz)
        assert H._L().oseh_write_osnap(str(path).encode(), fmt, lane, 17, 1, box, raw) == 0
        got, meta = H.read_osnap(path)
        assert np.array_equal(got, dec)
        assert meta == dict(format=fmt, lane=lane, step=17, shot=1, x0=2, y0=3, z0=4, nx=nx, ny=ny, nz=nz)
        assert len(open(path, "rb").read()) == 64 + lb
    assert (H.decode_snapshot(fmt, nx, ny, nz, H.snapshot_lane(fmt, 2, 1, nx, ny, nz, data))[1, 2] == 0).all()
    # a planted mistake in the restatement (scales of the wrong lane) is caught
    if fmt >= 2:
        bad = sc.encode(fmt, [q])[:4 * ny * nz] + sc.encode(fmt, [p])[4 * ny * nz:]
        good = sc.encode(fmt, [p])
        assert not np.array_equal(H.decode_snapshot(fmt, nx, ny, nz, bad), H.decode_snapshot(fmt, nx, ny, nz, good))


def test_gaussian_source_profile(H):
    """Normalised, symmetric, +-ceil(4 sigma) wide, clipped at the grid edge;
    sigma 0 is the unit point."""
    a, p = H.gaussian_profile(2.5, 30, 64)
    assert a == 20 and p.size == 21
    assert abs(float(p.sum(dtype=np.float64)) - 1.0) < 1e-6
    assert np.allclose(p, p[::-1]) and p.argmax() == 10
    assert np.isclose(p[10 + 3] / p[10], np.exp(-0.5 * (3 / 2.5) ** 2), rtol=1e-5)
    a, p = H.gaussian_profile(2.5, 3, 64)
    assert a == 0 and p.size == 14 and abs(float(p.sum(dtype=np.float64)) - 1.0) < 1e-6
    a, p = H.gaussian_profile(0.0, 7, 16)
    assert a == 7 and np.array_equal(p, [1.0])
    with pytest.raises(H.OseError):
        H.gaussian_profile(9.0, 40, 128)      # 73 cells > 64


def test_osnap_reader_rejects_bad_files(H, tmp_path):
    import _snapcodec as sc
    v = np.arange(2 * 3 * 16, dtype=np.float32).reshape(2, 3, 16)
    raw = sc.encode(sc.Q8, [v])
    good = tmp_path / "g.osnap"
    box = (ctypes.c_int * 6)(0, 0, 0, 16, 3, 2)
    assert H._L().oseh_write_osnap(str(good).encode(), sc.Q8, 0, 1, 0, box, raw) == 0
    data = open(good, "rb").read()
    for name, blob in (("magic", b"XSESNAP1" + data[8:]), ("short", data[:-1]), ("long", data + b"\0"),
                       ("size", data[:48] + (123).to_bytes(8, "little") + data[56:])):
        p = tmp_path / f"{name}.osnap"
        open(p, "wb").write(blob)
        with pytest.raises(H.OseError):
            H.read_osnap(p)
    got, meta = H.read_osnap(good)
    assert got.shape == (2, 3, 16) and meta["format"] == sc.Q8
