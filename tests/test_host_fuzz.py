"""Mutation check of the host library's file readers (CPU): files written
by the library's own writers, then truncated, byte-flipped and with header
fields overwritten; every read must return a status (never crash) and an
accepted read must fit the caller's buffer.  The same mutations ran 20000
times under AddressSanitizer with no report."""
import ctypes
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def L():
    from ose_amd import _lib
    return _lib.host()


def _files(L, d):
    from ose_amd import host as H
    tr = (np.arange(7 * 33, dtype=np.float32).reshape(7, 33) * 0.25) - 3
    H.write_su(os.path.join(d, "a.su"), tr, 1e-3)
    H.write_segy(os.path.join(d, "a.segy"), tr, 1e-3)
    box = (ctypes.c_int * 6)(1, 2, 3, 5, 4, 3)
    for fmt in range(4):
        n = L.oseh_snapshot_lane_bytes(fmt, 5, 4, 3)
        lane = (ctypes.c_ubyte * max(1, n))(*([7] * max(1, n)))
        assert L.oseh_write_osnap(os.path.join(d, f"f{fmt}.osnap").encode(), fmt, 0, 3, 0, box, lane) == 0
    np.save(os.path.join(d, "p.npy"), np.arange(10, dtype=np.float32))
    np.full(4 * 5 * 6, 1500.0, dtype=np.float32).tofile(os.path.join(d, "v.bin"))
    return ["a.su", "a.segy", "f0.osnap", "f1.osnap", "f2.osnap", "f3.osnap", "p.npy", "v.bin"]


def test_readers_survive_corrupt_files(L, tmp_path):
    d = str(tmp_path)
    names = _files(L, d)
    rng = np.random.default_rng(2024)
    out = np.zeros(1 << 14, dtype=np.float32)
    fp = out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    for it in range(1600):
        name = names[it % len(names)]
        b = bytearray(open(os.path.join(d, name), "rb").read())
        kind = it % 4
        if kind == 0 and b:
            del b[int(rng.integers(0, len(b))):]
        else:
            for _ in range(int(rng.integers(1, 9))):
                if b:
                    b[int(rng.integers(0, min(len(b), 3700)))] = int(rng.integers(0, 256))
            if kind == 3 and len(b) >= 8:
                pos = 4 * int(rng.integers(0, min(len(b), 3600) // 4))
                if pos + 4 <= len(b):
                    b[pos:pos + 4] = int(rng.integers(-2 ** 31, 2 ** 31)).to_bytes(4, "little", signed=True)
        p = os.path.join(d, "mut")
        with open(p, "wb") as f:
            f.write(b)
        cap = out.size >> int(rng.integers(0, 12))
        a, c, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        dt = ctypes.c_double()
        meta = (ctypes.c_int * 10)()
        if name == "a.su" or name == "a.segy":
            f = L.oseh_read_su if name == "a.su" else L.oseh_read_segy
            rc = f(p.encode(), fp, cap, ctypes.byref(a), ctypes.byref(c), ctypes.byref(dt))
            if rc == 0:
                assert a.value * c.value <= cap
        elif name == "p.npy":
            rc = L.oseh_read_params(p.encode(), fp, cap, ctypes.byref(n))
            if rc == 0:
                assert n.value <= cap
        elif name == "v.bin":
            rc = L.oseh_read_velocity_xyz(p.encode(), 4, 5, 6, fp)
        else:
            rc = L.oseh_read_osnap(p.encode(), fp if it % 2 else None, cap, meta)
            if rc == 0 and it % 2:
                assert meta[7] * meta[8] * meta[9] <= cap
        assert isinstance(rc, int)


def test_trace_files_cut_inside_a_header_are_rejected(L, tmp_path):
    """An SU or SEG-Y file that ends inside a trace header is cut short: an
    error, like one that ends inside a trace's samples; complete files read
    back whole."""
    from ose_amd import host as H
    tr = np.arange(7 * 33, dtype=np.float32).reshape(7, 33)
    out = np.zeros(1 << 12, dtype=np.float32)
    fp = out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    for write, read, head in ((H.write_su, L.oseh_read_su, 0), (H.write_segy, L.oseh_read_segy, 3600)):
        p = str(tmp_path / "t")
        write(p, tr, 1e-3)
        b = open(p, "rb").read()
        trace = 240 + 4 * 33
        for cut, ok, ntr in ((len(b), True, 7), (head + trace, True, 1), (head + 100, False, 0),
                             (head + trace + 239, False, 1), (head + 2 * trace - 5, False, 1)):
            with open(p, "wb") as f:
                f.write(b[:cut])
            a, c, dt = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
            rc = read(p.encode(), fp, out.size, ctypes.byref(a), ctypes.byref(c), ctypes.byref(dt))
            assert (rc == 0) == ok, (write.__name__, cut, rc)
            assert a.value == ntr


@pytest.mark.parametrize("params", [
    "nx=-5", "nz=0", "boundary=-3", "nx=5000", "dt=0", "dt=-1e-3", "dx=nan", "dz=inf", "dt=abc", "nt=0",
    "nt=12x", "nshot=100000000", "nshot=0", "srcgrid=100000", "srcgrid=0", "korder=ref dt=1e9", "ktol=0",
    "seed=99999999999", "model=const v0=0", "model=const v0=-100",
])
def test_bad_parameters_are_rejected(params):
    """Values that are not numbers, out-of-range sizes, shot counts, steps
    and spacings end in an error from the host's build, never in a crash or
    an unbounded allocation (found by random parameter strings under
    AddressSanitizer and UndefinedBehaviorSanitizer)."""
    from ose_amd.host import Simulation
    sim = Simulation(f"nx=12 ny=10 nz=8 nt=6 boundary=4 {params}")
    try:
        with pytest.raises(Exception):
            sim.build()
    finally:
        sim.close()


def test_extreme_delay_builds():
    """delay is any int: the wavelet's (i - delay) is formed in double."""
    from ose_amd.host import Simulation
    sim = Simulation("nx=12 ny=10 nz=8 nt=6 boundary=4 delay=-2147483648").build()
    sim.close()
