"""Error paths of the C-ABI on a real device: every bad input is turned into
an OSE_* status with a message (nothing exits), and a context that failed
can still be used or destroyed (the reference's checkCudaErrors/exit(0)
paths, ose_3d_lib.cu:18-34, 891-895)."""
import numpy as np
import pytest

from _cases import Case, hilbert_unnorm, layered, ricker

pytestmark = pytest.mark.gpu


def _err(fn):
    import ose_amd
    with pytest.raises(ose_amd.OseError) as e:
        fn()
    return e.value.code, str(e.value)


def _case():
    return Case((36, 20, 36), (6, 6, 6, 6, 6, 6), layered(36, 20, 36, seed=21))


def test_model_arguments_are_checked():
    import ose_amd
    from ose_amd._lib import OSE_EINVAL, OSE_ENUMERIC, OSE_EUNSUPPORTED
    print("model argument errors", flush=True)
    case = _case()
    with ose_amd.Context(0) as c:
        kw = dict(interior=(case.nx, case.ny, case.nz), origin=(case.bl, case.bf, case.bu), d=case.d, dt=case.dt,
                  R=case.R, coef=case.coef, taper=(case.gx, case.gy, case.gz))
        v = case.v.copy()
        v[3, 4, 5] = np.nan
        code, msg = _err(lambda: c.add_model(v, **kw))
        assert code == OSE_ENUMERIC and "velocity" in msg
        v = case.v.copy()
        v[0, 0, 0] = -100.0
        assert _err(lambda: c.add_model(v, **kw))[0] == OSE_ENUMERIC
        bad = dict(kw, coef=np.r_[case.coef, np.inf])
        assert _err(lambda: c.add_model(case.v, **bad))[0] == OSE_ENUMERIC
        bad = dict(kw, coef=case.coef[:2])                       # K = 1
        assert _err(lambda: c.add_model(case.v, **bad))[0] == OSE_EINVAL
        bad = dict(kw, R=-1.0)
        assert _err(lambda: c.add_model(case.v, **bad))[0] == OSE_EINVAL
        assert _err(lambda: c.add_model(case.v, rec_z=case.nz, **kw))[0] == OSE_EINVAL
        bad = dict(kw, interior=(case.nxb, case.ny, case.nz))          # origin + interior beyond the grid
        assert _err(lambda: c.add_model(case.v, **bad))[0] == OSE_EINVAL
        # every extent in [2, 4096] runs (17: the runtime-length path); 4098 does not
        vbig = np.full((48, 32, 4098), 2000.0, dtype=np.float32)
        kwbig = dict(kw, interior=(5, case.ny, case.nz), taper=(np.ones(4098), case.gy, case.gz))
        code, msg = _err(lambda: c.add_model(vbig, **kwbig))
        assert code == OSE_EUNSUPPORTED and "outside [2, 4096]" in msg
        # the context is still usable after the failures
        m = case.add_to(c)
        s = ricker(10, 30.0, case.dt)
        c.add_shot(m, (case.bl + 5, case.bf + 5, case.bu + 5), s, hilbert_unnorm(s))
        c.run(2)
        assert c.step == 2


def test_shot_and_state_errors():
    import ose_amd
    from ose_amd._lib import OSE_EINVAL, OSE_ENUMERIC, OSE_ESTATE, OSE_EUNSUPPORTED
    print("shot and state errors", flush=True)
    case = _case()
    s = ricker(20, 30.0, case.dt)
    h = hilbert_unnorm(s)
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        assert _err(lambda: c.run(1))[0] == OSE_ESTATE                    # no shots yet
        assert _err(lambda: c.add_shot(m, (case.nxb, 0, 0), s, h))[0] == OSE_EINVAL
        assert _err(lambda: c.add_shot(m + 7, (1, 1, 1), s, h))[0] == OSE_EINVAL
        sn = s.copy()
        sn[3] = np.inf
        assert _err(lambda: c.add_shot(m, (10, 10, 10), sn, h))[0] == OSE_ENUMERIC
        sid = c.add_shot(m, (case.bl + 18, case.bf + 10, case.bu + 2), s, h)
        c.run(5)
        assert _err(lambda: case.add_to(c))[0] == OSE_ESTATE                # models are fixed once running
        assert _err(lambda: c.add_shot(m, (10, 10, 10), s, h))[0] == OSE_ESTATE
        assert _err(lambda: c.traces(sid, 15, 10))[0] == OSE_EINVAL         # beyond nt
        re, im = c.wavefield(sid)
        re[1, 2, 3] = np.nan
        assert _err(lambda: c.set_wavefield(sid, re, im))[0] == OSE_ENUMERIC
        assert _err(lambda: c.fft3d(np.zeros((16, 16, 1), np.complex64)))[0] == OSE_EUNSUPPORTED   # extent 1
        assert _err(lambda: c.set_streams(3))[0] == OSE_EINVAL
        c.set_streams(1)
        c.run(1)
        tr = c.traces(sid, 0, 6)
        assert np.all(np.isfinite(tr))


def test_oversize_batch_reports_enomem():
    """16 shots of an 864^3 grid need 330 GB of fields: the allocation fails
    with OSE_ENOMEM on a 288 GB device and the context is still destroyed
    cleanly."""
    import ose_amd
    from ose_amd._lib import OSE_ENOMEM
    print("oversize allocation", flush=True)
    n = 864
    v = np.full((n, n, n), 3000.0, dtype=np.float32)
    one = np.ones(n, dtype=np.float32)
    with ose_amd.Context(0) as c:
        m = c.add_model(v, interior=(n - 80, n - 80, n - 80), origin=(40, 40, 40), d=(10.0, 10.0, 10.0), dt=1e-3,
                        R=3000.0 * np.pi * np.sqrt(0.03), coef=np.array([0.1, 0.2, 0.3]), taper=(one, one, one))
        del v
        s = np.zeros(4, dtype=np.float32)
        for k in range(16):
            c.add_shot(m, (100 + k, 100, 100), s, s)
        code, msg = _err(lambda: c.run(1))
        assert code == OSE_ENOMEM, msg
