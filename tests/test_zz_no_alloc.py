"""The entry point that allocates nothing (ose_apply_operator_noalloc)
against ose_apply_operator through the C-ABI: the same operator application
(memcpy_between_device_cr -> FFT -> scale_kxkz_3D -> IFFT -> scal_vr_3d,
ose_3d_lib.cu:649-687) on its static scratch instead of freshly allocated
fields, bit for bit, on two calls in a row (the second on a smaller model,
so the scratch holds the first call's larger contents), and models beyond
the scratch limit rejected."""
import numpy as np
import pytest

from _cases import Case, layered

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def test_noalloc_operator_equals_allocating_operator():
    import ose_amd
    from ose_amd import _lib
    print("no-alloc operator: 152^3 then 64 x 48 x 80", flush=True)
    big = Case((140, 140, 140), (6, 6, 6, 6, 6, 6), layered(140, 140, 140, seed=3))
    small = Case((52, 36, 68), (6, 6, 6, 6, 6, 6), layered(52, 36, 68, seed=4))
    with ose_amd.Context(0) as c:
        mb = big.add_to(c)
        ms = small.add_to(c)
        ub = _field(big.shape, 1)
        us = _field(small.shape, 2)
        ref_b = c.apply_operator(mb, ub)
        ref_s = c.apply_operator(ms, us)
        got_b = c.apply_operator_noalloc(mb, ub)
        got_s = c.apply_operator_noalloc(ms, us)
        # and once more on the large model after the small one
        got_b2 = c.apply_operator_noalloc(mb, ub)
    assert np.abs(ref_b).max() > 0 and np.abs(ref_s).max() > 0
    assert np.array_equal(got_b.view(np.uint32), ref_b.view(np.uint32))
    assert np.array_equal(got_s.view(np.uint32), ref_s.view(np.uint32))
    assert np.array_equal(got_b2.view(np.uint32), ref_b.view(np.uint32))


def test_noalloc_operator_rejects_models_beyond_the_scratch():
    import ose_amd
    from ose_amd import _lib
    print("no-alloc operator: 192^3 rejected", flush=True)
    case = Case((180, 180, 180), (6, 6, 6, 6, 6, 6), np.full((180, 180, 180), 2000.0, np.float32))
    assert case.nxb * case.nyb * case.nzb > 1 << 22
    with ose_amd.Context(0) as c:
        m = case.add_to(c)
        u = np.zeros(case.shape, dtype=np.complex64)
        with pytest.raises(_lib.OseError) as e:
            c.apply_operator_noalloc(m, u)
        assert e.value.code == _lib.OSE_EINVAL
        assert "scratch limit" in str(e.value)
