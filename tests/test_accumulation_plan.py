"""The step's deferred accumulator schedule (ose_accumulation_plan, host
only) forms the reference's series sum.

wavefield_extro_ori2 (ose_3d_lib.cu:741-859) adds c_0 W_0 + c_1 W_1 at order
0, then on every order j = 1..K-1 adds c_{j+1} W_{j+1} and multiplies the
accumulator by the taper g (apply_ABC_Xt).  The HIP step instead updates the
accumulator only on the orders the plan marks, adding the terms that pass
holds (W_{j-1} read, W_{j+1} written) with powers of g.  Replaying the plan
on random scalars must give the reference's sum to FP64 rounding for every
K the API accepts, with every term added exactly once by a pass that holds
it, and with fewer accumulator passes than orders."""
import ctypes

import numpy as np
import pytest

from ose_amd import _lib


def plan(K):
    out = (ctypes.c_int * (6 * K))()
    assert _lib.hip().ose_accumulation_plan(K, out) == _lib.OSE_OK
    return np.array(out[:], dtype=int).reshape(K, 6)


def reference_sum(K, g, c, W):
    # ose_3d_lib.cu:741-743 (order 0), 845-858 (orders 1..M, M = K-1)
    A = c[0] * W[0] + c[1] * W[1]
    for j in range(1, K):
        A = g * (A + c[j + 1] * W[j + 1])
    return A


def replay(K, p, g, c, W):
    A = None
    used = []
    for j in range(K):
        touch, add_prev, add_next, ea, eb, ec = p[j]
        if not touch:
            assert not add_prev and not add_next
            continue
        prev_m = 0 if j == 0 else j - 1      # pass 0 holds W_0 (it reads u)
        new = 0.0
        if add_prev:
            new += g ** eb * c[prev_m] * W[prev_m]
            used.append(prev_m)
        if add_next:
            new += g ** ec * c[j + 1] * W[j + 1]
            used.append(j + 1)
        if ea < 0:
            assert A is None
            A = new
        else:
            assert A is not None
            A = g ** ea * A + new
    return A, sorted(used)


@pytest.mark.parametrize("K", list(range(2, 65)))
def test_plan_reproduces_reference_sum(K):
    p = plan(K)
    assert p[K - 1, 0] == 1, "the last order forms u"
    assert p[K - 2, 0] == 1, "W_{K-1} is held only by pass K-2"
    assert (p[:, 3:][p[:, 0] == 1] <= 3).all(), "kernels raise g to powers 0..3"
    rng = np.random.default_rng(K)
    for _ in range(4):
        g = rng.uniform(0.3, 1.0)
        c = rng.standard_normal(K + 1)
        W = rng.standard_normal(K + 1)
        got, used = replay(K, p, g, c, W)
        assert used == list(range(K + 1)), "every term added exactly once"
        ref = reference_sum(K, g, c, W)
        assert abs(got - ref) <= 1e-12 * (1 + abs(ref)), (K, got, ref)


def test_plan_touches_about_half_the_orders():
    # K = 11 (the 768^3 workload): passes 1 2 5 6 9 10 update the accumulator
    p = plan(11)
    assert list(np.nonzero(p[:, 0])[0]) == [1, 2, 5, 6, 9, 10]
    for K in range(6, 65):
        assert plan(K)[:, 0].sum() <= K // 2 + 2


def test_planted_mistake_is_caught():
    """A plan whose exponent is off by one (an extra taper on one term) no
    longer reproduces the reference sum."""
    K = 11
    p = plan(K).copy()
    j = 5
    assert p[j, 0] == 1
    p[j, 4] += 1
    g, c, W = 0.7, np.linspace(1, 2, K + 1), np.linspace(-1, 1, K + 1) + 0.3
    got, _ = replay(K, p, g, c, W)
    assert abs(got - reference_sum(K, g, c, W)) > 1e-6


def test_rejects_out_of_range_K():
    out = (ctypes.c_int * 6)()
    assert _lib.hip().ose_accumulation_plan(1, out) == _lib.OSE_EINVAL
    assert _lib.hip().ose_accumulation_plan(65, out) == _lib.OSE_EINVAL
    assert _lib.hip().ose_accumulation_plan(11, None) == _lib.OSE_EINVAL
