"""Pins of the FP64 numpy reference of the boundary network (CPU only):
finite-difference gradients, the conv adjoint identity, and the shell sets."""
import numpy as np

import bnet_ref as B


def _case(seed=0):
    rng = np.random.default_rng(seed)
    geo = B.Geometry((12, 14, 16), (4, 3, 2), (8, 9, 7))   # shape (nz,ny,nx), origin (bl,bf,bu), interior (nx,ny,nz)
    u = rng.standard_normal(geo.pad.shape) + 1j * rng.standard_normal(geo.pad.shape)
    v = rng.uniform(1500, 4000, geo.pad.shape)
    t = u + 0.3 * (rng.standard_normal(u.shape) + 1j * rng.standard_normal(u.shape))
    p = 0.2 * rng.standard_normal(B.NPAR)
    return geo, u, v, t, p


def test_shell_sets():
    geo, *_ = _case()
    assert geo.pad.sum() == 12 * 14 * 16 - 8 * 9 * 7
    # R_d is the pad dilated by d in the Chebyshev norm on the torus
    assert geo.R[3].sum() == 12 * 14 * 16 - (8 - 6) * (9 - 6) * (7 - 6)
    assert geo.R[1].sum() == 12 * 14 * 16 - 6 * 7 * 5


def test_conv_adjoint():
    rng = np.random.default_rng(1)
    h = rng.standard_normal((5, 6, 7, 16))
    g = rng.standard_normal((5, 6, 7, 16))
    W = rng.standard_normal((27, 16, 16))
    assert abs(np.sum(B.conv(h, W) * g) - np.sum(h * B.conv_T(g, W))) < 1e-9 * np.abs(np.sum(h * B.conv_T(g, W)))
    dW = B.conv_wgrad(h, g)
    E = rng.standard_normal(W.shape)
    assert abs(np.sum(B.conv(h, E) * g) - np.sum(dW * E)) < 1e-9 * abs(np.sum(dW * E))


def test_gradient_matches_finite_differences():
    geo, u, v, t, p = _case(2)
    L, g = B.loss_and_grad(geo, p, u, v, 4000.0, t)
    rng = np.random.default_rng(3)
    idx = rng.choice(B.NPAR, 40, replace=False)
    # include biases of every layer
    idx = np.concatenate([idx, [27 * 256 + 3, 2 * (27 * 256 + 16) - 5, B.NPAR - 16, B.NPAR - 15]])
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (B.loss_and_grad(geo, pp, u, v, 4000.0, t)[0] - B.loss_and_grad(geo, pm, u, v, 4000.0, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_zero_last_layer_is_the_plain_pad():
    geo, u, v, t, p = _case(4)
    p[-(27 * 256 + 16):] = 0.0
    delta, _ = B.forward(geo, p, u, v, 4000.0)
    assert np.all(delta == 0)
