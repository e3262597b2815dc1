"""Closed-form checks of the OSE step on sums of Fourier modes (checker side).

With a constant velocity v and no taper the operator L = (v/R)|k| is
diagonal in the Fourier basis, so one step maps a grid mode
e_m(x) = exp(2 pi i (mx x/nx + my y/ny + mz z/nz)) to
exp(-i dt v |k_m|) e_m, up to the series tail 2 sum_{k>K} |J_k(R dt)|
(Jacobi-Anger; ose_3d_lib.cu:639-874 sums the same series).  |k_m| uses the
folded wavenumbers of scale_kxkz_3D (ose_3d_kernel.cu:101-126).  Fields are
built and projected slab by slab along z, so grids of 864^3 and more need
no full-size temporary beyond the float32 lanes themselves.
"""
import numpy as np


def folded_k(m, n, d):
    m = m % n
    return 2.0 * np.pi / (d * n) * min(m, n - m)


def omega(mode, shape, d, v):
    """v |k| of a mode (mx, my, mz) on a (nz, ny, nx) grid with spacing d."""
    nz, ny, nx = shape
    kx = folded_k(mode[0], nx, d[0])
    ky = folded_k(mode[1], ny, d[1])
    kz = folded_k(mode[2], nz, d[2])
    return v * np.sqrt(kx * kx + ky * ky + kz * kz)


def _axis(m, n):
    return np.exp(2j * np.pi * (m % n) * np.arange(n) / n)


def _axes(modes, shape, conj=False):
    nz, ny, nx = shape
    ex = np.stack([_axis(m[0], nx) for m in modes], axis=1)   # (nx, M)
    ey = np.stack([_axis(m[1], ny) for m in modes], axis=1)
    ez = np.stack([_axis(m[2], nz) for m in modes], axis=1)
    return (np.conj(ex), np.conj(ey), np.conj(ez)) if conj else (ex, ey, ez)


def modes_field(shape, modes, amps, chunk=32):
    """(re, im) float32 lanes of sum_m a_m e_m: each z plane is the rank-M
    product (a_m ez_m[z] ey_m) (ex_m)^T."""
    nz, ny, nx = shape
    re = np.empty(shape, dtype=np.float32)
    im = np.empty(shape, dtype=np.float32)
    ex, ey, ez = _axes(modes, shape)
    a = np.asarray(amps, dtype=np.complex128)
    for z0 in range(0, nz, chunk):
        z1 = min(nz, z0 + chunk)
        left = (ez[z0:z1, None, :] * ey[None, :, :] * a).astype(np.complex64)   # (zc, ny, M)
        u = left.reshape(-1, len(modes)) @ ex.T.astype(np.complex64)              # (zc*ny, nx)
        re[z0:z1] = u.real.reshape(z1 - z0, ny, nx)
        im[z0:z1] = u.imag.reshape(z1 - z0, ny, nx)
    return re, im


def project(re, im, modes, chunk=32):
    """Coefficients c_m = <e_m, u> / N of u = re + i im and the relative
    L2 norm of what lies outside the listed modes (the modes are
    orthogonal, so ||u||^2 = N sum |c_m|^2 + ||rest||^2).  FP64 sums."""
    nz, ny, nx = re.shape
    n = float(nz) * ny * nx
    ex, ey, ez = _axes(modes, re.shape, conj=True)
    exr, exi = np.ascontiguousarray(ex.real), np.ascontiguousarray(ex.imag)
    c = np.zeros(len(modes), dtype=np.complex128)
    norm2 = 0.0
    for z0 in range(0, nz, chunk):
        z1 = min(nz, z0 + chunk)
        r = re[z0:z1].astype(np.float64).reshape(-1, nx)
        i = im[z0:z1].astype(np.float64).reshape(-1, nx)
        norm2 += float(np.dot(r.ravel(), r.ravel()) + np.dot(i.ravel(), i.ravel()))
        ux = (r @ exr - i @ exi) + 1j * (r @ exi + i @ exr)          # (zc*ny, M): sum over x
        uy = np.einsum("zym,ym->zm", ux.reshape(z1 - z0, ny, -1), ey)   # sum over y
        c += np.einsum("zm,zm->m", uy, ez[z0:z1])
    c /= n
    rest2 = max(0.0, norm2 - n * float(np.sum(np.abs(c) ** 2)))
    return c, float(np.sqrt(rest2 / norm2)), norm2


def check_modes(re, im, modes, amps, nsteps, dt, shape, d, v, tail, z):
    """Compare a stepped field with the closed form.  Returns a dict of the
    measured errors and their bounds:
      amp    max_m ||c_m / a_m| - 1|: the series tail per step plus the FP32
             rounding of the coefficients (about 2e-7 per step), plus 2e-5;
      phase  max_m |arg(c_m / (a_m exp(-i n dt v |k_m|)))|: the FP32 rounding
             of the symbol (v/R)|k| (a few ulp, times R dt per step) plus the
             tail, plus 2e-5;
      rest   relative L2 norm outside the modes: FFT rounding noise (no
             leakage into other modes, which a wrong index or stride gives
             at O(1)), bounded by 1e-4 + 2e-6 sqrt(n)."""
    c, rest, _ = project(re, im, modes)
    amp_err, ph_err = 0.0, 0.0
    for cm, a, m in zip(c, amps, modes):
        ratio = cm / (a * np.exp(-1j * nsteps * dt * omega(m, shape, d, v)))
        amp_err = max(amp_err, abs(abs(ratio) - 1.0))
        ph_err = max(ph_err, abs(np.angle(ratio)))
    return dict(amp=amp_err, amp_bound=nsteps * (tail + 2e-7) + 2e-5,
                phase=ph_err, phase_bound=nsteps * (tail + 6e-7 * z) + 2e-5,
                rest=rest, rest_bound=1e-4 + 2e-6 * np.sqrt(nsteps))


D = (10.0, 10.0, 10.0)
DT = 1e-3


def add_sim_model(ctx, sim, mi, v=None):
    """Model mi of a host Simulation's geometry added to a Context (v in
    place of the preset's velocity when given).  Returns (id, model info)."""
    m = sim.model(mi)
    nx, ny, nz, bl, br, bf, bb, bu, bd, nxb, nyb, nzb = m["dims"]
    return ctx.add_model(m["v"] if v is None else v, interior=(nx, ny, nz), origin=(bl, bf, bu), d=D, dt=DT,
                         R=m["R"], coef=m["coef"], taper=(m["gx"], m["gy"], m["gz"]), rec_z=m["rec_z"]), m


def series_tail(z, K):
    """2 sum_{k>K} |J_k(z)|: what the truncated series drops per step."""
    import scipy.special as sp
    return 2.0 * sum(abs(sp.jv(k, z)) for k in range(K + 1, K + 40))


def shot_modes(k, shape):
    """Four modes per shot, different for every shot: a low one, two high
    ones and the Nyquist corner (the top of the operator's spectrum)."""
    nz, ny, nx = shape
    modes = [(3 + k, 5 + 2 * k, 7 + k),
             (nx // 3 + 13 * k, ny // 2 - 7 * k - 1, nz // 5 + 11 * k),
             (nx - 40 - 9 * k, 31 + 17 * k, nz // 2 - 3 - 5 * k),
             (nx // 2, ny // 2, nz // 2)]
    amps = [1.0, 0.5 - 0.25j, 0.3j + 0.05 * k, 0.2 * (-1) ** k]
    return modes, amps


def const_batch(preset, v, nsteps, lanes=None, extra="", chunk=0, log=print):
    """Step every shot of a preset's geometry on a constant velocity with no
    taper, each from its own sum of modes (shot_modes), and check each shot
    against the closed form.  chunk > 0 runs the steps in chunks of that
    many with a progress line each.  Returns one check_modes dict per shot,
    with the shot's index, padded extents and K added."""
    import ose_amd
    from ose_amd.host import Simulation
    sim = Simulation(f"preset={preset} model=const v0={v} taper=none amp=0 nt={nsteps} {extra}").build()
    results = []
    try:
        with ose_amd.Context(0) as c:
            if lanes:
                c.set_streams(lanes)
            shots = []
            models = []
            for mi in range(sim.nmodels()):
                m = sim.model(mi)
                nxb, nyb, nzb = m["dims"][9:12]
                vc = np.full((nzb, nyb, nxb), v, dtype=np.float32)
                mid, mm = add_sim_model(c, sim, mi, vc)
                assert np.all(mm["gx"] == 1) and np.all(mm["gz"] == 1)
                models.append((mid, mm))
            for k in range(sim.nshots()):
                sh = sim.shot(k)
                mid, mm = models[sh["model"]]
                shots.append((c.add_shot(mid, tuple(sh["src"]), sh["s"], sh["h"]), mm, k))
            for sid, mm, k in shots:
                shape = tuple(mm["v"].shape)
                modes, amps = shot_modes(k, shape)
                re, im = modes_field(shape, modes, amps)
                c.set_wavefield(sid, re, im)
                del re, im
            done = 0
            while done < nsteps:
                n = nsteps - done if chunk <= 0 else min(chunk, nsteps - done)
                c.run(n)
                c.sync()
                done += n
                if chunk > 0:
                    log(f"  {done}/{nsteps} steps")
            for sid, mm, k in shots:
                shape = tuple(mm["v"].shape)
                modes, amps = shot_modes(k, shape)
                re, im = c.wavefield(sid)
                assert np.isfinite(re).all() and np.isfinite(im).all()
                z = mm["R"] * DT
                r = check_modes(re, im, modes, amps, nsteps, DT, shape, D, v, series_tail(z, mm["K"]), z)
                del re, im
                log(f"  shot {k} {shape[::-1]} K={mm['K']}: amp {r['amp']:.2e} (bound {r['amp_bound']:.2e}) "
                    f"phase {r['phase']:.2e} (bound {r['phase_bound']:.2e}) rest {r['rest']:.2e} "
                    f"(bound {r['rest_bound']:.2e})")
                results.append(dict(r, shot=k, padded=list(shape[::-1]), K=int(mm["K"])))
    finally:
        sim.close()
    return results


def within(r):
    return bool(r["amp"] < r["amp_bound"] and r["phase"] < r["phase_bound"] and r["rest"] < r["rest_bound"])


def weighted_energy(re, im, v, chunk=32):
    """sum |u|^2 / v over the grid, FP64, slab by slab.  The OSE operator
    L = v |k| (v multiplying after |k|) is self-adjoint in the inner
    product <a, b> = sum conj(a) b / v, so exp(-i dt L) conserves this norm
    (up to the series tail) on any velocity model; sum |u|^2 is conserved
    only where v is constant."""
    e = 0.0
    for z0 in range(0, re.shape[0], chunk):
        r = re[z0:z0 + chunk].astype(np.float64)
        i = im[z0:z0 + chunk].astype(np.float64)
        e += float(np.sum((r * r + i * i) / v[z0:z0 + chunk]))
    return e
