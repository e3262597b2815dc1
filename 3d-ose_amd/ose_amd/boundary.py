"""Training and evaluation of the learned thin-pad boundary network.

The network (ose_set_boundary_net / ose_bnet_*, libose_hip.so) replaces a
wide absorbing sponge with a thin pad.  It is trained on wavefields the OSE
method itself generates on wide-padded domains:

  1. a wide run (interior n^3, `wide`-cell reference taper) of a seeded random
     layered/anomaly model gives states u_w(t);
  2. for sampled t, the thin window (interior + `thin` cells) of u_w(t) is
     stepped once on the thin grid, giving u_thin(t+1);
  3. the network learns the correction of the thin pad cells towards
     u_w(t+1) restricted to the window (ose_bnet_train, Adam on the GPU).

With replace=True the network is the whole absorbing layer (the thin pad
has no taper) and later rounds aggregate pairs from rollouts (dataset
aggregation): the thin run proceeds with the current network in the loop,
so the inputs carry what the network lets wrap around the periodic grid,
and the wide run still supplies the targets; the thin state is reset to the
wide window every `reset` steps.

Evaluation runs the same held-out models three ways -- wide sponge
(reference), thin pad with the plain taper, thin pad with the network -- and
reports the seismogram rel-L2 of each thin run against the wide run and the
reflected energy (interior error energy relative to the interior energy at
the end of the run).  Everything runs through the C-ABI; the host-side
arithmetic (models, wavelets, Bessel coefficients, tapers) is libose_host.so.
"""
import json
import os
import time

import numpy as np

from . import host as H
from .backend import BoundaryNet, Context, boundary_net_size

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "bnet_thin8.npy")


class Grid:
    """A padded model built by the host library: interior velocity, pads
    (before, after) per axis, order K, coefficients, taper."""

    def __init__(self, vi, pad_before, pad_after, dt=1e-3, d=10.0, taper="ref", ktol=1e-7, vmax=None):
        vi = np.ascontiguousarray(vi, dtype=np.float32)
        nz, ny, nx = vi.shape
        (bl, bf, bu), (br, bb, bd) = pad_before, pad_after
        self.interior = (nx, ny, nz)
        self.origin = (bl, bf, bu)
        self.pads = (bl, br, bf, bb, bu, bd)
        self.v = H.pad_velocity(vi, self.pads)
        vm = float(self.v.max()) if vmax is None else vmax
        self.R = H.spectral_radius(vm, d, d, d)
        self.dt = dt
        self.d = d
        z = self.R * dt
        self.K = H.order_tail(z, ktol)
        self.coef = H.coefficients(z, self.K)
        self.taper = (H.taper(bl, nx, br, taper), H.taper(bf, ny, bb, taper), H.taper(bu, nz, bd, taper))

    @property
    def shape(self):
        return self.v.shape

    def add_to(self, ctx, rec_z=0):
        return ctx.add_model(self.v, interior=self.interior, origin=self.origin, d=(self.d,) * 3, dt=self.dt,
                             R=self.R, coef=self.coef, taper=self.taper, rec_z=rec_z)

    def window(self, other):
        """Slices [z, y, x] of this (wide) grid that cover `other` (thin) grid;
        both share the interior."""
        sl = []
        for a in (2, 1, 0):                       # z, y, x
            start = self.origin[a] - other.origin[a]
            sl.append(slice(start, start + other.shape[2 - a]))
        return tuple(sl)


def extent(n, pad_before, pad_after_min):
    """After-pad that makes the padded extent a compiled FFT size."""
    from .backend import supported_extent
    nb = supported_extent(n + pad_before + pad_after_min)
    return nb - n - pad_before


def make_pair_grids(vi, thin, wide, vmax=None, taper="ref", thin_taper=None, wide_taper=None):
    n = vi.shape[::-1]
    ta = [extent(n[a], thin, thin) for a in range(3)]
    wa = [extent(n[a], wide, wide) for a in range(3)]
    vm = float(vi.max()) if vmax is None else vmax
    thin_g = Grid(vi, (thin,) * 3, tuple(ta), taper=thin_taper or taper, vmax=vm)
    wide_g = Grid(vi, (wide,) * 3, tuple(wa), taper=wide_taper or taper, vmax=vm)
    return thin_g, wide_g


def wide_setup(wide, n, thin, nt):
    """(pad, taper) of the runs that supply the training targets: wide > 0,
    a sponge of that many cells with the reference taper; wide = 0, a
    taper-free domain wide enough that nothing leaving the thin window
    returns to it within nt steps (the targets are then free of any
    absorber's reflections)."""
    if wide > 0:
        return wide, None
    return unreflected_pad(n, nt) + thin, "none"


def _wavelet(nt, fm, dt, delay=40):
    s = H.ricker(nt, fm, dt, 100.0, delay)
    return s, H.hilbert(s)


def generate_pairs(net, seeds, n=64, thin=8, wide=64, nt=400, t0=120, every=8, fm=20.0, replace=True, log=print):
    """Add (u_thin(t+1), u_wide(t+1)|window) pairs of the given seeded models to net."""
    count = 0
    wide, wtaper = wide_setup(wide, n, thin, nt)
    for seed in seeds:
        vi = H.model("train", n, n, n, seed=seed)
        # a network that replaces the taper of the thin pad learns from untapered steps
        thin_g, wide_g = make_pair_grids(vi, thin, wide, thin_taper="none" if replace else None, wide_taper=wtaper)
        rng = np.random.default_rng(seed)
        src = (int(rng.integers(n // 4, 3 * n // 4)), int(rng.integers(n // 4, 3 * n // 4)),
               int(rng.integers(0, n // 3)))
        s, h = _wavelet(nt + 1, fm, wide_g.dt, delay=int(rng.integers(10, 40)))
        win = wide_g.window(thin_g)
        with Context(0) as cw, Context(0) as ct:
            mw = wide_g.add_to(cw)
            sw = cw.add_shot(mw, (wide_g.origin[0] + src[0], wide_g.origin[1] + src[1], wide_g.origin[2] + src[2]), s, h)
            mt = thin_g.add_to(ct)
            st = ct.add_shot(mt, (thin_g.origin[0] + src[0], thin_g.origin[1] + src[1], thin_g.origin[2] + src[2]), s, h)
            cw.run(t0)
            re, im = cw.wavefield(sw)
            for t in range(t0, nt, every):
                uw = (re + 1j * im)[win]
                ct.set_wavefield(st, uw.real, uw.imag)
                ct.step = t
                ct.run(1)
                rt, it_ = ct.wavefield(st)
                cw.run(1)
                re, im = cw.wavefield(sw)
                target = (re + 1j * im)[win]
                count = net.add_sample(rt + 1j * it_, target)
                cw.run(every - 1)
                re, im = cw.wavefield(sw)
        log(f"[boundary] seed {seed}: {count} pairs")
    return count


def generate_rollout_pairs(net, params, seeds, n=64, thin=8, wide=64, nt=400, t0=120, every=4, reset=48,
                           fm=20.0, replace=True, guard=4.0, log=print):
    """Pairs whose inputs come from thin runs with the network in the loop."""
    count = 0
    wide, wtaper = wide_setup(wide, n, thin, nt + every)
    for seed in seeds:
        vi = H.model("train", n, n, n, seed=seed)
        thin_g, wide_g = make_pair_grids(vi, thin, wide, thin_taper="none" if replace else None, wide_taper=wtaper)
        rng = np.random.default_rng(seed)
        src = (int(rng.integers(n // 4, 3 * n // 4)), int(rng.integers(n // 4, 3 * n // 4)),
               int(rng.integers(0, n // 3)))
        s, h = _wavelet(nt + every + 1, fm, wide_g.dt, delay=int(rng.integers(10, 40)))
        win = wide_g.window(thin_g)
        with Context(0) as cw, Context(0) as ct, Context(0) as cn:
            mw = wide_g.add_to(cw)
            sw = cw.add_shot(mw, tuple(wide_g.origin[a] + src[a] for a in range(3)), s, h)
            mn = thin_g.add_to(cn)
            sn = cn.add_shot(mn, tuple(thin_g.origin[a] + src[a] for a in range(3)), s, h)
            mt = thin_g.add_to(ct)
            st = ct.add_shot(mt, tuple(thin_g.origin[a] + src[a] for a in range(3)), s, h)
            ct.set_boundary_net(mt, params, guard, replace_taper=replace)
            cw.run(t0)
            for t in range(t0, nt, every):
                re, im = cw.wavefield(sw)
                if (t - t0) % reset == 0:
                    uw = (re + 1j * im)[win]
                    ct.set_wavefield(st, uw.real, uw.imag)
                    ct.step = t
                rt, it_ = ct.wavefield(st)
                cn.set_wavefield(sn, rt, it_)
                cn.step = t
                cn.run(1)
                rn, in_ = cn.wavefield(sn)
                cw.run(1)
                re, im = cw.wavefield(sw)
                count = net.add_sample(rn + 1j * in_, (re + 1j * im)[win])
                ct.run(every)
                cw.run(every - 1)
            if ct.boundary_status(mt)["tripped"]:
                log(f"[boundary] seed {seed}: guard tripped during the rollout")
        log(f"[boundary] seed {seed}: {count} pairs (rollout)")
    return count


def train(seeds=range(1000, 1016), n=64, thin=8, wide=64, iters=4000, lr=2e-3, log=print, out=None,
          replace=True, rounds=3, init=None):
    """init: parameters to continue from; training then starts with a
    rollout round (no teacher-forced pairs)."""
    vi = H.model("train", n, n, n, seed=int(seeds[0]))
    thin_g, _ = make_pair_grids(vi, thin, wide, thin_taper="none")
    net = BoundaryNet(thin_g.v, thin_g.origin, thin_g.interior)
    first = 0
    if init is not None:
        net.set_params(init)
        first = 1
    else:
        t = time.time()
        generate_pairs(net, seeds, n=n, thin=thin, wide=wide, replace=replace, log=log)
        log(f"[boundary] pairs generated in {time.time() - t:.1f} s")
    losses = []
    for r in range(first, (first + rounds) if replace else 1):
        if r > 0:
            t = time.time()
            generate_rollout_pairs(net, net.params(), seeds, n=n, thin=thin, wide=wide, replace=replace, log=log)
            log(f"[boundary] round {r}: rollout pairs in {time.time() - t:.1f} s")
        t = time.time()
        lo = net.train(iters, lr=lr, seed=7 + r)
        w = max(1, iters // 10)
        curve = " ".join(f"{lo[i:i + w].mean():.4g}" for i in range(0, iters, w))
        log(f"[boundary] round {r}: {iters} Adam steps (lr {lr}) in {time.time() - t:.1f} s; loss by tenths: {curve}")
        losses.append(lo)
    losses = np.concatenate(losses)
    p = net.params()
    net.close()
    if out:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        np.save(out, p)
    return p, losses


def evaluate(params, seeds, n=64, thin=8, wide=64, nt=400, fm=20.0, model_kind="train", guard=4.0, replace=True,
             log=None):
    """Rel-L2 of surface seismograms and reflected energy of thin runs
    (plain taper, learned pad) against the wide-sponge run."""
    rows = []
    for seed in seeds:
        vi = H.model(model_kind, n, n, n, seed=seed)
        thin_g, wide_g = make_pair_grids(vi, thin, wide)
        s, h = _wavelet(nt, fm, wide_g.dt)
        src = (n // 2, n // 2, 2)
        res = {}
        runs = [("wide", wide_g, None), ("plain", thin_g, None)]
        if params is not None:
            runs.append(("learned", thin_g, params))
        for name, g, p in runs:
            with Context(0) as c:
                m = g.add_to(c)
                sid = c.add_shot(m, (g.origin[0] + src[0], g.origin[1] + src[1], g.origin[2] + src[2]), s, h)
                if p is not None:
                    c.set_boundary_net(m, p, guard, replace_taper=replace)
                t0 = time.time()
                c.run(nt)
                c.sync()
                dt = time.time() - t0
                tr = c.traces(sid)
                re, im = c.wavefield(sid)
                st = c.boundary_status(m) if p is not None else None
            bl, bf, bu = g.origin
            u_int = (re + 1j * im)[bu:bu + n, bf:bf + n, bl:bl + n]
            res[name] = dict(tr=tr, u=u_int, s=dt, status=st)
        ref_tr, ref_u = res["wide"]["tr"], res["wide"]["u"]
        row = {"seed": int(seed)}
        row["thin"] = thin
        for name in [r[0] for r in runs[1:]]:
            row[f"{name}_rel_l2"] = float(np.linalg.norm(res[name]["tr"] - ref_tr) / np.linalg.norm(ref_tr))
            row[f"{name}_reflected"] = float(np.linalg.norm(res[name]["u"] - ref_u) ** 2 / np.linalg.norm(ref_u) ** 2)
        if params is not None:
            row["learned_status"] = res["learned"]["status"]
        row["run_seconds"] = {k: round(v["s"], 3) for k, v in res.items()}
        rows.append(row)
        if log:
            log(json.dumps(row))
    return rows


def unreflected_pad(n, nt, dt=1e-3, d=10.0, vmax=4500.0, margin=16):
    """Pad width of a taper-free reference run on which nothing that leaves
    the interior returns to the surface receivers within nt steps.  The
    periodic domain's nearest image of a source at the surface centre lies
    n + 2W away along x or y, i.e. n/2 + 2W from the farthest receiver; a
    wave covers at most vmax nt dt / d cells, so W > (that - n/2) / 2; the
    margin covers the spectral operator's tail ahead of the wavefront."""
    reach = vmax * nt * dt / d
    return max(8, int(np.ceil((reach - n / 2) / 2)) + margin)


def run_traces(g, src, s, h, nt, params=None, guard=4.0, replace=True):
    """Surface seismogram of one run on grid g (and the learned pad's status)."""
    with Context(0) as c:
        m = g.add_to(c)
        sid = c.add_shot(m, tuple(g.origin[a] + src[a] for a in range(3)), s, h)
        if params is not None:
            c.set_boundary_net(m, params, guard, replace_taper=replace)
        t0 = time.time()
        c.run(nt)
        c.sync()
        dt = time.time() - t0
        tr = c.traces(sid)
        st = c.boundary_status(m) if params is not None else None
    return tr, st, dt


def compare_pads(seeds, boundaries, n=64, nt=400, fm=20.0, model_kind="train", log=None, check=False):
    """Seismogram rel-L2 of thin boundaries against an unreflected reference.

    boundaries: list of (name, width, params) -- params None for the plain
    reference taper of that width (generate_att), else the learned pad's
    network (replacing the taper).  The reference is a taper-free run on a
    domain wide enough (unreflected_pad) that nothing leaving the interior
    returns to the receivers within the window; with check=True it is
    compared with one 32 cells wider, which must agree to FP32 rounding."""
    rows = []
    for seed in seeds:
        vi = H.model(model_kind, n, n, n, seed=seed)
        vm = float(vi.max())
        wref = unreflected_pad(n, nt, vmax=vm)
        s, h = _wavelet(nt, fm, 1e-3)
        src = (n // 2, n // 2, 2)
        row = {"seed": int(seed), "vmax": vm, "ref_pad": wref}
        refs = {}
        for w in ((wref, wref + 32) if check else (wref,)):
            g = Grid(vi, (w,) * 3, tuple(extent(vi.shape[::-1][a], w, w) for a in range(3)), taper="none", vmax=vm)
            refs[w], _, t = run_traces(g, src, s, h, nt)
            row.setdefault("ref_padded", list(g.shape[::-1]))
        ref = refs[wref]
        nref = np.linalg.norm(ref)
        if check:
            row["ref_vs_wider"] = float(np.linalg.norm(refs[wref + 32] - ref) / nref)
        for name, w, p in boundaries:
            g = Grid(vi, (w,) * 3, tuple(extent(vi.shape[::-1][a], w, w) for a in range(3)),
                     taper="none" if p is not None else "ref", vmax=vm)
            tr, st, t = run_traces(g, src, s, h, nt, params=p)
            row[name] = float(np.linalg.norm(tr - ref) / nref)
            if st is not None:
                row[name + "_status"] = st
        rows.append(row)
        if log:
            log(json.dumps(row))
    return rows


def main():
    import argparse
    ap = argparse.ArgumentParser(description="train / evaluate the learned thin-pad network")
    ap.add_argument("cmd", choices=["train", "eval", "compare"])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--thin", type=int, default=8)
    ap.add_argument("--wide", type=int, default=64,
                    help="train: sponge width of the target runs; 0: taper-free targets on an unreflected domain")
    ap.add_argument("--models", type=int, default=16)
    ap.add_argument("--iters", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--nt", type=int, default=400)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--seeds", default="2000:2004")
    ap.add_argument("--guard", type=float, default=4.0)
    ap.add_argument("--keep-taper", action="store_true", help="the network acts on top of the thin taper")
    ap.add_argument("--rounds", type=int, default=3, help="rounds of rollout pair aggregation")
    ap.add_argument("--plain-only", action="store_true", help="eval: wide sponge and thin taper only")
    ap.add_argument("--init", default=None, help="train: parameters to continue from")
    ap.add_argument("--seed0", type=int, default=1000, help="train: first model seed")
    ap.add_argument("--plain", default="8,16,24,32,40", help="compare: plain taper widths")
    ap.add_argument("--learned", default="", help="compare: width=weights.npy pairs, comma separated")
    ap.add_argument("--check-ref", action="store_true", help="compare: also run a 32-cell wider reference")
    args = ap.parse_args()
    if args.cmd == "compare":
        a, b = (int(x) for x in args.seeds.split(":"))
        bnd = [(f"plain{w}", int(w), None) for w in args.plain.split(",") if w]
        for item in (x for x in args.learned.split(",") if x):
            w, path = item.split("=")
            stem = os.path.splitext(os.path.basename(path))[0]
            bnd.append((f"learned{w}_{stem}", int(w), np.load(path)))
        compare_pads(range(a, b), bnd, n=args.n, nt=args.nt, log=lambda m: print(m, flush=True),
                     check=args.check_ref)
        return
    if args.cmd == "train":
        init = np.load(args.init) if args.init else None
        train(range(args.seed0, args.seed0 + args.models), n=args.n, thin=args.thin, wide=args.wide, iters=args.iters,
              lr=args.lr, out=args.out, replace=not args.keep_taper, rounds=args.rounds, init=init)
    else:
        a, b = (int(x) for x in args.seeds.split(":"))
        p = None
        if not args.plain_only:
            p = np.load(args.weights)
            assert p.size == boundary_net_size()
        evaluate(p, range(a, b), n=args.n, thin=args.thin, wide=args.wide, nt=args.nt, guard=args.guard,
                 replace=not args.keep_taper, log=lambda m: print(m, flush=True))


if __name__ == "__main__":
    main()
