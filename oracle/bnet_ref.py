"""FP64 numpy reference of the learned thin-pad boundary network.

TEST INFRASTRUCTURE ONLY (the checker of the HIP kernels in
csrc/hip/ose_bnet.hip).  The reference (Littleke/3D-OSE) has no learned
boundary; this module defines the network the learned thin pad uses so that
the GPU forward pass and gradients can be checked against FP64.

Grid: padded [nzb][nyb][nxb], interior box origin (bl, bf, bu), size
(nx, ny, nz), periodic (the FFT torus).  The pad P is every cell outside the
interior box.  R_d is the set of cells within Chebyshev distance d of P
(R_0 = P).  The network:

  features on R_3 (16 channels, 9 used):
     u/s (re, im), v/vmax, and per axis a: outward distance into the pad
     along a divided by the pad width of a, and the side (-1 before,
     +1 after, 0 inside the interior range of a);
     s = sqrt(mean_{R_3} |u|^2)
  h1 = gelu(conv(f,  W1) + b1)   on R_2
  h2 = gelu(conv(h1, W2) + b2)   on R_1
  y  =      conv(h2, W3) + b3    on R_0, channels 0, 1
  u'[c] = u[c] G(y[c]),  G(y) = q / max(1, |q|),  q = 1 + y0 + i y1,  c in P
The correction is a learned passive complex gain on the pad (|G| <= 1, so
re-applying it every step cannot grow energy): it vanishes where u does and
is invariant to the scale of u (the features see u/s).

conv(h, W)[c, o] = sum_{tap} sum_i h[c + off(tap), i] W[tap, o, i] with
tap = (dz+1)*9 + (dy+1)*3 + (dx+1), offsets taken modulo the grid.
gelu(x) = x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
Training loss: L = mean_{c in P} |(u[c]/s) G(y[c]) - t[c]/s|^2.

Parameters, flat float32 (the C-ABI order): W1[27][16][16], b1[16],
W2[27][16][16], b2[16], W3[27][16][16], b3[16].
"""
import numpy as np

C = 16
NTAP = 27
NPAR = 3 * (NTAP * C * C + C)
K_GELU = np.sqrt(2.0 / np.pi)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = [], 0
    for _ in range(3):
        W = p[o:o + NTAP * C * C].reshape(NTAP, C, C)
        o += NTAP * C * C
        b = p[o:o + C]
        o += C
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers]).astype(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(K_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K_GELU * (1.0 + 3 * 0.044715 * x * x)


class Geometry:
    def __init__(self, shape, origin, interior):
        self.nzb, self.nyb, self.nxb = shape
        self.bl, self.bf, self.bu = origin
        self.nx, self.ny, self.nz = interior
        z, y, x = np.meshgrid(np.arange(self.nzb), np.arange(self.nyb), np.arange(self.nxb), indexing="ij")
        self.inside = ((x >= self.bl) & (x < self.bl + self.nx) & (y >= self.bf) & (y < self.bf + self.ny)
                       & (z >= self.bu) & (z < self.bu + self.nz))
        self.pad = ~self.inside
        self.R = [self.pad]
        for d in range(1, 4):
            self.R.append(self._dilate(self.R[-1]))
        feats = []
        for (c, lo, n, nb) in ((x, self.bl, self.nx, self.nxb), (y, self.bf, self.ny, self.nyb),
                               (z, self.bu, self.nz, self.nzb)):
            width = float(nb - n)
            after = c >= lo + n
            before = c < lo
            dist = np.where(after, c - (lo + n - 1), np.where(before, lo - c, 0)) / width
            side = np.where(after, 1.0, np.where(before, -1.0, 0.0))
            feats.append((dist, side))
        self.geo_feats = feats

    def _dilate(self, m):
        out = m.copy()
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    out |= np.roll(m, (dz, dy, dx), axis=(0, 1, 2))
        return out


def conv(h, W):
    """h [nz][ny][nx][C] -> [nz][ny][nx][C] (periodic), W[tap][o][i]."""
    out = np.zeros(h.shape[:3] + (W.shape[1],))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                sh = np.roll(h, (-dz, -dy, -dx), axis=(0, 1, 2))   # sh[c] = h[c + off]
                out += sh @ W[tap].T
    return out


def conv_T(g, W):
    """Adjoint of conv: d/dh of sum(conv(h, W) * g)."""
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                out += np.roll(g @ W[tap], (dz, dy, dx), axis=(0, 1, 2))
    return out


def conv_wgrad(h, g):
    """dW[tap][o][i] = sum_c h[c + off(tap)][i] g[c][o]."""
    dW = np.zeros((NTAP, g.shape[3], h.shape[3]))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                sh = np.roll(h, (-dz, -dy, -dx), axis=(0, 1, 2))
                dW[tap] = np.einsum("zyxo,zyxi->oi", g, sh)
    return dW


def gain(y0, y1):
    """G and its partial derivatives dG/dy0, dG/dy1 (complex arrays)."""
    q = 1.0 + y0 + 1j * y1
    a = np.abs(q)
    big = a > 1.0
    G = np.where(big, q / np.where(big, a, 1.0), q)
    a3 = np.where(big, a ** 3, 1.0)
    ga = np.where(big, 1.0 / np.where(big, a, 1.0) - q * q.real / a3, 1.0 + 0j)
    gb = np.where(big, 1j / np.where(big, a, 1.0) - q * q.imag / a3, 1j)
    return G, ga, gb


def features(geo, u, v, vmax):
    r3 = geo.R[3]
    s = np.sqrt(np.mean(np.abs(u[r3]) ** 2))
    f = np.zeros(u.shape + (C,))
    if s > 0:
        f[..., 0] = u.real / s
        f[..., 1] = u.imag / s
    f[..., 2] = v / vmax
    for a, (dist, side) in enumerate(geo.geo_feats):
        f[..., 3 + 2 * a] = dist
        f[..., 4 + 2 * a] = side
    f[~r3] = 0.0
    return f, s


def forward(geo, params, u, v, vmax):
    (W1, b1), (W2, b2), (W3, b3) = unpack(params)
    f, s = features(geo, u, v, vmax)
    y1 = (conv(f, W1) + b1) * geo.R[2][..., None]
    h1 = gelu(y1) * geo.R[2][..., None]
    y2 = (conv(h1, W2) + b2) * geo.R[1][..., None]
    h2 = gelu(y2) * geo.R[1][..., None]
    y3 = (conv(h2, W3) + b3) * geo.R[0][..., None]
    G, _, _ = gain(y3[..., 0], y3[..., 1])
    delta = np.where(geo.R[0], u * (G - 1.0), 0.0)
    return delta, dict(f=f, s=s, y1=y1, h1=h1, y2=y2, h2=h2, y3=y3)


def loss_and_grad(geo, params, u, v, vmax, t):
    (W1, b1), (W2, b2), (W3, b3) = unpack(params)
    delta, c = forward(geo, params, u, v, vmax)
    s = c["s"]
    P = geo.R[0]
    M = P.sum()
    un = u / s
    G, ga, gb = gain(c["y3"][..., 0], c["y3"][..., 1])
    er = np.where(P, un * G - t / s, 0.0)
    L = float(np.sum(np.abs(er) ** 2) / M)
    # dL/dy_k = 2 Re(conj(e) u_n dG/dy_k) / M
    g3 = np.zeros(u.shape + (C,))
    g3[..., 0] = 2.0 * np.real(np.conj(er) * un * ga) / M
    g3[..., 1] = 2.0 * np.real(np.conj(er) * un * gb) / M
    dW3 = conv_wgrad(c["h2"], g3)
    db3 = g3.sum(axis=(0, 1, 2))
    g2 = conv_T(g3, W3) * geo.R[1][..., None] * gelu_grad(c["y2"])
    dW2 = conv_wgrad(c["h1"], g2)
    db2 = g2.sum(axis=(0, 1, 2))
    g1 = conv_T(g2, W2) * geo.R[2][..., None] * gelu_grad(c["y1"])
    dW1 = conv_wgrad(c["f"], g1)
    db1 = g1.sum(axis=(0, 1, 2))
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return L, grad
