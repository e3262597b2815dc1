"""Generate straight-line register DFT kernels for the HIP FFT passes.

Writes csrc/hip/dft_gen.h: for every length N used as a factor of a
supported FFT extent, two device functions (forward sign -1, inverse +1)

    template <> __device__ __forceinline__ void rdft<N, S>(float2* v)

that transform v[0..N-1] in place (natural order in and out, unnormalised)
entirely in registers.  Composite N use a decimation-in-time split N = p*m
(radix 4 first, then 2, 3, 5, 7, ...); odd primes use the symmetric pair
form X_r = x0 + sum_q [(x_q + x_{p-q}) cos + i s (x_q - x_{p-q}) sin], so a
p-point DFT costs (p-1)^2 real multiply-adds.  Twiddle constants are
computed in double precision here and rounded once to float.

Run: python gen_dft.py <out_header> <N> [<N> ...]
"""
import math
import os
import sys


class Emitter:
    def __init__(self):
        self.lines = []
        self.n = 0

    def tmp(self):
        self.n += 1
        return f"t{self.n}"

    def let(self, expr_x, expr_y):
        t = self.tmp()
        self.lines.append(f"const float2 {t} = make_float2({expr_x}, {expr_y});")
        return t


def flit(c):
    if c == 0.0:
        return "0.0f"
    return repr(float(c)) + "f"


def add(e, a, b):
    return e.let(f"{a}.x + {b}.x", f"{a}.y + {b}.y")


def sub(e, a, b):
    return e.let(f"{a}.x - {b}.x", f"{a}.y - {b}.y")


def mul_i(e, a, s):
    """a * (s*i), s = +-1."""
    if s > 0:
        return e.let(f"-{a}.y", f"{a}.x")
    return e.let(f"{a}.y", f"-{a}.x")


def twiddle(e, a, num, den, sign):
    """a * exp(sign * 2 pi i num/den), with exact special angles."""
    num %= den
    if num == 0:
        return a
    # reduce: multiples of quarter turns
    if (4 * num) % den == 0:
        q = (4 * num) // den  # 1,2,3 quarter turns
        if q == 2:
            return e.let(f"-{a}.x", f"-{a}.y")
        s = sign if q == 1 else -sign
        return mul_i(e, a, s)
    ang = sign * 2.0 * math.pi * num / den
    c, s = math.cos(ang), math.sin(ang)
    if (8 * num) % den == 0:
        h = math.sqrt(0.5)
        cs = flit(h)
        # (a.x + i a.y)(c + i s) with |c| = |s| = sqrt(1/2):
        # re = c*ax - s*ay, im = s*ax + c*ay
        re = f"{cs} * ({'' if c > 0 else '-'}{a}.x {'-' if s > 0 else '+'} {a}.y)"
        im = f"{cs} * ({'' if s > 0 else '-'}{a}.x {'+' if c > 0 else '-'} {a}.y)"
        return e.let(re, im)
    return e.let(f"{a}.x * {flit(c)} - {a}.y * {flit(s)}", f"{a}.x * {flit(s)} + {a}.y * {flit(c)}")


def dft_prime(e, x, sign):
    p = len(x)
    if p == 1:
        return list(x)
    if p == 2:
        return [add(e, x[0], x[1]), sub(e, x[0], x[1])]
    h = (p - 1) // 2
    a = [None] * (h + 1)
    b = [None] * (h + 1)
    for q in range(1, h + 1):
        a[q] = add(e, x[q], x[p - q])
        b[q] = sub(e, x[q], x[p - q])
    # X0
    acc = x[0]
    for q in range(1, h + 1):
        acc = add(e, acc, a[q])
    out = [None] * p
    out[0] = acc
    for r in range(1, h + 1):
        rex = [f"{x[0]}.x"]
        rey = [f"{x[0]}.y"]
        sx = []
        sy = []
        for q in range(1, h + 1):
            th = 2.0 * math.pi * ((q * r) % p) / p
            c, s = math.cos(th), math.sin(th)
            rex.append(f"{flit(c)} * {a[q]}.x")
            rey.append(f"{flit(c)} * {a[q]}.y")
            sx.append(f"{flit(s)} * {b[q]}.x")
            sy.append(f"{flit(s)} * {b[q]}.y")
        Rr = e.let(" + ".join(rex), " + ".join(rey))
        Sr = e.let(" + ".join(sx), " + ".join(sy))
        iS = mul_i(e, Sr, sign)          # i*sign*S
        out[r] = add(e, Rr, iS)
        out[p - r] = sub(e, Rr, iS)
    return out


def dft4(e, x, sign):
    a0 = add(e, x[0], x[2])
    a1 = sub(e, x[0], x[2])
    b0 = add(e, x[1], x[3])
    b1 = sub(e, x[1], x[3])
    rb1 = mul_i(e, b1, sign)
    return [add(e, a0, b0), add(e, a1, rb1), sub(e, a0, b0), sub(e, a1, rb1)]


def smallest_factor(n):
    for p in range(2, n + 1):
        if n % p == 0:
            return p
    return n


def choose_radix(n):
    if n % 4 == 0 and n != 4:
        return 4
    if n == 4:
        return 4
    if n % 2 == 0:
        return 2
    return smallest_factor(n)


def dft(e, x, sign):
    n = len(x)
    if n == 1:
        return list(x)
    if n == 4:
        return dft4(e, x, sign)
    p = choose_radix(n)
    if p == n:
        return dft_prime(e, x, sign)
    m = n // p
    # DIT: X[k2 + m k1] = sum_{n1} w_p^{n1 k1} (w_n^{n1 k2} Y_{n1}[k2])
    Y = [dft(e, [x[p * n2 + n1] for n2 in range(m)], sign) for n1 in range(p)]
    out = [None] * n
    for k2 in range(m):
        col = [twiddle(e, Y[n1][k2], n1 * k2, n, sign) for n1 in range(p)]
        if p == 4:
            Z = dft4(e, col, sign)
        else:
            Z = dft_prime(e, col, sign)
        for k1 in range(p):
            out[k2 + m * k1] = Z[k1]
    return out


def gen_one(n, sign):
    e = Emitter()
    xs = [f"v{i}" for i in range(n)]
    body = [f"const float2 v{i} = v[{i}];" for i in range(n)]
    out = dft(e, xs, sign)
    body += e.lines
    body += [f"v[{i}] = {out[i]};" for i in range(n)]
    s = "-1" if sign < 0 else "1"
    hdr = f"template <> OSE_RDFT_QUAL void rdft<{n}, {s}>(float2* __restrict__ v)\n{{\n"
    return hdr + "\n".join("    " + l for l in body) + "\n}\n"


def main():
    out = sys.argv[1]
    ns = sorted(set(int(a) for a in sys.argv[2:]))
    parts = [
        "// Generated by csrc/gen/gen_dft.py -- do not edit.\n",
        "// Straight-line register DFTs: rdft<N, S>(v) transforms v[0..N-1] in place,\n",
        "// natural order, unnormalised, kernel exp(S * 2 pi i n k / N).\n",
        "#pragma once\n#if defined(__HIPCC__)\n#include <hip/hip_runtime.h>\n",
        "#define OSE_RDFT_QUAL __device__ __forceinline__\n#else\n#define OSE_RDFT_QUAL inline\n#endif\n\n",
        "template <int N, int S> OSE_RDFT_QUAL void rdft(float2* __restrict__ v);\n\n",
    ]
    for n in ns:
        parts.append(gen_one(n, -1))
        parts.append(gen_one(n, +1))
    text = "".join(parts)
    old = open(out).read() if os.path.exists(out) else None
    if old != text:
        with open(out, "w") as f:
            f.write(text)


if __name__ == "__main__":
    main()
