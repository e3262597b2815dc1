"""Generate the per-extent kernel instances of the HIP backend.

For every supported FFT extent N = N1 * N2 this writes inst_<N>.hip, which
instantiates the strided (y/z) and x-line kernels of ose_kernels.h for that
split and fills an ose::FftEntry with their addresses, block sizes and LDS
sizes.  It also writes dft_gen.h (the register DFTs for every N1, N2) and
ose_sizes.inc (the registry list).

SIZES are the extents the configurations use (plus 152 = 8*19, the
reference's own 101+25+26 grid, and 832, where the 768^3 workload's 24-cell
taper lands: 816 has the prime factor 17), compiled in every launch shape the
autotuner chooses from; LEAN are the other even extents up to 1024 with
small prime factors, one shape each.  Every other extent up to 4096 runs
on the runtime-length path (ose_generic.h); the host's extent=smooth rounds
a padded size up to the next compiled one.

Run: python gen_instances.py <outdir>
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# N: (N1, N2).  N1 = register DFT length of step A (run by N2 threads per
# pencil), N2 = register DFT length of step B (run by N1 threads).
SIZES = {
    16: (4, 4), 24: (4, 6), 32: (4, 8), 48: (6, 8), 64: (8, 8), 80: (8, 10),
    96: (8, 12), 112: (8, 14), 128: (8, 16), 140: (10, 14), 144: (12, 12),
    152: (8, 19), 160: (10, 16), 192: (12, 16), 200: (10, 20), 224: (14, 16),
    240: (15, 16), 256: (16, 16), 272: (16, 17), 280: (14, 20), 288: (16, 18), 320: (16, 20),
    336: (16, 21), 378: (18, 21), 384: (16, 24), 400: (20, 20), 448: (16, 28),
    480: (20, 24), 512: (16, 32), 576: (24, 24), 600: (20, 30), 640: (20, 32),
    672: (21, 32), 720: (24, 30), 768: (24, 32), 784: (28, 28), 800: (25, 32),
    832: (26, 32), 864: (27, 32), 960: (30, 32), 1024: (32, 32),
}

# radices of the runtime-length path's register butterflies (ose_generic.h)
GENERIC_RADICES = (2, 3, 4, 5, 7, 8, 11, 13, 16)

def _prime_factors(n):
    out, p = [], 2
    while n > 1:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    return out


def _split(n):
    """The most balanced N1 * N2 = n with 2 <= N1 <= N2 <= 32, or None."""
    best = None
    for a in range(2, 33):
        b, r = divmod(n, a)
        if r == 0 and a <= b <= 32 and (best is None or b - a < best[1] - best[0]):
            best = (a, b)
    return best


# Lean instances: every other even extent up to 1024 whose prime factors are
# at most 13 and that splits into two register DFTs, compiled with a single
# launch shape per pass (no autotuner choice) to keep the library small.
# Extents outside both lists take the runtime-length path (ose_generic.h).
LEAN = {n: _split(n) for n in range(16, 1025, 2)
        if n not in SIZES and _split(n) and max(_prime_factors(n)) <= 13}

STRIDED_TS = (8, 16, 32)
LDS_MAX = 160 * 1024


def strided_variants(n1, n2):
    """(T, split) launch shapes: T x columns per workgroup; split = 1
    exchanges real and imaginary parts separately (half the LDS).  The split
    form is compiled where the whole-complex exchange needs more than a
    quarter of a CU's LDS (it then doubles the workgroups per CU)."""
    nt = max(n1, n2)
    out = []
    for t in STRIDED_TS:
        full = 8 * n1 * n2 * (t + 1)
        if t * nt > 1024:
            continue
        if t * nt > 896:
            # 1024-thread workgroups (N1 or N2 of 29..32) spill 20-30 VGPRs of
            # the register DFTs; compiled, split form only, for the long
            # extents, where 256-byte rows halve the z pass's page visits
            # per byte; the autotuner keeps them only where they are faster
            if n1 * n2 >= 600 and full // 2 <= LDS_MAX:
                out.append((t, 1))
            continue
        if full <= LDS_MAX:
            out.append((t, 0))
        if full > LDS_MAX // 4 and full // 2 <= LDS_MAX:
            out.append((t, 1))
    return out


def x_variants(n1, n2):
    """(P, split, twb) launch shapes: P lines per workgroup; the split
    (half-LDS) exchange where a line's whole-complex exchange exceeds 2 KiB;
    twb = 1 (scheduling barrier before the twiddle multiplies) for extents
    of 600 and more, where register pressure decides the occupancy."""
    nt = max(n1, n2)
    pn = n1 | 1
    out = []
    big = n1 * n2 >= 600
    ps = {max(1, 128 // nt), max(1, 256 // nt)} | ({max(1, 64 // nt)} if big else set())
    for twb in ((0, 1) if big else (0,)):
        for p in sorted(ps):
            full = 8 * p * n2 * pn
            if p * nt <= 1024 and full <= LDS_MAX:
                out.append((p, 0, twb))
                if 8 * n2 * pn > 2048:
                    out.append((p, 1, twb))
    return out


HEAD = """// Generated by csrc/gen/gen_instances.py -- do not edit.
#include "ose_kernels.h"
#include "ose_entry.h"

namespace ose {{
void fill_entry_{N}(FftEntry& e)
{{
    constexpr int N1 = {N1}, N2 = {N2};
    e.N = {N}; e.N1 = N1; e.N2 = N2;
"""
STRIDED = """    {{
        constexpr int T = {T}, SPLIT = {SPLIT};
        using SC = StridedCfg<N1, N2, T, SPLIT>;
        StridedKern k;
        k.T = T; k.split = SPLIT; k.threads = SC::THREADS; k.lds = SC::LDS_BYTES;
        k.fn = (const void*)&k_strided<N1, N2, T, -1, SPLIT>; e.yf[e.ns] = k;
        k.fn = (const void*)&k_strided<N1, N2, T, 1, SPLIT>; e.yi[e.ns] = k;
        k.fn = (const void*)&k_zsym<N1, N2, T, SPLIT>; k.lds = ZsymCfg<N1, N2, T, SPLIT>::LDS_BYTES; e.zs[e.ns] = k;
        ++e.ns;
    }}
"""
XPASS = """    {{
        constexpr int P = {P}, SPLIT = {SPLIT}, TWB = {TWB};
        using XC = XCfg<N1, N2, P, SPLIT>;
        XKern k;
        k.P = P; k.split = SPLIT; k.twb = TWB; k.threads = XC::THREADS; k.lds = XC::LDS_BYTES;
        k.fn[XM_ORDER0] = (const void*)&k_xpass<N1, N2, P, XM_ORDER0, SPLIT, TWB>;
        k.fn[XM_MID] = (const void*)&k_xpass<N1, N2, P, XM_MID, SPLIT, TWB>;
        k.fn[XM_LAST] = (const void*)&k_xpass<N1, N2, P, XM_LAST, SPLIT, TWB>;
        k.fn[XM_FWD] = (const void*)&k_xpass<N1, N2, P, XM_FWD, SPLIT, TWB>;
        k.fn[XM_INV] = (const void*)&k_xpass<N1, N2, P, XM_INV, SPLIT, TWB>;
        k.fn[XM_LASTA] = (const void*)&k_xpass<N1, N2, P, XM_LASTA, SPLIT, TWB>;
        k.fn[XM_SKIP] = (const void*)&k_xpass<N1, N2, P, XM_SKIP, SPLIT, TWB>;
        k.fn[XM_ORDER0N] = (const void*)&k_xpass<N1, N2, P, XM_ORDER0N, SPLIT, TWB>;
        k.fn[XM_MIDF] = (const void*)&k_xpass<N1, N2, P, XM_MIDF, SPLIT, TWB>;
        e.xk[e.nx++] = k;
    }}
"""
ZCHUNK = """    {{
        constexpr int T = {T};
        using ZC = ZChunkCfg<N1, N2, T>;
        StridedKern k;
        k.T = T; k.split = 0; k.threads = ZC::THREADS; k.lds = ZC::LDS_BYTES;
        k.fn = (const void*)&k_zsym_chunk<N1, N2, T>;
        for (int m : {{{MS}}}) {{
            k.chunk = m;
            e.zs[e.ns] = k;
            e.yf[e.ns] = StridedKern();
            e.yi[e.ns] = StridedKern();
            ++e.ns;
        }}
    }}
"""
TAIL = """}}
}}  // namespace ose
"""

# chunked z pass (k_zsym_chunk): T = 16 columns, whole-complex exchange plus
# the twiddle rows and |k_z|^2 in LDS, one workgroup per CU; compiled for the
# autotuned extents from 256 up, where the z pass streams from HBM
ZCHUNK_T = 16
ZCHUNK_MS = (2, 3, 4, 6)


def zchunk_fits(n1, n2):
    n = n1 * n2
    lds = 8 * (n * (ZCHUNK_T + (n1 % 2 == 0)) + 2 * n) + 4 * n   # exchange pitch: XchgPitch
    return n >= 256 and n1 <= n2 and ZCHUNK_T * n2 <= 1024 and lds <= LDS_MAX


def lean_variants(n1, n2):
    """One strided and one x-pass shape: T = 16 (split exchange above 40 KiB
    of LDS) and P = 128 / NT lines (split above 2 KiB per line)."""
    nt = max(n1, n2)
    full = 8 * n1 * n2 * 17
    strided = [(16, 1 if full > 40 * 1024 else 0)]
    p = max(1, 128 // nt)
    xs = [(p, 1 if 8 * n2 * (n1 | 1) > 2048 else 0, 0)]
    return strided, xs


def instance_source(n, n1, n2, lean=False):
    src = HEAD.format(N=n, N1=n1, N2=n2)
    strided, xs = lean_variants(n1, n2) if lean else (strided_variants(n1, n2), x_variants(n1, n2))
    for t, split in strided:
        src += STRIDED.format(T=t, SPLIT=split)
    if not lean and zchunk_fits(n1, n2):
        src += ZCHUNK.format(T=ZCHUNK_T, MS=", ".join(str(m) for m in ZCHUNK_MS))
    for p, split, twb in xs:
        src += XPASS.format(P=p, SPLIT=split, TWB=twb)
    return src + TAIL.format()


def main():
    out = sys.argv[1]
    only = [int(a) for a in sys.argv[2:]]
    allsizes = {**SIZES, **LEAN}
    sizes = {n: allsizes[n] for n in (only or allsizes)}
    os.makedirs(out, exist_ok=True)
    # register DFTs of the compiled splits and of the runtime-length path's
    # radices (ose_generic.h)
    factors = sorted({f for (a, b) in allsizes.values() for f in (a, b)} | set(GENERIC_RADICES))
    subprocess.check_call([sys.executable, os.path.join(HERE, "gen_dft.py"),
                           os.path.join(out, "dft_gen.h")] + [str(f) for f in factors])
    # the registry always lists every extent; `only` limits which instance
    # sources are (re)written
    with open(os.path.join(out, "ose_sizes.inc"), "w") as f:
        f.write("// Generated by csrc/gen/gen_instances.py -- do not edit.\n")
        for n in sorted(allsizes):
            f.write(f"OSE_SIZE({n})\n")
    for n in sorted(sizes):
        n1, n2 = sizes[n]
        src = instance_source(n, n1, n2, lean=n in LEAN)
        path = os.path.join(out, f"inst_{n}.hip")
        old = open(path).read() if os.path.exists(path) else None
        if old != src:
            with open(path, "w") as f:
                f.write(src)


if __name__ == "__main__":
    main()
