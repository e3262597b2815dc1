"""numpy restatement of the snapshot encodings (checker side).

Layout of one encoded snapshot of the lanes `lanes` (bit 0 Re u, bit 1
Im u) over a box [nz][ny][nx]: for Q16/Q8 the per-x-line scales first,
float32 [lane][z][y], then the values [lane][z][y][x]; F32 and BF16 hold
only the values.
"""
import numpy as np

F32, BF16, Q16, Q8 = 0, 1, 2, 3
QMAX = {Q16: 32767.0, Q8: 127.0}


def bf16_bits(v):
    """float32 -> bfloat16 bits, round to nearest even (NaN kept quiet)."""
    b = np.ascontiguousarray(v, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (b & 0x7FFFFFFF) > 0x7F800000
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((b >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def encode_lane(fmt, v):
    """v: float32 [nz][ny][nx] -> (scales or None, values array)."""
    v = np.ascontiguousarray(v, dtype=np.float32)
    if fmt == F32:
        return None, v
    if fmt == BF16:
        return None, bf16_bits(v)
    s = np.abs(v).max(axis=2).astype(np.float32)
    qmax = np.float32(QMAX[fmt])
    with np.errstate(divide="ignore"):
        inv = np.where(s > 0, qmax / np.where(s > 0, s, np.float32(1)), np.float32(0)).astype(np.float32)
    q = np.clip(np.rint(v * inv[:, :, None]), -qmax, qmax)
    return s, q.astype(np.int8 if fmt == Q8 else np.int16)


def encode(fmt, lanes_data):
    """Bytes of a snapshot of the given lanes (list of [nz][ny][nx] arrays)."""
    parts = [encode_lane(fmt, v) for v in lanes_data]
    out = b""
    if fmt in QMAX:
        out += b"".join(p[0].tobytes() for p in parts)
    return out + b"".join(p[1].tobytes() for p in parts)


def decode_lane(fmt, shape, scales, values):
    nz, ny, nx = shape
    if fmt == F32:
        return values.reshape(shape).astype(np.float32)
    if fmt == BF16:
        return (values.astype(np.uint32) << 16).view(np.float32).reshape(shape)
    step = (scales / np.float32(QMAX[fmt])).astype(np.float32)
    return (values.reshape(shape).astype(np.float32) * step.reshape(nz, ny, 1)).astype(np.float32)


def bound(fmt, v):
    """Largest decode error the encoding allows for each value of v."""
    v = np.asarray(v, dtype=np.float32)
    if fmt == F32:
        return np.zeros_like(v)
    if fmt == BF16:
        return np.abs(v) * 2.0 ** -8 + 1e-38
    s = np.abs(v).max(axis=2, keepdims=True)
    return np.broadcast_to(0.5 * s / QMAX[fmt] * (1 + 1e-5) + 1e-6 * s, v.shape)
