"""Independent reader of the GPZ byte stream, written from the format
(header, u64 offsets, per-group meta, delta / count planes MSB-first over
run index, rel planes LSB-first over sorted rank).  Tests use it to check
stream contents against quantities computed with numpy, without going
through the oracle's decoder."""
import struct

import numpy as np


def header(s):
    s = bytes(s[:32])
    dim, n, bits = struct.unpack_from("<IQI", s, 0)
    if bits == 64:
        eb, seg, ps = struct.unpack_from("<dII", s, 16)
        return dict(dim=dim, n=n, bits=bits, eb=eb, seg=seg, P=32, M=64)
    eb, seg = struct.unpack_from("<fI", s, 16)
    return dict(dim=dim, n=n, bits=bits, eb=eb, seg=seg, P=24, M=36)


def _pad(v, a):
    return (v + a - 1) // a * a


def section_sizes(bits, dim, U, db, cb, rb):
    B = (U + 7) // 8
    if bits == 32:
        return _pad(B * db, 4), _pad(B * cb, 4), 128 * rb
    return _pad(B * db, 8), _pad(B * cb, 8 if dim == 3 else 4), 128 * rb


def _msb_planes(buf, nbits, B, count):
    out = np.zeros(count, np.uint64)
    for p in range(nbits):
        plane = np.unpackbits(np.frombuffer(buf[p * B:(p + 1) * B], np.uint8),
                              bitorder="big")[:count].astype(np.uint64)
        out |= plane << np.uint64(nbits - 1 - p)
    return out


def _lsb_planes(buf, nbits):
    out = np.zeros(1024, np.uint64)
    for p in range(nbits):
        plane = np.unpackbits(np.frombuffer(buf[p * 128:(p + 1) * 128], np.uint8),
                              bitorder="little").astype(np.uint64)
        out |= plane << np.uint64(nbits - 1 - p)
    return out


def group(s, c):
    """Decoded fields of group c: meta, deltas[U], counts[U], rel[1024]."""
    s = bytes(s)
    h = header(s)
    G = (h["n"] + 1023) // 1024
    P, M = h["P"], h["M"]
    off = struct.unpack_from("<Q", s, P + 8 * c)[0]
    m0 = P + 8 * G + M * c
    if h["bits"] == 32:
        v = struct.unpack_from("<6fI4H", s, m0)
    else:
        v = struct.unpack_from("<6dQ4H", s, m0)
    mn, mx, outlier = v[0:3], v[3:6], v[6]
    db, cb, rb, U = v[7:11]
    D, C, R = section_sizes(h["bits"], h["dim"], U, db, cb, rb)
    base = P + (8 + M) * G + off
    B = (U + 7) // 8
    deltas = _msb_planes(s[base:base + D], db, B, U)
    counts = _msb_planes(s[base + D:base + D + C], cb, B, U)
    rel = _lsb_planes(s[base + D + C:base + D + C + R], rb)
    return dict(mn=mn, mx=mx, outlier=outlier, db=db, cb=cb, rb=rb, U=U,
                off=off, D=D, C=C, R=R, deltas=deltas, counts=counts, rel=rel)


def total_size(s):
    s = bytes(s)
    h = header(s)
    G = (h["n"] + 1023) // 1024
    last = G - 1
    g = group(s, last)
    return h["P"] + (8 + h["M"]) * G + g["off"] + g["D"] + g["C"] + g["R"]
