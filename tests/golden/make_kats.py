"""Writes tests/golden/kats.json: hand-derived known-answer streams.

The reference repository holds no test vectors, so these are derived by hand
from its source (src/cudaWarpCompress.cu, src/cudaWarpCompressDouble.cu):
each stream below is assembled field by field with struct.pack from the
derivation written next to it, without running any compressor.

    python tests/golden/make_kats.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def hdr32(dim, n, eb, seg):
    # dim u32 | N u64 (offset 4) | float_bits u32 | eb f32 | seg u32
    return struct.pack("<IQIfI", dim, n, 32, eb, seg)


def hdr64(dim, n, eb, seg):
    # dim u32 | N u64 | float_bits u32 | eb f64 (offset 16) | seg u32 | 32
    return struct.pack("<IQIdII", dim, n, 64, eb, seg, 32)


def meta32(mn, mx, outlier, db, cb, rb, U):
    return struct.pack("<6fI4H", *mn, *mx, outlier, db, cb, rb, U)


def meta64(mn, mx, outlier, db, cb, rb, U):
    return struct.pack("<6dQ4H", *mn, *mx, outlier, db, cb, rb, U)


def rel_plane_all_invalid_but(nvalid_zero_bits):
    """128-byte rel plane whose first ranks are 0 and the rest (invalid
    slots, rel = all ones) are 1."""
    bits = [0] * nvalid_zero_bits + [1] * (1024 - nvalid_zero_bits)
    return bytes(sum(bits[8 * t + u] << u for u in range(8)) for t in range(128))


kats = []

# KAT-1: N = 1, point (1,2,3), abs eb 0.5, seg 1, FP32 3D.
#   min = max = point; inv = 1; nbx = nby = 1; q = 0; bid = rel = 0.
#   Invalid slots count 1 bit -> bidbits = 1, rb = 1.  U = 1, count 1 ->
#   cb = 1; delta 0 -> db = 0.  D = 0, C = 4, R = 128.
kat1 = (hdr32(3, 1, 0.5, 1) + struct.pack("<Q", 0) +
        meta32((1, 2, 3), (1, 2, 3), 0, 0, 1, 1, 1) +
        bytes([0x80, 0, 0, 0]) + rel_plane_all_invalid_but(1))
kats.append(dict(name="kat1_f32_3d_n1", bits=32, dim=3, eb=0.5, seg=1,
                 x=[1.0], y=[2.0], z=[3.0], stream=kat1.hex(),
                 decoded=[[1.5], [2.5], [3.5]]))

# KAT-2: N = 2, (0,0,0) and (3,0,0), eb 0.5, seg 1, FP32 3D.
#   nbx = ceil(3/1)+1 = 4, nby = 1; bids 0 and 3; U = 2; deltas {0,3} ->
#   db = 2 (planes 0x40, 0x40); counts {1,1} -> cb = 1 (0xC0); rb = 1.
kat2 = (hdr32(3, 2, 0.5, 1) + struct.pack("<Q", 0) +
        meta32((0, 0, 0), (3, 0, 0), 0, 2, 1, 1, 2) +
        bytes([0x40, 0x40, 0, 0]) + bytes([0xC0, 0, 0, 0]) +
        rel_plane_all_invalid_but(2))
kats.append(dict(name="kat2_f32_3d_n2", bits=32, dim=3, eb=0.5, seg=1,
                 x=[0.0, 3.0], y=[0.0, 0.0], z=[0.0, 0.0], stream=kat2.hex(),
                 decoded=[[0.5, 3.5], [0.5, 0.5], [0.5, 0.5]]))

# KAT-3: same points, FP32 2D: identical sections, meta z extrema 0.
kat3 = (hdr32(2, 2, 0.5, 1) + struct.pack("<Q", 0) +
        meta32((0, 0, 0), (3, 0, 0), 0, 2, 1, 1, 2) +
        bytes([0x40, 0x40, 0, 0]) + bytes([0xC0, 0, 0, 0]) +
        rel_plane_all_invalid_but(2))
kats.append(dict(name="kat3_f32_2d_n2", bits=32, dim=2, eb=0.5, seg=1,
                 x=[0.0, 3.0], y=[0.0, 0.0], z=None, stream=kat3.hex(),
                 decoded=[[0.5, 3.5], [0.5, 0.5], None]))

# KAT-4: KAT-2 in FP64 3D: 32-byte header, 64-byte meta, sections padded to
#   8 bytes: D = 8, C = 8, R = 128 (src/cudaWarpCompressDouble.cu:325-329).
kat4 = (hdr64(3, 2, 0.5, 1) + struct.pack("<Q", 0) +
        meta64((0, 0, 0), (3, 0, 0), 0, 2, 1, 1, 2) +
        bytes([0x40, 0x40] + [0] * 6) + bytes([0xC0] + [0] * 7) +
        rel_plane_all_invalid_but(2))
kats.append(dict(name="kat4_f64_3d_n2", bits=64, dim=3, eb=0.5, seg=1,
                 x=[0.0, 3.0], y=[0.0, 0.0], z=[0.0, 0.0], stream=kat4.hex(),
                 decoded=[[0.5, 3.5], [0.5, 0.5], [0.5, 0.5]]))

# KAT-5: FP64 2D: the count section is padded to 4 bytes only
#   (src/cudaWarpCompressDouble.cu:691): D = 8, C = 4, R = 128.
kat5 = (hdr64(2, 2, 0.5, 1) + struct.pack("<Q", 0) +
        meta64((0, 0, 0), (3, 0, 0), 0, 2, 1, 1, 2) +
        bytes([0x40, 0x40] + [0] * 6) + bytes([0xC0, 0, 0, 0]) +
        rel_plane_all_invalid_but(2))
kats.append(dict(name="kat5_f64_2d_n2", bits=64, dim=2, eb=0.5, seg=1,
                 x=[0.0, 3.0], y=[0.0, 0.0], z=None, stream=kat5.hex(),
                 decoded=[[0.5, 3.5], [0.5, 0.5], None]))

# KAT-6: tie order.  N = 64, eb 0.5, seg 4, point j = (j%4, (j/4)%4, j/16):
#   inv = 1, q = coordinates, every q>>2 = 0 so all 64 points share bid 0,
#   rel = qx + 4qy + 16qz = j.  CUB sorts thread t item i (= slot 32i+t) at
#   position 32t+i, stably; with equal keys the sorted ranks follow that
#   position: slot j < 32 sits at 32j, slot j >= 32 at 32(j-32)+1, so rank
#   2i holds slot i and rank 2i+1 holds slot 32+i.  U = 1, count 64 -> cb = 7
#   (C = 8: 0x80 then zeros), db = 0, rb = 6 (R = 768).
rel = [(r // 2) + 32 * (r % 2) for r in range(64)] + [0xFFFFFFFF] * 960
planes = b""
for p in range(6):
    b = 5 - p
    planes += bytes(sum(((rel[8 * t + u] >> b) & 1) << u for u in range(8))
                    for t in range(128))
xs = [float(j % 4) for j in range(64)]
ys = [float((j // 4) % 4) for j in range(64)]
zs = [float(j // 16) for j in range(64)]
kat6 = (hdr32(3, 64, 0.5, 4) + struct.pack("<Q", 0) +
        meta32((0, 0, 0), (3, 3, 3), 0, 0, 7, 6, 1) +
        bytes([0x80] + [0] * 7) + planes)
order = [(r // 2) + 32 * (r % 2) for r in range(64)]
kats.append(dict(name="kat6_f32_3d_tie_order", bits=32, dim=3, eb=0.5, seg=4,
                 x=xs, y=ys, z=zs, stream=kat6.hex(),
                 decoded=[[xs[j] + 0.5 for j in order],
                          [ys[j] + 0.5 for j in order],
                          [zs[j] + 0.5 for j in order]]))

# KAT-7: two groups, N = 1025, seg 1, eb 0.5.  Group 0: 1024 copies of
#   (0,0,0): bid = rel = 0 everywhere -> bidbits 0 (no sort), rb = 0
#   (R = 0), U = 1, count 1024 -> cb = 11 (C = 12: 0x80 + 11 zeros).
#   Group 1 (last): one point (5,5,5) -> as KAT-1.  Offsets {0, 12}.
g0 = bytes([0x80] + [0] * 11)
g1 = bytes([0x80, 0, 0, 0]) + rel_plane_all_invalid_but(1)
kat7 = (hdr32(3, 1025, 0.5, 1) + struct.pack("<2Q", 0, 12) +
        meta32((0, 0, 0), (0, 0, 0), 0, 0, 11, 0, 1) +
        meta32((5, 5, 5), (5, 5, 5), 0, 0, 1, 1, 1) + g0 + g1)
kats.append(dict(name="kat7_f32_3d_two_groups", bits=32, dim=3, eb=0.5,
                 seg=1, x=[0.0] * 1024 + [5.0], y=[0.0] * 1024 + [5.0],
                 z=[0.0] * 1024 + [5.0], stream=kat7.hex(),
                 decoded=[[0.5] * 1024 + [5.5]] * 3))

if __name__ == "__main__":
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    for k in kats:
        print(k["name"], len(k["stream"]) // 2, "bytes")
