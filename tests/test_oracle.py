"""Pins the CPU oracle (oracle/gpz_oracle.cpp) before it is trusted.

The reference holds no tests or vectors, so every function of the oracle is
pinned by checks that do not restate its own code:

  encode stream layout   hand-derived known-answer streams (golden/kats.json:
                         FP32/FP64, 2D/3D, tie order, two groups, rb = 0)
  quantise (stage 1)     closed form q = floor((x - min) / 2eb) when eb is a
                         power of two (fma and inverse exact)
  block id / rel pos     inverse map: bid, rel decode back to (q >> s, q & m)
  sort (stage 3)         numpy.lexsort over (key, CUB blocked position)
  runs / deltas / counts numpy.unique of the sorted valid ids, read back from
                         the stream with an independent parser (streamfmt)
  bit planes             the same parser, rel per rank = rel[perm]
  decode                 KAT decodes + pointwise error bound through perm
  seg-size heuristic     hand-derived outcomes (cube -> 32 / 2, elongated ->
                         4 / 1, N < 256 -> 4 / 1, guard doubling -> 32,
                         exit case, FP64 never doubles)
  rel -> abs             closed form fl32(max_range * fl32(e))
  stream size            P + (8 + M) G + sum(payload)
Parity unpinned (needs the reference on NVIDIA hardware): the contraction
of the decoder's min + float(2q+1)*eb into an FMA, and the outcome of the
reference's intra-warp races; both are fixed by definition in the oracle.
"""
import numpy as np
import pytest

import streamfmt as SF


def _arr(v, dt):
    return None if v is None else np.array(v, dt)


def test_kats_encode_decode(oracle, kats):
    for k in kats:
        dt = np.float64 if k["bits"] == 64 else np.float32
        x, y, z = _arr(k["x"], dt), _arr(k["y"], dt), _arr(k["z"], dt)
        s = oracle.compress(x, y, z, dt(k["eb"]), k["seg"])
        assert s.tobytes().hex() == k["stream"], k["name"]
        d = oracle.decompress(s)
        for a in range(3):
            if k["decoded"][a] is None:
                assert d[a] is None
            else:
                np.testing.assert_array_equal(d[a], np.array(k["decoded"][a], dt))


def _L():
    j = np.arange(1024)
    return (j % 32) * 32 + j // 32


def _cases():
    # (bits, dim, n, clusters, box, eb, seg)
    return [
        (32, 3, 5000, 3, 1000.0, 0.25, 4),
        (32, 3, 3000, 2, 100.0, 0.5, 32),
        (32, 2, 4097, 4, 50.0, 0.125, 2),
        (64, 3, 2500, 2, 1000.0, 0.25, 8),
        (64, 2, 2049, 3, 10.0, 0.0625, 1),
    ]


@pytest.mark.parametrize("case", _cases())
def test_stages_against_numpy(oracle, case):
    from gpz_amd import datagen
    bits, dim, n, cl, box, eb, seg = case
    x, y, z = datagen.clustered(n, dim=dim, bits=bits, clusters=cl, box=box,
                                seed=11)
    dt = x.dtype.type
    s = oracle.compress(x, y, z, dt(eb), seg)
    G = (n + 1023) // 1024
    L = _L()
    sh = int(np.log2(seg))
    P = [x, y, z]
    for c in range(G):
        st = oracle.group_stages(x, y, z, dt(eb), seg, c)
        nv = st["nvalid"]
        base = 1024 * c
        # stage 1: closed form with exact power-of-two arithmetic
        for a in range(dim):
            v = P[a][base:base + nv].astype(np.float64)
            q = np.floor((v - v.min()) / (2 * eb)).astype(np.uint64)
            np.testing.assert_array_equal(st["q"][a][:nv], q)
        # stage 2: inverse map of block id / relative position
        bid, rel = st["bid"][:nv], st["rel"][:nv]
        nbx, nby = np.uint64(st["nbx"]), np.uint64(st["nby"])
        qx, qy = st["q"][0][:nv], st["q"][1][:nv]
        np.testing.assert_array_equal(bid % nbx, qx >> np.uint64(sh))
        m = np.uint64(seg - 1)
        np.testing.assert_array_equal(rel & m, qx & m)
        np.testing.assert_array_equal((rel >> np.uint64(sh)) & m, qy & m)
        if dim == 3:
            qz = st["q"][2][:nv]
            np.testing.assert_array_equal((bid // nbx) % nby, qy >> np.uint64(sh))
            np.testing.assert_array_equal(bid // (nbx * nby), qz >> np.uint64(sh))
            np.testing.assert_array_equal(rel >> np.uint64(2 * sh), qz & m)
        else:
            np.testing.assert_array_equal(bid // nbx, qy >> np.uint64(sh))
        # stage 3: stable order by key, ties by CUB blocked position
        perm = np.lexsort((L, st["bid"]))
        np.testing.assert_array_equal(st["perm"], perm)
        # stage 4: runs, deltas, counts, planes read back independently
        g = SF.group(s, c)
        sb = st["bid"][perm][:nv]
        uniq, cnt = np.unique(sb, return_counts=True)
        assert g["U"] == len(uniq) == st["U"]
        np.testing.assert_array_equal(g["counts"], cnt)
        np.testing.assert_array_equal(g["deltas"][1:], np.diff(uniq))
        assert g["deltas"][0] == 0 and g["outlier"] == uniq[0]
        rel_sorted = st["rel"][perm]
        exp_rel = np.where(np.arange(1024) < nv, rel_sorted,
                           np.uint64((1 << g["rb"]) - 1))
        np.testing.assert_array_equal(g["rel"], exp_rel)
        assert g["rb"] == max(int(st["rel"][:nv].max()).bit_length(),
                              1 if nv < 1024 else 0)
    assert SF.total_size(s) == len(s)


def _bound_check(x, y, z, eb, seg, oracle):
    """Pointwise |x - x_hat| <= eb (1 + 2^-20) + ulp through the permutation,
    over groups whose block / rel spaces fit the 31-bit masks."""
    dt = x.dtype.type
    s = oracle.compress(x, y, z, dt(eb), seg)
    d = oracle.decompress(s)
    n = x.size
    G = (n + 1023) // 1024
    P = [x, y, z]
    bad = 0
    for c in range(G):
        st = oracle.group_stages(x, y, z, dt(eb), seg, c)
        nv = st["nvalid"]
        perm = st["perm"][:nv].astype(np.int64)
        ok_space = (st["bidbits"] <= 31 and st["rb"] <= 31)
        for a in range(3 if z is not None else 2):
            orig = P[a][1024 * c + perm]
            rec = d[a][1024 * c:1024 * c + nv]
            tol = float(eb) * (1 + 2.0 ** -20) + np.spacing(np.abs(orig)).astype(np.float64)
            err = np.abs(rec.astype(np.float64) - orig.astype(np.float64))
            if ok_space:
                bad += int((err > tol).sum())
    return bad


@pytest.mark.parametrize("bits,dim", [(32, 3), (32, 2), (64, 3), (64, 2)])
def test_round_trip_bound(oracle, bits, dim):
    from gpz_amd import datagen
    x, y, z = datagen.clustered(20_000, dim=dim, bits=bits, clusters=8,
                                box=1000.0, seed=5)
    mm = oracle.min_max(x, y, z)
    dt = x.dtype.type
    eb = dt(oracle.rel_to_abs(dt(1e-4), mm))
    seg, _ = oracle.seg_size(x, y, z, 0, True, dt(1e-4), eb)
    assert _bound_check(x, y, z, eb, seg, oracle) == 0


def test_identical_points_zero_bound(oracle):
    # rel mode over a zero range gives eb = 0: q = NaN -> 0, nb = 1; decode
    # returns the group minimum exactly.
    x = np.full(3000, 7.25, np.float32)
    s, eb, seg, adj = oracle.compress_auto(x, x.copy(), x.copy(), "rel", 1e-4)
    assert eb == 0.0
    d = oracle.decompress(s)
    for a in range(3):
        np.testing.assert_array_equal(d[a], x)


def test_seg_heuristic_outcomes(oracle):
    rng = np.random.default_rng(0)
    n = 200_000
    cube = [rng.random(n, dtype=np.float32) * 100 for _ in range(3)]
    eb = np.float32(100 * 1e-4)
    assert oracle.seg_size(*cube, 0, True, np.float32(1e-4), eb) == (32, False)
    assert oracle.seg_size(*cube, 0, True, np.float32(1e-3), eb * 10) == (2, False)
    slab = [cube[0], cube[1] * 0.01, cube[2] * 0.01]
    assert oracle.seg_size(*slab, 0, True, np.float32(1e-4), eb) == (4, False)
    assert oracle.seg_size(*slab, 0, True, np.float32(1e-3), eb * 10) == (1, False)
    # N < 256: every sample window is skipped and reads zeros -> zero extents,
    # NaN ratio -> the "not cubic" branch.
    tiny = [c[:100] for c in cube]
    assert oracle.seg_size(*tiny, 0, True, np.float32(1e-4), eb)[0] == 4
    # block-count guard: range 100, eb 0.001, seg 1 -> 50001^3 > 2^32; doubling
    # to 32 gives ceil(100/0.064)+1 = 1564, 1564^3 = 3.83e9 <= 2^32 - 1.
    assert oracle.seg_size(*cube, 1, False, np.float32(1e-3), np.float32(0.001)) == (32, True)
    # ... and abs 1e-5 cannot be fixed within 11 doublings -> error.
    with pytest.raises(ValueError):
        oracle.seg_size(*cube, 1, False, np.float32(1e-5), np.float32(1e-5))
    # FP64 compares against UINT64_MAX and never doubles.
    c64 = [c.astype(np.float64) for c in cube]
    assert oracle.seg_size(*c64, 1, False, 1e-5, 1e-5) == (1, False)
    assert oracle.seg_size(*c64, 0, True, 1e-4, 0.01)[0] == 32


def test_rel_to_abs_closed_form(oracle):
    rng = np.random.default_rng(1)
    x, y, z = (rng.normal(size=1000).astype(np.float32) * s for s in (1, 3, 2))
    mm = oracle.min_max(x, y, z)
    assert mm[0] == x.min() and mm[4] == y.max()
    r = max(np.float32(x.max() - x.min()), np.float32(y.max() - y.min()),
            np.float32(z.max() - z.min()))
    assert oracle.rel_to_abs(np.float32(1e-4), mm) == np.float32(r * np.float32(1e-4))
    mm2 = oracle.min_max(x, y, None)
    assert mm2[2] == 0 and mm2[5] == 0


def test_stream_size_formula(oracle):
    from gpz_amd import datagen
    for bits in (32, 64):
        x, y, z = datagen.clustered(7000, bits=bits, clusters=4, seed=2)
        s = oracle.compress(x, y, z, x.dtype.type(0.5), 4)
        h = SF.header(s)
        G = (7000 + 1023) // 1024
        tot = sum(SF.group(s, c)["D"] + SF.group(s, c)["C"] + SF.group(s, c)["R"]
                  for c in range(G))
        assert len(s) == h["P"] + (8 + h["M"]) * G + tot
