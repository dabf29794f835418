"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: byte-identical streams, bit-identical decoded coordinates, identical
segment-size decisions, on the same seeded inputs; plus the known-answer
streams of tests/golden/kats.json.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _first_diff(a, b):
    n = min(a.size, b.size)
    d = np.nonzero(a[:n] != b[:n])[0]
    return int(d[0]) if d.size else (n if a.size != b.size else -1)


def _parity(oracle, x, y, z, mode, e, seg_in=0, check_decode=True):
    from gpz_amd import api
    s_o, eb, seg_o, adj_o = oracle.compress_auto(x, y, z, mode, e, seg_in)
    r = api.compress(_dev(x), _dev(y), _dev(z), mode=mode, e=e, seg_size=seg_in)
    assert r.eb == float(eb)
    assert (r.seg_size, r.adjusted) == (seg_o, adj_o)
    s_g = r.to_numpy()
    assert _first_diff(s_g, s_o) == -1, f"stream differs at byte {_first_diff(s_g, s_o)}"
    d_g = None
    if check_decode:
        d_g = api.decompress(r.bytes())
        d_o = oracle.decompress(s_o)
        for a in range(3 if z is not None else 2):
            np.testing.assert_array_equal(d_g[a].cpu().numpy(), d_o[a])
    return s_o, r, d_g


def test_native_library_is_in_tree():
    from gpz_amd import _native as N
    L = N.lib()
    assert L.gpz_version() >= 100
    assert torch.cuda.is_available()


def test_kats_on_gpu(kats):
    from gpz_amd import api
    for k in kats:
        dt = np.float64 if k["bits"] == 64 else np.float32
        x = np.array(k["x"], dt)
        y = np.array(k["y"], dt)
        z = np.array(k["z"], dt) if k["z"] is not None else None
        r = api.compress(_dev(x), _dev(y), _dev(z), mode="abs", e=k["eb"],
                         seg_size=k["seg"])
        assert r.to_numpy().tobytes().hex() == k["stream"], k["name"]
        d = api.decompress(r.bytes())
        for a in range(3):
            if k["decoded"][a] is not None:
                np.testing.assert_array_equal(d[a].cpu().numpy(),
                                              np.array(k["decoded"][a], dt))


def test_small_config_full(oracle):
    """1,048,576 clustered FP32 3D points, seed 0, rel 1e-4, auto seg."""
    from gpz_amd import datagen
    x, y, z, spec = datagen.workload("small_f32_3d")
    s, r, d = _parity(oracle, x, y, z, "rel", 1e-4)
    assert len(s) < x.nbytes * 3
    for orig, dec in zip((x, y, z), d):
        assert _group_sorted_bound(_dev(orig), dec, r.eb) == 0


@pytest.mark.parametrize("n", [1, 2, 100, 255, 256, 257, 1023, 1024, 1025,
                               1279, 1280, 1281, 4096, 70_001])
def test_sizes_f32_3d(oracle, n):
    from gpz_amd import datagen
    x, y, z = datagen.clustered(n, clusters=max(1, n // 2048), seed=n)
    _parity(oracle, x, y, z, "rel", 1e-4)


@pytest.mark.parametrize("bits,dim,mode,e,seg", [
    (32, 2, "rel", 1e-5, 0),
    (32, 2, "rel", 1e-3, 0),
    (32, 3, "rel", 1e-3, 0),
    (32, 3, "abs", 0.05, 64),
    (32, 3, "abs", 0.5, 1),
    (32, 3, "abs", 0.5, 3),
    (32, 3, "abs", 2.0, 7),
    (32, 3, "abs", 0.01, 1024),
    (64, 3, "abs", 1e-3, 64),
    (64, 3, "rel", 1e-4, 0),
    (64, 2, "rel", 1e-4, 0),
    (64, 2, "abs", 0.01, 16),
])
def test_variants(oracle, bits, dim, mode, e, seg):
    from gpz_amd import datagen
    n = 150_000
    x, y, z = datagen.clustered(n, dim=dim, bits=bits, clusters=64,
                                box=1000.0 if dim == 3 else 1.0,
                                sigma=0.0 if dim == 3 else 1e-3,
                                frac=0.7, seed=bits + dim + seg)
    _parity(oracle, x, y, z, mode, e, seg)


def test_seg_guard_doubles_on_gpu(oracle):
    rng = np.random.default_rng(3)
    x, y, z = (rng.random(50_000, dtype=np.float32) * 100 for _ in range(3))
    _parity(oracle, x, y, z, "abs", 1e-3, 1)


def test_seg_too_large_reports_error(oracle):
    from gpz_amd import api, GpzError
    rng = np.random.default_rng(4)
    x, y, z = (rng.random(50_000, dtype=np.float32) * 100 for _ in range(3))
    with pytest.raises(GpzError):
        api.compress(_dev(x), _dev(y), _dev(z), mode="abs", e=1e-5, seg_size=1)


def test_identical_points(oracle):
    x = np.full(5000, -3.5, np.float32)
    _parity(oracle, x, x.copy(), x.copy(), "rel", 1e-4)


def test_negative_and_offset_coordinates(oracle):
    from gpz_amd import datagen
    x, y, z = datagen.clustered(100_000, clusters=32, seed=21)
    _parity(oracle, x - 1e4, -y, z * 0.001 + 5e3, "rel", 1e-4)


@pytest.mark.parametrize("bits,dim,n", [
    (32, 3, 3 * 1048576 + 5000),
    (32, 3, 8 * 1048576 + 777),
    (64, 3, 4 * 1048576 + 1),
    (32, 2, 17 * 1048576 - 3),
])
def test_multi_million_points(oracle, bits, dim, n):
    """Millions of points (thousands of groups, ragged last group): payload
    offsets accumulate over many groups."""
    from gpz_amd import datagen
    x, y, z = datagen.clustered(n, dim=dim, bits=bits, clusters=256, seed=n % 1000)
    _parity(oracle, x, y, z, "rel", 1e-4)


def _group_sorted_bound(orig, dec, eb):
    """Per group and axis, sorted original vs sorted decoded within the bound
    (holds whenever every point is within the bound)."""
    n = orig.numel()
    G = (n + 1023) // 1024
    pad = G * 1024 - n
    big = torch.finfo(orig.dtype).max
    o = torch.cat([orig, orig.new_full((pad,), big)]).view(G, 1024)
    d = torch.cat([dec, dec.new_full((pad,), big)]).view(G, 1024)
    o, _ = torch.sort(o, dim=1)
    d, _ = torch.sort(d, dim=1)
    err = (o.double() - d.double()).abs()
    tol = eb * (1 + 2.0 ** -20) + 2 * torch.finfo(orig.dtype).eps * o.double().abs()
    return int((err > tol).sum())


@pytest.mark.slow
def test_large_config_full(oracle):
    """143,735,721 FP32 3D points, rel 1e-4: byte parity and bound."""
    from gpz_amd import datagen
    x, y, z, spec = datagen.workload("large_f32_3d")
    s_o, r, d = _parity(oracle, x, y, z, "rel", 1e-4)
    for orig, dec in zip((x, y, z), d):
        assert _group_sorted_bound(_dev(orig), dec, r.eb) == 0
    assert len(s_o) == r.nbytes


@pytest.mark.parametrize("bits,dim", [(32, 3), (32, 2), (64, 3), (64, 2)])
@pytest.mark.parametrize("n", [1, 1023, 1025, 3 * 1024, 5000])
def test_decode_stays_inside_outputs(oracle, bits, dim, n):
    """Decoding into slices of a sentinel-filled buffer: every point of the
    output is written, nothing past its end (the last group may be
    partial)."""
    import ctypes
    from gpz_amd import api, datagen, _native as N
    x, y, z = datagen.clustered(n, dim=dim, bits=bits, clusters=8, seed=n)
    zz = z if dim == 3 else None
    r = api.compress(_dev(x), _dev(y), _dev(zz), mode="rel", e=1e-4)
    hdr = api.read_header(r.bytes())
    dt = torch.float64 if bits == 64 else torch.float32
    pad = 4096
    big = torch.full((3 * (n + pad),), -12345.0, dtype=dt, device="cuda")
    outs = [big[i * (n + pad): i * (n + pad) + n] for i in range(3)]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = N.lib().gpz_decompress(ctypes.byref(hdr), ctypes.c_void_p(r.bytes().data_ptr()),
                                ctypes.c_uint64(r.nbytes),
                                *(ctypes.c_void_p(o.data_ptr()) for o in outs[:dim]),
                                *([ctypes.c_void_p(0)] if dim == 2 else []), s)
    torch.cuda.synchronize()
    assert rc == 0
    b = big.cpu().numpy()
    d_o = oracle.decompress(r.to_numpy())
    for a in range(3):
        seg = b[a * (n + pad): (a + 1) * (n + pad)]
        assert (seg[n:] == -12345.0).all(), f"axis {a}: write past the end"
        if a < dim:
            np.testing.assert_array_equal(seg[:n], d_o[a])
        else:
            assert (seg[:n] == -12345.0).all()


@pytest.mark.parametrize("bits,dim", [(32, 3), (64, 2)])
def test_all_distinct_blocks(oracle, bits, dim):
    """Every point in its own block (U = points of the group): the
    decoder's rank-is-run path, full and partial groups."""
    rng = np.random.default_rng(5)
    n = 4 * 1024 + 333
    dt = np.float64 if bits == 64 else np.float32
    x, y, z = (rng.uniform(0, 1000, n).astype(dt) for _ in range(3))
    zz = z if dim == 3 else None
    s_o, r, d = _parity(oracle, x, y, zz, "abs", 1e-3, seg_in=1)
    from gpz_amd import api
    h = api.parse_header(s_o)
    assert h.groups == 5


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("dim,e,seg,sortbits", [
    (3, 1.0, 2, 24),      # pair pass over 4 low bits, then packed passes
    (3, 0.4, 2, 28),      # 8 low bits
    (2, 0.0076, 4, 28),   # 2D, 4 rel bits
    (3, 0.02, 4, 31),     # 12 low bits (31-bit masked block ids)
])
def test_wide_sort_keys(oracle, bits, dim, e, seg, sortbits):
    """Groups whose block ids need 23-32 sort bits (the encoder's pair-then-
    packed sort), uniform in [0, 1000)^dim, with a partial last group."""
    rng = np.random.default_rng(sortbits * 10 + dim)
    n = 5 * 1024 + 77
    dt = np.float64 if bits == 64 else np.float32
    pts = [rng.uniform(0, 1000, n).astype(dt) for _ in range(dim)]
    x, y = pts[0], pts[1]
    z = pts[2] if dim == 3 else None
    # the workload reaches the intended sort width in its first group
    nb = math.ceil((float(x[:1024].max() - x[:1024].min())) / (2 * seg * e)) + 1
    assert (nb ** dim).bit_length() >= sortbits - 1
    _parity(oracle, x, y, z, "abs", e, seg_in=seg)
