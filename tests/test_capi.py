"""C-ABI library: loads, exports every symbol include/gpz.h declares, parses
headers and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from gpz_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "gpz.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gpz_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    decl = _declared()
    assert len(decl) >= 11
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(N.EXPORTS)
    assert L.gpz_version() >= 100


def test_parse_header_kats(kats):
    for k in kats:
        b = np.frombuffer(bytes.fromhex(k["stream"]), np.uint8)
        h = N.Header()
        assert N.lib().gpz_parse_header(b.ctypes.data_as(ctypes.c_void_p),
                                        b.size, ctypes.byref(h)) == N.GPZ_OK
        assert (h.dim, h.float_bits, h.seg_size) == (k["dim"], k["bits"], k["seg"])
        assert h.n == len(k["x"]) and h.eb == k["eb"]
        assert h.header_bytes == (32 if k["bits"] == 64 else 24)
        assert h.groups == (h.n + 1023) // 1024


@pytest.mark.parametrize("field,value", [(0, 5), (12, 16), (4, 0)])
def test_parse_header_rejects_corrupt(kats, field, value):
    b = bytearray(bytes.fromhex(kats[0]["stream"]))
    b[field:field + 4] = value.to_bytes(4, "little")
    if field == 4:
        b[8:12] = bytes(4)
    a = np.frombuffer(bytes(b), np.uint8)
    h = N.Header()
    assert N.lib().gpz_parse_header(a.ctypes.data_as(ctypes.c_void_p), a.size,
                                    ctypes.byref(h)) == N.GPZ_ERR_CORRUPT


def _params(**kw):
    p = N.Params()
    d = dict(dim=3, float_bits=32, n=1000, eb=0.1, rel_eb=1e-4,
             mode=N.GPZ_MODE_REL, seg_size=0)
    d.update(kw)
    for k, v in d.items():
        setattr(p, k, v)
    return p


def test_bounds_and_workspace_sizes():
    L = N.lib()
    p = _params(n=143_735_721)
    G = 140_367
    assert L.gpz_stream_bound(ctypes.byref(p)) == 24 + 44 * G + 9344 * G
    p64 = _params(n=1 << 26, float_bits=64)
    assert L.gpz_stream_bound(ctypes.byref(p64)) == 32 + 72 * 65536 + 9344 * 65536
    assert L.gpz_compress_workspace_bytes(ctypes.byref(p)) >= 9344 * G


@pytest.mark.parametrize("kw", [dict(dim=4), dict(float_bits=16), dict(n=0),
                                dict(mode=N.GPZ_MODE_ABS, seg_size=0),
                                dict(eb=-1.0), dict(eb=float("nan"))])
def test_compress_rejects_bad_params_without_gpu(kw):
    p = _params(**kw)
    dummy = ctypes.c_void_p(16)
    rc = N.lib().gpz_compress(ctypes.byref(p), dummy, dummy, dummy, dummy,
                              1 << 40, dummy, 1 << 40, None)
    assert rc == N.GPZ_ERR_INVALID_ARG


def test_compress_rejects_small_buffers_without_gpu():
    p = _params()
    dummy = ctypes.c_void_p(16)
    L = N.lib()
    assert L.gpz_compress(ctypes.byref(p), dummy, dummy, dummy, dummy, 1 << 40,
                          dummy, 10, None) == N.GPZ_ERR_WORKSPACE
    assert L.gpz_compress(ctypes.byref(p), dummy, dummy, dummy, dummy, 10,
                          dummy, 1 << 40, None) == N.GPZ_ERR_CAPACITY


def test_rel_to_abs_matches_reference_arithmetic():
    mm = np.array([0, -1, 2, 10, 4, 3], np.float32)
    e = np.float32(1e-4)
    got = N.lib().gpz_rel_to_abs(32, float(e), mm.ctypes.data_as(ctypes.c_void_p))
    assert np.float32(got) == np.float32(10) * e
    mm64 = mm.astype(np.float64)
    got = N.lib().gpz_rel_to_abs(64, 1e-4, mm64.ctypes.data_as(ctypes.c_void_p))
    assert got == 10.0 * 1e-4


def test_datagen_deterministic():
    from gpz_amd import datagen
    a = datagen.clustered(50_000, clusters=16, seed=9)
    b = datagen.clustered(50_000, clusters=16, seed=9)
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u, v)
    c = datagen.clustered(50_000, clusters=16, seed=10)
    assert not np.array_equal(a[0], c[0])
    # uniform tail lies in the box
    tail = a[0][int(50_000 * 0.7):]
    assert tail.min() >= 0 and tail.max() < 1000
    spec = datagen.batch_mix()
    assert len(spec) == 256
    assert all(1e4 * 0.99 <= s["n"] <= 4e6 * 1.01 for s in spec)
