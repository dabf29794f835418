"""ctypes binding of the CPU oracle (oracle/gpz_oracle.cpp).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only to check or time the GPU path.
The product library (gpz_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s",
                        "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, u32, i32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        vp = ctypes.c_void_p
        L.oracle_num_threads.argtypes = [i32]
        L.oracle_num_threads.restype = i32
        L.oracle_stream_bound.argtypes = [i32, u64]
        L.oracle_stream_bound.restype = u64
        for t, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            getattr(L, f"oracle_min_max_{t}").argtypes = [i32, u64, vp, vp, vp, vp]
            getattr(L, f"oracle_min_max_{t}").restype = None
            getattr(L, f"oracle_rel_to_abs_{t}").argtypes = [ct, vp]
            getattr(L, f"oracle_rel_to_abs_{t}").restype = ct
            getattr(L, f"oracle_seg_size_{t}").argtypes = [
                i32, u64, vp, vp, vp, u32, i32, ct, ct,
                ctypes.POINTER(u32), ctypes.POINTER(i32)]
            getattr(L, f"oracle_seg_size_{t}").restype = i32
            getattr(L, f"oracle_compress_{t}").argtypes = [
                i32, u64, vp, vp, vp, ct, u32, vp, u64, ctypes.POINTER(u64)]
            getattr(L, f"oracle_compress_{t}").restype = i32
            getattr(L, f"oracle_decompress_{t}").argtypes = [vp, u64, vp, vp, vp]
            getattr(L, f"oracle_decompress_{t}").restype = i32
            getattr(L, f"oracle_group_stages_{t}").argtypes = [
                i32, u64, vp, vp, vp, ct, u32, u64, vp, vp, vp, vp, vp]
            getattr(L, f"oracle_group_stages_{t}").restype = i32
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _tag(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def num_threads(want=0):
    return lib().oracle_num_threads(int(want))


def min_max(x, y, z=None):
    dim = 3 if z is not None else 2
    mm = np.zeros(6, dtype=x.dtype)
    getattr(lib(), f"oracle_min_max_{_tag(x.dtype)}")(
        dim, x.size, _p(x), _p(y), _p(z), _p(mm))
    return mm


def rel_to_abs(e, mm):
    mm = np.ascontiguousarray(mm)
    return getattr(lib(), f"oracle_rel_to_abs_{_tag(mm.dtype)}")(e, _p(mm))


def seg_size(x, y, z, seg_in, is_rel, rel_eb, eb):
    """Returns (seg, adjusted); raises on the reference's 'too large' exit."""
    dim = 3 if z is not None else 2
    seg = ctypes.c_uint32(0)
    adj = ctypes.c_int(0)
    rc = getattr(lib(), f"oracle_seg_size_{_tag(x.dtype)}")(
        dim, x.size, _p(x), _p(y), _p(z), int(seg_in), int(bool(is_rel)),
        rel_eb, eb, ctypes.byref(seg), ctypes.byref(adj))
    if rc != 0:
        raise ValueError("segment size too large for this error bound")
    return seg.value, bool(adj.value)


def compress(x, y, z, eb, seg):
    """Stream bytes (np.uint8) for SoA planes x, y[, z] at absolute eb."""
    dim = 3 if z is not None else 2
    tag = _tag(x.dtype)
    cap = lib().oracle_stream_bound(64 if tag == "f64" else 32, x.size)
    out = np.empty(cap, dtype=np.uint8)
    size = ctypes.c_uint64(0)
    rc = getattr(lib(), f"oracle_compress_{tag}")(
        dim, x.size, _p(x), _p(y), _p(z), eb, int(seg), _p(out), cap,
        ctypes.byref(size))
    if rc != 0:
        raise ValueError(f"oracle compress failed ({rc})")
    return out[: size.value].copy()


def compress_auto(x, y, z, mode, e, seg_in=0):
    """Reference CLI semantics: rel -> abs over the full data, auto seg."""
    dt = x.dtype.type
    if mode == "rel":
        mm = min_max(x, y, z)
        eb = dt(rel_to_abs(dt(e), mm))
        rel_eb = dt(e)
    else:
        eb = dt(e)
        rel_eb = dt(e)
    seg, adj = seg_size(x, y, z, seg_in, mode == "rel", rel_eb, eb)
    return compress(x, y, z, eb, seg), eb, seg, adj


def parse_header(s):
    dim = int(np.frombuffer(s[0:4].tobytes(), "<u4")[0])
    n = int(np.frombuffer(s[4:12].tobytes(), "<u8")[0])
    bits = int(np.frombuffer(s[12:16].tobytes(), "<u4")[0])
    if bits == 64:
        eb = float(np.frombuffer(s[16:24].tobytes(), "<f8")[0])
        seg = int(np.frombuffer(s[24:28].tobytes(), "<u4")[0])
    else:
        eb = float(np.frombuffer(s[16:20].tobytes(), "<f4")[0])
        seg = int(np.frombuffer(s[20:24].tobytes(), "<u4")[0])
    return dim, n, bits, eb, seg


def decompress(s):
    dim, n, bits, _, _ = parse_header(s)
    dt = np.float64 if bits == 64 else np.float32
    x = np.zeros(n, dt)
    y = np.zeros(n, dt)
    z = np.zeros(n, dt) if dim == 3 else None
    s = np.ascontiguousarray(s)
    rc = getattr(lib(), f"oracle_decompress_{_tag(dt)}")(
        _p(s), s.size, _p(x), _p(y), _p(z))
    if rc != 0:
        raise ValueError(f"oracle decompress failed ({rc})")
    return x, y, z


def group_stages(x, y, z, eb, seg, c):
    """Per-slot stage outputs of group c: dict of q, bid, rel, perm, info."""
    dim = 3 if z is not None else 2
    q = np.zeros(3 * 1024, np.uint64)
    bid = np.zeros(1024, np.uint64)
    rel = np.zeros(1024, np.uint64)
    perm = np.zeros(1024, np.uint16)
    info = np.zeros(9, np.uint32)
    rc = getattr(lib(), f"oracle_group_stages_{_tag(x.dtype)}")(
        dim, x.size, _p(x), _p(y), _p(z), eb, int(seg), int(c), _p(q),
        _p(bid), _p(rel), _p(perm), _p(info))
    if rc != 0:
        raise ValueError("bad group")
    keys = ("nbx", "nby", "shift", "bidbits", "rb", "U", "db", "cb", "nvalid")
    return {"q": q.reshape(3, 1024), "bid": bid, "rel": rel, "perm": perm,
            **{k: int(v) for k, v in zip(keys, info)}}
