"""Python API over the C ABI, for device-resident torch tensors.

Functions take SoA coordinate planes (1-D CUDA tensors of float32/float64)
and byte streams as uint8 CUDA tensors.  Work is issued on the current torch
stream; calls that return host values synchronise that stream.
"""
import ctypes

import numpy as np
import torch

from . import _native as N


def _vp(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(s=None):
    s = s if s is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _bits(t):
    if t.dtype == torch.float32:
        return 32
    if t.dtype == torch.float64:
        return 64
    raise TypeError(f"coordinates must be float32 or float64, got {t.dtype}")


def _check_planes(x, y, z):
    for t in (x, y) + ((z,) if z is not None else ()):
        if not t.is_cuda or t.dim() != 1 or not t.is_contiguous():
            raise ValueError("planes must be contiguous 1-D CUDA tensors")
        if t.dtype != x.dtype or t.numel() != x.numel():
            raise ValueError("planes must share dtype and length")


def make_params(dim, float_bits, n, eb, rel_eb, mode, seg_size=0):
    p = N.Params()
    p.dim, p.float_bits, p.n = int(dim), int(float_bits), int(n)
    p.eb, p.rel_eb = float(eb), float(rel_eb)
    p.mode = N.GPZ_MODE_REL if mode == "rel" else N.GPZ_MODE_ABS
    p.seg_size = int(seg_size)
    return p


def stream_bound(p):
    return int(N.lib().gpz_stream_bound(ctypes.byref(p)))


def workspace_bytes(p):
    return int(N.lib().gpz_compress_workspace_bytes(ctypes.byref(p)))


def min_max(x, y, z=None, stream=None):
    """Full-data extrema {min_x,min_y,min_z,max_x,max_y,max_z} (numpy)."""
    _check_planes(x, y, z)
    bits = _bits(x)
    ws = torch.empty(int(N.lib().gpz_min_max_workspace_bytes()),
                     dtype=torch.uint8, device=x.device)
    mm = np.zeros(6, dtype=np.float64 if bits == 64 else np.float32)
    N.check(N.lib().gpz_min_max(3 if z is not None else 2, bits, x.numel(),
                                _vp(x), _vp(y), _vp(z), _vp(ws),
                                mm.ctypes.data_as(ctypes.c_void_p),
                                _stream(stream)), "gpz_min_max")
    return mm


def rel_to_abs(float_bits, e, mm):
    mm = np.ascontiguousarray(mm, dtype=np.float64 if float_bits == 64 else np.float32)
    return float(N.lib().gpz_rel_to_abs(float_bits, float(e),
                                        mm.ctypes.data_as(ctypes.c_void_p)))


def compress_into(p, x, y, z, out, ws, stream=None):
    """Asynchronous compress into preallocated `out` / `ws` (uint8 tensors)."""
    N.check(N.lib().gpz_compress(ctypes.byref(p), _vp(x), _vp(y), _vp(z),
                                 _vp(out), out.numel(), _vp(ws), ws.numel(),
                                 _stream(stream)), "gpz_compress")


def compress_result(ws, stream=None):
    r = N.Result()
    N.check(N.lib().gpz_compress_result(_vp(ws), ctypes.byref(r),
                                        _stream(stream)), "gpz_compress_result")
    return r


class Compressed:
    def __init__(self, data, nbytes, eb, seg_size, adjusted):
        self.data = data            # uint8 CUDA tensor (capacity >= nbytes)
        self.nbytes = nbytes
        self.eb = eb
        self.seg_size = seg_size
        self.adjusted = adjusted

    def bytes(self):
        return self.data[: self.nbytes]

    def to_numpy(self):
        return self.bytes().cpu().numpy()


def resolve_eb(x, y, z, mode, e):
    """Absolute bound exactly as the reference CLI derives it."""
    bits = _bits(x)
    if mode == "rel":
        return rel_to_abs(bits, e, min_max(x, y, z))
    return float(np.float32(e)) if bits == 32 else float(e)


def compress(x, y, z=None, mode="rel", e=1e-4, seg_size=0, eb=None,
             stream=None):
    """Compress planes; returns a Compressed (synchronises the stream)."""
    _check_planes(x, y, z)
    bits = _bits(x)
    if eb is None:
        eb = resolve_eb(x, y, z, mode, e)
    rel = float(np.float32(e)) if bits == 32 else float(e)
    p = make_params(3 if z is not None else 2, bits, x.numel(), eb, rel,
                    mode, seg_size)
    out = torch.empty(stream_bound(p), dtype=torch.uint8, device=x.device)
    ws = torch.empty(workspace_bytes(p), dtype=torch.uint8, device=x.device)
    compress_into(p, x, y, z, out, ws, stream)
    r = compress_result(ws, stream)
    N.check(r.status, "gpz_compress (device status)")
    return Compressed(out, int(r.stream_bytes), eb, int(r.seg_size),
                      bool(r.adjusted))


def parse_header(host_bytes):
    b = np.ascontiguousarray(np.frombuffer(bytes(host_bytes[:32]), np.uint8))
    h = N.Header()
    N.check(N.lib().gpz_parse_header(b.ctypes.data_as(ctypes.c_void_p),
                                     b.size, ctypes.byref(h)),
            "gpz_parse_header")
    return h


def read_header(stream_tensor, stream=None):
    h = N.Header()
    N.check(N.lib().gpz_read_header(_vp(stream_tensor), ctypes.byref(h),
                                    _stream(stream)), "gpz_read_header")
    return h


def decompress_into(h, stream_tensor, nbytes, x, y, z=None, stream=None):
    N.check(N.lib().gpz_decompress(ctypes.byref(h), _vp(stream_tensor),
                                   int(nbytes), _vp(x), _vp(y), _vp(z),
                                   _stream(stream)), "gpz_decompress")


def decompress(stream_tensor, nbytes=None, header=None, stream=None):
    """Decompress a device stream; returns (x, y, z) CUDA tensors."""
    h = header if header is not None else read_header(stream_tensor, stream)
    nbytes = stream_tensor.numel() if nbytes is None else nbytes
    dt = torch.float64 if h.float_bits == 64 else torch.float32
    dev = stream_tensor.device
    x = torch.empty(h.n, dtype=dt, device=dev)
    y = torch.empty(h.n, dtype=dt, device=dev)
    z = torch.empty(h.n, dtype=dt, device=dev) if h.dim == 3 else None
    decompress_into(h, stream_tensor, nbytes, x, y, z, stream)
    return x, y, z
