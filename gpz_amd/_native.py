"""ctypes binding of libgpz_amd.so (the C ABI declared in include/gpz.h).

There is no fallback: if the HIP library has not been built, importing the
binding raises NativeLibraryMissing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libgpz_amd.so")
DATAGEN_PATH = os.path.join(LIB_DIR, "libgpz_datagen.so")

GPZ_OK = 0
GPZ_ERR_INVALID_ARG = -1
GPZ_ERR_WORKSPACE = -2
GPZ_ERR_CAPACITY = -3
GPZ_ERR_CORRUPT = -4
GPZ_ERR_SEG_TOO_LARGE = -5
GPZ_ERR_HIP = -6
GPZ_MODE_ABS = 0
GPZ_MODE_REL = 1

ERRORS = {
    GPZ_ERR_INVALID_ARG: "invalid argument",
    GPZ_ERR_WORKSPACE: "workspace missing or too small",
    GPZ_ERR_CAPACITY: "output buffer too small",
    GPZ_ERR_CORRUPT: "corrupt stream header",
    GPZ_ERR_SEG_TOO_LARGE: "auto segment size would exceed 1024; "
                           "adjust the error bound",
    GPZ_ERR_HIP: "HIP runtime error",
}

# Every symbol include/gpz.h declares.
EXPORTS = (
    "gpz_version", "gpz_stream_bound", "gpz_compress_workspace_bytes",
    "gpz_compress", "gpz_compress_result", "gpz_parse_header",
    "gpz_read_header", "gpz_stream_bytes", "gpz_decompress", "gpz_min_max_workspace_bytes",
    "gpz_min_max", "gpz_rel_to_abs", "gpz_timing_enable", "gpz_timing_reset",
    "gpz_timing_read",
    "gpz_batch_compress_workspace_bytes", "gpz_batch_decompress_workspace_bytes",
    "gpz_batch_plan_compress", "gpz_batch_plan_decompress", "gpz_batch_run",
    "gpz_batch_compress_results", "gpz_batch_graph_create",
    "gpz_batch_graph_launch", "gpz_batch_graph_destroy",
    "gpz_batch_autotune_compress", "gpz_batch_autotune_decompress",
)


class NativeLibraryMissing(ImportError):
    pass


class GpzError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: {ERRORS.get(code, 'error')} ({code})")


class Params(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_uint32), ("float_bits", ctypes.c_uint32),
                ("n", ctypes.c_uint64), ("eb", ctypes.c_double),
                ("rel_eb", ctypes.c_double), ("mode", ctypes.c_uint32),
                ("seg_size", ctypes.c_uint32)]


class Header(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_uint32), ("float_bits", ctypes.c_uint32),
                ("n", ctypes.c_uint64), ("eb", ctypes.c_double),
                ("seg_size", ctypes.c_uint32),
                ("header_bytes", ctypes.c_uint32),
                ("groups", ctypes.c_uint64),
                ("min_stream_bytes", ctypes.c_uint64)]


class Result(ctypes.Structure):
    _fields_ = [("stream_bytes", ctypes.c_uint64),
                ("seg_size", ctypes.c_uint32), ("adjusted", ctypes.c_uint32),
                ("status", ctypes.c_int32), ("reserved", ctypes.c_uint32)]


class BatchItem(ctypes.Structure):
    _fields_ = [("p", Params), ("x", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("stream", ctypes.c_void_p),
                ("cap", ctypes.c_uint64)]


class BatchDItem(ctypes.Structure):
    _fields_ = [("h", Header), ("stream", ctypes.c_void_p),
                ("bytes", ctypes.c_uint64), ("x", ctypes.c_void_p),
                ("y", ctypes.c_void_p), ("z", ctypes.c_void_p)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("seg_groups", ctypes.c_uint32), ("waves_per_wg", ctypes.c_uint32)]


class BatchPlan(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("kind", ctypes.c_uint32),
                ("seg_groups", ctypes.c_uint32), ("waves_per_wg", ctypes.c_uint32),
                ("cls_items", ctypes.c_uint32 * 4), ("cls_first", ctypes.c_uint32 * 4),
                ("cls_segs", ctypes.c_uint32 * 4), ("cls_seg_first", ctypes.c_uint32 * 4),
                ("grid", ctypes.c_uint32 * 4), ("total_groups", ctypes.c_uint64),
                ("ws_bytes", ctypes.c_uint64)]


def _bind(L):
    vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
    P = ctypes.POINTER
    sig = {
        "gpz_version": ([], i32),
        "gpz_stream_bound": ([P(Params)], u64),
        "gpz_compress_workspace_bytes": ([P(Params)], u64),
        "gpz_compress": ([P(Params), vp, vp, vp, vp, u64, vp, u64, vp], i32),
        "gpz_compress_result": ([vp, P(Result), vp], i32),
        "gpz_parse_header": ([vp, u64, P(Header)], i32),
        "gpz_read_header": ([vp, P(Header), vp], i32),
        "gpz_stream_bytes": ([P(Header), vp, P(u64), vp], i32),
        "gpz_decompress": ([P(Header), vp, u64, vp, vp, vp, vp], i32),
        "gpz_min_max_workspace_bytes": ([], u64),
        "gpz_min_max": ([u32, u32, u64, vp, vp, vp, vp, vp, vp], i32),
        "gpz_rel_to_abs": ([u32, ctypes.c_double, vp], ctypes.c_double),
        "gpz_timing_enable": ([i32], i32),
        "gpz_timing_reset": ([], i32),
        "gpz_timing_read": ([ctypes.c_char_p, P(ctypes.c_double), P(u64)], i32),
        "gpz_batch_compress_workspace_bytes": ([P(BatchItem), u32], u64),
        "gpz_batch_decompress_workspace_bytes": ([P(BatchDItem), u32], u64),
        "gpz_batch_plan_compress": ([P(BatchItem), u32, P(BatchTuning), vp, u64,
                                     P(BatchPlan), vp], i32),
        "gpz_batch_plan_decompress": ([P(BatchDItem), u32, P(BatchTuning), vp, u64,
                                       P(BatchPlan), vp], i32),
        "gpz_batch_run": ([P(BatchPlan), vp, vp], i32),
        "gpz_batch_compress_results": ([P(BatchPlan), vp, P(Result), vp], i32),
        "gpz_batch_graph_create": ([P(BatchPlan), vp, P(vp)], i32),
        "gpz_batch_graph_launch": ([vp, vp], i32),
        "gpz_batch_graph_destroy": ([vp], i32),
        "gpz_batch_autotune_compress": ([P(BatchItem), u32, vp, u64, vp, u32,
                                         P(BatchTuning), P(ctypes.c_float)], i32),
        "gpz_batch_autotune_decompress": ([P(BatchDItem), u32, vp, u64, vp, u32,
                                           P(BatchTuning), P(ctypes.c_float)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


_lib = None


def lib():
    """The loaded HIP library; raises NativeLibraryMissing if not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing: build it with "
                "`python -m gpz_amd.build` (there is no CPU fallback)")
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


_datagen = None


def datagen():
    global _datagen
    if _datagen is None:
        if not os.path.exists(DATAGEN_PATH):
            raise NativeLibraryMissing(f"{DATAGEN_PATH} is missing")
        L = ctypes.CDLL(DATAGEN_PATH)
        L.gpz_gen_clustered.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.gpz_gen_clustered.restype = ctypes.c_int
        L.gpz_gen_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64,
                                      ctypes.c_uint64]
        L.gpz_gen_uniform.restype = ctypes.c_double
        _datagen = L
    return _datagen


def check(rc, what):
    if rc != GPZ_OK:
        raise GpzError(rc, what)
