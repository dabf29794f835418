"""gpz_amd -- MI355X-native GPZ error-bounded lossy particle compressor.

The hot path (quantise, block id, LDS radix sort, delta/bit-plane encode and
the inverse) runs as hand-written HIP kernels for gfx950 in
gpz_amd/lib/libgpz_amd.so behind the C ABI of include/gpz.h.  This package
binds that ABI for torch device tensors and mirrors the reference's
ICudaCompressor / CompressorManager interface (gpz_amd.compressor).
"""
from . import _native  # noqa: F401
from ._native import GpzError, NativeLibraryMissing  # noqa: F401

__all__ = ["GpzError", "NativeLibraryMissing", "api", "compressor", "datagen"]


def __getattr__(name):
    if name in ("api", "compressor", "datagen", "build"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
