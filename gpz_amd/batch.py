"""Batched compression and decompression over the C ABI (include/gpz.h,
gpz_batch_*): many snapshots of uneven size, dimension and precision per
launch, with device-side segment queues, hipGraph replay and an autotuner.

The reference compresses one snapshot per call; a batch writes, for every
snapshot, the same stream that `api.compress` writes for it alone.

    b = BatchCompressor([(x0, y0, z0), (x1, y1, None)], mode="rel", e=1e-4)
    b.run()                       # asynchronous, current torch stream
    streams = b.streams()         # uint8 CUDA tensors, one per snapshot
    g = b.graph(); g.launch()     # the same work as one hipGraph launch
"""
import ctypes

import numpy as np
import torch

from . import _native as N
from . import api


def _tuning(t):
    if t is None:
        return None
    if isinstance(t, N.BatchTuning):
        return t
    seg_groups, waves = t
    r = N.BatchTuning()
    r.seg_groups, r.waves_per_wg = int(seg_groups), int(waves)
    return r


def _ptr(t):
    return None if t is None else t.data_ptr()


class BatchGraph:
    """A captured hipGraph of one planned batch (gpz_batch_graph_*)."""

    def __init__(self, plan, ws):
        self._h = ctypes.c_void_p()
        N.check(N.lib().gpz_batch_graph_create(ctypes.byref(plan), api._vp(ws),
                                               ctypes.byref(self._h)),
                "gpz_batch_graph_create")

    def launch(self, stream=None):
        N.check(N.lib().gpz_batch_graph_launch(self._h, api._stream(stream)),
                "gpz_batch_graph_launch")

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.lib().gpz_batch_graph_destroy(self._h)
            self._h = ctypes.c_void_p()


class BatchCompressor:
    """Plans the compression of `snapshots`, each (x, y, z) with z None in
    2D, or (x, y, z, mode, e, seg_size) to override the defaults.  The
    absolute bound of every rel-mode snapshot is resolved once here (as the
    reference does on the host before compressing); eb= overrides it."""

    def __init__(self, snapshots, mode="rel", e=1e-4, seg_size=0, tuning=None,
                 stream=None):
        dev = snapshots[0][0].device
        self.planes, self.params, self.ebs = [], [], []
        for snap in snapshots:
            x, y, z = snap[:3]
            m, ee, sg = (snap[3], snap[4], snap[5]) if len(snap) > 3 else (mode, e, seg_size)
            api._check_planes(x, y, z)
            bits = api._bits(x)
            eb = api.resolve_eb(x, y, z, m, ee)
            rel = float(np.float32(ee)) if bits == 32 else float(ee)
            p = api.make_params(3 if z is not None else 2, bits, x.numel(), eb, rel, m, sg)
            self.planes.append((x, y, z))
            self.params.append(p)
            self.ebs.append(eb)
        caps = [api.stream_bound(p) for p in self.params]
        self.offsets = [0]
        for c in caps:
            self.offsets.append(self.offsets[-1] + ((c + 255) // 256) * 256)
        self.out = torch.empty(self.offsets[-1], dtype=torch.uint8, device=dev)
        items = (N.BatchItem * len(snapshots))()
        for i, (p, (x, y, z)) in enumerate(zip(self.params, self.planes)):
            it = items[i]
            it.p = p
            it.x, it.y, it.z = _ptr(x), _ptr(y), _ptr(z)
            it.stream = self.out.data_ptr() + self.offsets[i]
            it.cap = caps[i]
        self.items = items
        L = N.lib()
        wsb = int(L.gpz_batch_compress_workspace_bytes(items, len(snapshots)))
        self.ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        self.plan = N.BatchPlan()
        self.replan(tuning, stream)

    def __len__(self):
        return len(self.params)

    def replan(self, tuning=None, stream=None):
        t = _tuning(tuning)
        N.check(N.lib().gpz_batch_plan_compress(
            self.items, len(self), ctypes.byref(t) if t else None, api._vp(self.ws),
            self.ws.numel(), ctypes.byref(self.plan), api._stream(stream)),
            "gpz_batch_plan_compress")

    def run(self, stream=None):
        N.check(N.lib().gpz_batch_run(ctypes.byref(self.plan), api._vp(self.ws),
                                      api._stream(stream)), "gpz_batch_run")

    def graph(self):
        return BatchGraph(self.plan, self.ws)

    def results(self, stream=None):
        res = (N.Result * len(self))()
        N.check(N.lib().gpz_batch_compress_results(ctypes.byref(self.plan),
                                                   api._vp(self.ws), res,
                                                   api._stream(stream)),
                "gpz_batch_compress_results")
        return list(res)

    def streams(self, stream=None):
        out = []
        for i, r in enumerate(self.results(stream)):
            N.check(r.status, f"snapshot {i}")
            out.append(self.out[self.offsets[i]: self.offsets[i] + r.stream_bytes])
        return out

    def autotune(self, reps=3, stream=None):
        """Times every (groups per segment, waves per workgroup) pair and
        keeps the fastest.  Returns ((seg_groups, waves), table) with table
        [i][j] = ms per run at seg_groups 2^i, waves 2^j."""
        best = N.BatchTuning()
        table = (ctypes.c_float * 16)()
        N.check(N.lib().gpz_batch_autotune_compress(
            self.items, len(self), api._vp(self.ws), self.ws.numel(),
            api._stream(stream), int(reps), ctypes.byref(best), table),
            "gpz_batch_autotune_compress")
        self.replan(best, stream)
        return (best.seg_groups, best.waves_per_wg), [[table[4 * i + j] for j in range(4)]
                                                      for i in range(4)]


class BatchDecompressor:
    """Plans the decompression of `streams` (uint8 CUDA tensors holding
    whole streams) into freshly allocated planes."""

    def __init__(self, streams, tuning=None, stream=None, headers=None):
        dev = streams[0].device
        self.headers = headers or [api.read_header(s, stream) for s in streams]
        self.outputs = []
        items = (N.BatchDItem * len(streams))()
        for i, (s, h) in enumerate(zip(streams, self.headers)):
            dt = torch.float64 if h.float_bits == 64 else torch.float32
            planes = [torch.empty(h.n, dtype=dt, device=dev) for _ in range(h.dim)]
            self.outputs.append(tuple(planes) + ((None,) if h.dim == 2 else ()))
            it = items[i]
            it.h = h
            it.stream, it.bytes = s.data_ptr(), s.numel()
            it.x, it.y = planes[0].data_ptr(), planes[1].data_ptr()
            it.z = planes[2].data_ptr() if h.dim == 3 else None
        self.items = items
        self.streams = streams
        wsb = int(N.lib().gpz_batch_decompress_workspace_bytes(items, len(streams)))
        self.ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        self.plan = N.BatchPlan()
        self.replan(tuning, stream)

    def __len__(self):
        return len(self.headers)

    def replan(self, tuning=None, stream=None):
        t = _tuning(tuning)
        N.check(N.lib().gpz_batch_plan_decompress(
            self.items, len(self), ctypes.byref(t) if t else None, api._vp(self.ws),
            self.ws.numel(), ctypes.byref(self.plan), api._stream(stream)),
            "gpz_batch_plan_decompress")

    def run(self, stream=None):
        N.check(N.lib().gpz_batch_run(ctypes.byref(self.plan), api._vp(self.ws),
                                      api._stream(stream)), "gpz_batch_run")

    def graph(self):
        return BatchGraph(self.plan, self.ws)

    def autotune(self, reps=3, stream=None):
        best = N.BatchTuning()
        table = (ctypes.c_float * 16)()
        N.check(N.lib().gpz_batch_autotune_decompress(
            self.items, len(self), api._vp(self.ws), self.ws.numel(),
            api._stream(stream), int(reps), ctypes.byref(best), table),
            "gpz_batch_autotune_decompress")
        self.replan(best, stream)
        return (best.seg_groups, best.waves_per_wg), [[table[4 * i + j] for j in range(4)]
                                                      for i in range(4)]
