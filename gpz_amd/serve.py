"""Serving loop: continuous batching of compress and decompress requests
over the batched kernels (include/gpz.h, gpz_batch_*), with inputs, codes
and outputs resident in HBM and admission against a device-memory budget.

The reference has no server; this is the request loop the batched path is
for.  Requests arrive as a seeded Poisson process; a request is admitted
when device memory for its outputs and workspace fits the budget, admitted
requests are packed into one batch per direction, and up to `max_inflight`
batches run at once on their own streams, so requests that arrive while
others are on the GPU join the next batch while earlier work is still
running.  A request's latency runs from its arrival to the host seeing its
batch complete.  Every result is byte-identical to what the per-call path
produces for the same request alone (checked on a sample).

    python -m gpz_amd.serve --requests 2000 --rate-gbps 40
"""
import argparse
import bisect
import ctypes
import json
import math
import time

import numpy as np
import torch

from . import _native as N
from . import api, datagen

ALIGN = 256


class Arena:
    """First-fit allocator over one device buffer; `budget` bytes at most
    are handed out (the admission limit)."""

    def __init__(self, nbytes, device="cuda"):
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.size = nbytes
        self.free = [(0, nbytes)]          # sorted (offset, length)
        self.used = 0

    def alloc(self, n):
        n = (n + ALIGN - 1) // ALIGN * ALIGN
        for i, (o, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (o + n, ln - n)
                self.used += n
                return o, n
        return None

    def release(self, o, n):
        self.used -= n
        i = bisect.bisect_left(self.free, (o, 0))
        self.free.insert(i, (o, n))
        # merge with neighbours
        if i + 1 < len(self.free) and o + n == self.free[i + 1][0]:
            self.free[i] = (o, n + self.free[i + 1][1])
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == self.free[i][0]:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            self.free.pop(i)

    def view(self, o, n):
        return self.buf[o: o + n]


def log_uniform(u, lo, hi):
    return int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))


class Server:
    """Resident pools: `bases` clustered FP32 3D snapshots of `max_points`
    points (a compress request of n points takes the first n of one base)
    and `codes` streams compressed at start-up (a decompress request takes
    one of them)."""

    def __init__(self, seed=4, bases=4, codes=32, min_points=10_000,
                 max_points=50_000_000, e=1e-4, budget_gb=200.0, max_inflight=2,
                 device="cuda"):
        self.seed, self.e = seed, e
        self.min_points, self.max_points = min_points, max_points
        self.max_inflight = max_inflight
        self.L = N.lib()
        self.bases = []
        for b in range(bases):
            x, y, z = datagen.clustered(max_points, dim=3, bits=32,
                                        clusters=max(1, max_points // 2048),
                                        seed=seed * 1_000_003 + b)
            self.bases.append(tuple(torch.from_numpy(a).to(device) for a in (x, y, z)))
        self.ebs = {}
        # code pool: streams of clustered snapshots, resident
        self.codes = []
        for j in range(codes):
            n = log_uniform(datagen.uniform01(seed, 90, j), min_points, max_points)
            x, y, z = self._input(j % bases, n)
            c = api.compress(x, y, z, mode="rel", e=e)
            self.codes.append((c.bytes().clone(), api.read_header(c.bytes()), n))
        torch.cuda.synchronize()
        self.budget = int(budget_gb * 1e9)
        self.reserved = 0
        # the budget bounds the reservations; the arena keeps some slack for
        # fragmentation
        self.arena = Arena(int(self.budget * 1.25) + (64 << 20), device)
        self.streams = [torch.cuda.Stream() for _ in range(max_inflight)]
        self.plan_stream = torch.cuda.Stream()

    def _input(self, b, n):
        x, y, z = self.bases[b]
        return x[:n], y[:n], z[:n]

    def eb(self, b, n):
        """Absolute bound of a compress request (the client's rel -> abs,
        as the reference CLI resolves it before compressing)."""
        k = (b, n)
        if k not in self.ebs:
            self.ebs[k] = api.resolve_eb(*self._input(b, n), "rel", self.e)
        return self.ebs[k]

    def requests(self, count, rate_gbps, compress_frac=0.6):
        """Seeded request list: arrival time (s), kind, payload."""
        rate_bytes = rate_gbps * 1e9
        out, t = [], 0.0
        for i in range(count):
            u = [datagen.uniform01(self.seed, 100 + k, i) for k in range(4)]
            if u[0] < compress_frac:
                n = log_uniform(u[1], self.min_points, self.max_points)
                req = dict(kind="c", base=int(u[2] * len(self.bases)) % len(self.bases), n=n,
                           bytes=12 * n)
            else:
                j = int(u[2] * len(self.codes)) % len(self.codes)
                req = dict(kind="d", code=j, n=self.codes[j][2], bytes=12 * self.codes[j][2])
            out.append(req)
        mean_bytes = sum(r["bytes"] for r in out) / len(out)
        lam = rate_bytes / mean_bytes                      # requests per second
        for i, r in enumerate(out):
            r["idx"] = i
            u = datagen.uniform01(self.seed, 200, i)
            t += -math.log(max(1e-300, 1.0 - u)) / lam
            r["arrival"] = t
        return out

    # ---- admission and one batch
    def _admit(self, req):
        """Counts a request's device memory (outputs plus its share of the
        batch workspace) against the budget; False when it does not fit."""
        if req["kind"] == "c":
            p = api.make_params(3, 32, req["n"], self.eb(req["base"], req["n"]),
                                float(np.float32(self.e)), "rel", 0)
            req["params"] = p
            req["out_bytes"] = [api.stream_bound(p)]
            one = (N.BatchItem * 1)()
            one[0].p = p
            need = (req["out_bytes"][0] + 4 * ALIGN
                    + int(self.L.gpz_batch_compress_workspace_bytes(one, 1)))
        else:
            req["out_bytes"] = [4 * req["n"]] * 3
            need = 12 * req["n"] + 4 * ALIGN
        if need > self.budget:
            raise ValueError(f"request needs {need} bytes, more than the budget")
        if self.reserved + need > self.budget:
            return False
        self.reserved += need
        req["need"] = need
        return True

    def _unreserve(self, req):
        self.reserved -= req["need"]

    def _release(self, req):
        for o, ln in req.get("mem", []):
            self.arena.release(o, ln)
        req["mem"] = []

    def _alloc_batch(self, batch):
        """Output buffers of every request; None (nothing held) when the
        arena is too fragmented right now."""
        got = []
        for r in batch:
            r["mem"] = []
            for nb in r["out_bytes"]:
                a = self.arena.alloc(nb)
                if a is None:
                    for rr in batch:
                        self._release(rr)
                    return None
                r["mem"].append(a)
            got.append(r)
        return got

    def _launch(self, batch, stream):
        """Plans and launches one batch (compress and decompress parts);
        returns (event, workspaces held, compress plan, compress requests),
        or None when the workspaces do not fit the arena yet."""
        comp = [r for r in batch if r["kind"] == "c"]
        deco = [r for r in batch if r["kind"] == "d"]
        held = []
        ps = api._stream(self.plan_stream)
        citems = ditems = None
        if comp:
            citems = (N.BatchItem * len(comp))()
            for it, r in zip(citems, comp):
                x, y, z = self._input(r["base"], r["n"])
                it.p = r["params"]
                it.x, it.y, it.z = x.data_ptr(), y.data_ptr(), z.data_ptr()
                so, sl = r["mem"][0]
                it.stream, it.cap = self.arena.view(so, sl).data_ptr(), sl
        if deco:
            ditems = (N.BatchDItem * len(deco))()
            for it, r in zip(ditems, deco):
                st, h, _ = self.codes[r["code"]]
                it.h = h
                it.stream, it.bytes = st.data_ptr(), st.numel()
                it.x, it.y, it.z = [self.arena.view(o, ln).data_ptr() for o, ln in r["mem"]]
        sizes = []
        if comp:
            sizes.append(int(self.L.gpz_batch_compress_workspace_bytes(citems, len(comp))))
        if deco:
            sizes.append(int(self.L.gpz_batch_decompress_workspace_bytes(ditems, len(deco))))
        for nb in sizes:
            a = self.arena.alloc(nb)
            if a is None:
                for h in held:
                    self.arena.release(*h)
                return None
            held.append(a)
        ws = [ctypes.c_void_p(self.arena.view(*h).data_ptr()) for h in held]
        comp_plan = None
        k = 0
        if comp:
            plan = N.BatchPlan()
            N.check(self.L.gpz_batch_plan_compress(citems, len(comp), None, ws[k], sizes[k],
                                                   ctypes.byref(plan), ps), "plan")
            N.check(self.L.gpz_batch_run(ctypes.byref(plan), ws[k], api._stream(stream)), "run")
            comp_plan = (plan, ws[k])
            k += 1
        if deco:
            plan = N.BatchPlan()
            N.check(self.L.gpz_batch_plan_decompress(ditems, len(deco), None, ws[k], sizes[k],
                                                     ctypes.byref(plan), ps), "plan")
            N.check(self.L.gpz_batch_run(ctypes.byref(plan), ws[k], api._stream(stream)), "run")
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev, held, comp_plan, comp

    def run(self, reqs, check_every=100):
        """Serves `reqs` in real time; returns (statistics, requests).
        Every `check_every`-th request keeps its memory for verify()."""
        pending, inflight, done = [], [], []
        i, n = 0, len(reqs)
        t0 = time.perf_counter()
        busy_streams = [None] * self.max_inflight
        max_used = 0
        while len(done) < n:
            now = time.perf_counter() - t0
            while i < n and reqs[i]["arrival"] <= now:
                pending.append(reqs[i])
                i += 1
            # completions
            for k, slot in enumerate(busy_streams):
                if slot is not None and slot[0].query():
                    fin = time.perf_counter() - t0
                    ev, held, comp_plan, comp, batch = slot
                    if comp_plan is not None:
                        res = (N.Result * len(comp))()
                        N.check(self.L.gpz_batch_compress_results(
                            ctypes.byref(comp_plan[0]), comp_plan[1], res,
                            api._stream(self.streams[k])), "results")
                        for r, rr in zip(comp, res):
                            N.check(rr.status, "request")
                            r["stream_bytes"] = int(rr.stream_bytes)
                    for r in batch:
                        r["finish"] = fin
                        r["latency"] = fin - r["arrival"]
                        r["batch_size"] = len(batch)
                        r["check"] = r["idx"] % check_every == 0
                        done.append(r)
                        self._unreserve(r)
                        if not r["check"]:
                            self._release(r)
                    for o, ln in held:
                        self.arena.release(o, ln)
                    busy_streams[k] = None
            # admission and launch
            free_slot = next((k for k, s in enumerate(busy_streams) if s is None), None)
            if pending and free_slot is not None:
                batch = []
                while pending and self._admit(pending[0]):
                    batch.append(pending.pop(0))
                launched = None
                if batch and self._alloc_batch(batch) is not None:
                    with torch.cuda.stream(self.streams[free_slot]):
                        launched = self._launch(batch, self.streams[free_slot])
                    if launched is None:
                        for r in batch:
                            self._release(r)
                if launched is not None:
                    max_used = max(max_used, self.reserved)
                    busy_streams[free_slot] = launched + (batch,)
                elif batch:                  # arena fragmented: retry later
                    for r in reversed(batch):
                        self._unreserve(r)
                        pending.insert(0, r)
            elif not pending and all(s is None for s in busy_streams) and i < n:
                # idle until the next arrival
                wait = reqs[i]["arrival"] - (time.perf_counter() - t0)
                if wait > 2e-4:
                    time.sleep(wait - 1e-4)
        wall = time.perf_counter() - t0
        lat = np.array([r["latency"] for r in done])
        total = sum(r["bytes"] for r in done)
        return dict(requests=n, wall_s=round(wall, 3),
                    sustained_GBps=round(total / wall / 1e9, 3),
                    offered_GBps=round(total / reqs[-1]["arrival"] / 1e9, 3),
                    p50_ms=round(float(np.percentile(lat, 50)) * 1e3, 3),
                    p99_ms=round(float(np.percentile(lat, 99)) * 1e3, 3),
                    mean_batch=round(float(np.mean([r["batch_size"] for r in done])), 2),
                    max_device_bytes=int(max_used), budget_bytes=self.budget), done

    def verify(self, done):
        """Byte-identity of sampled results with the per-call path run on
        the same request alone."""
        checked = 0
        for r in done:
            if not r.get("check"):
                continue
            if r["kind"] == "c":
                x, y, z = self._input(r["base"], r["n"])
                one = api.compress(x, y, z, mode="rel", e=self.e, eb=self.eb(r["base"], r["n"]))
                got = self.arena.view(r["mem"][0][0], r["stream_bytes"])
                ok = one.nbytes == r["stream_bytes"] and torch.equal(one.bytes(), got)
            else:
                st, h, _ = self.codes[r["code"]]
                want = api.decompress(st)
                ok = all(torch.equal(self.arena.view(o, 4 * r["n"]).view(torch.float32), w)
                         for (o, _), w in zip(r["mem"], want))
            if not ok:
                raise AssertionError(f"request differs from its per-call result: {r}")
            self._release(r)
            checked += 1
        return checked


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=20000)
    ap.add_argument("--rate-gbps", type=float, default=40.0)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-points", type=int, default=50_000_000)
    ap.add_argument("--bases", type=int, default=4)
    ap.add_argument("--codes", type=int, default=32)
    ap.add_argument("--check-every", type=int, default=100)
    a = ap.parse_args()
    srv = Server(seed=a.seed, bases=a.bases, codes=a.codes, max_points=a.max_points,
                 budget_gb=a.budget_gb)
    reqs = srv.requests(a.requests, a.rate_gbps)
    for r in reqs:                       # clients send absolute bounds
        if r["kind"] == "c":
            srv.eb(r["base"], r["n"])
    torch.cuda.synchronize()
    stats, done = srv.run(reqs, a.check_every)
    stats["checked_bit_exact"] = srv.verify(done)
    print(json.dumps(stats), flush=True)


if __name__ == "__main__":
    main()
