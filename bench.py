#!/usr/bin/env python3
"""Benchmark of the GPZ particle compressor on one MI355X.

A step is one compress plus one decompress of the resident snapshot
(143,735,721 clustered FP32 3D points, rel 1e-4 by default).  The timed
region holds K steps bracketed by torch.cuda.synchronize(); inputs, stream,
workspace and outputs stay in HBM and nothing is allocated inside it.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload NAME]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("compress/decompress GB/s & % HBM roofline, 143.7M f32 3D pts; "
          "batched vs per-call")
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def kernel_bytes(name, in_bytes, stream_bytes, payload_bytes, G, M, P):
    """Algorithmic HBM bytes of one launch of each kernel."""
    return {
        "encode": in_bytes + payload_bytes + (M + 8) * G,
        "decode": stream_bytes + in_bytes,
        "compact": 2 * (payload_bytes + M * G) + 16 * G,
        "scan": 16 * G + P,
        "seg_sample": 1024 * 256 * (in_bytes // max(1, G * 1024)),
        "seg_decide": 6 * 1024 * 8,
    }.get(name, 0)


def traffic_from_profiles(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(f"{kernel}/{workload}")
    except (OSError, ValueError):
        return None


def cpu_baseline(x, y, z, spec, eb, seg, threads, sample_points, reps=2):
    """The oracle (scalar/OpenMP C++ restatement) on the host cores, over
    the first `sample_points` points of the snapshot (whole 1024-point
    groups, compressed with the snapshot's own eb and segment size, so they
    are exactly the snapshot's first groups)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    used = O.num_threads(threads)
    dt = x.dtype.type
    m = min(x.size, max(1024, sample_points // 1024 * 1024))
    xs, ys = x[:m], y[:m]
    zs = z[:m] if z is not None else None
    nbytes = xs.nbytes * (3 if zs is not None else 2)

    def once():
        t0 = time.perf_counter()
        s = O.compress(xs, ys, zs, dt(eb), seg)
        t1 = time.perf_counter()
        O.decompress(s)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    once()
    runs = sorted((once() for _ in range(reps)), key=lambda r: r[0] + r[1])
    tc, td = runs[len(runs) // 2]
    return dict(value=round(nbytes / (tc + td) / 1e9, 4), unit="GB/s",
                cores=used, kind="port",
                sample=(f"first {m} of the {x.size} points of {spec['name']} "
                        f"({m // 1024} groups), oracle compress+decompress, "
                        f"OpenMP {used} threads on {cpu_model()}, median of "
                        f"{reps} after 1 warm-up ({tc + td:.2f} s per run)"),
                compress_GBps=round(nbytes / tc / 1e9, 4),
                decompress_GBps=round(nbytes / td / 1e9, 4))


def batch_leg(steps, warmup):
    """Uneven mix: 256 snapshots (N log-uniform in [1e4, 4e6], 80 % 3D, 75 %
    FP32, rel 10^U(-5,-3), seed 3).  Per-call = one gpz_compress (or
    gpz_decompress) per snapshot, back to back on one stream with a single
    synchronisation at the end (the most favourable per-call schedule);
    batched = one hipGraph launch of the autotuned batch."""
    import ctypes
    import numpy as np
    import torch
    from gpz_amd import _native as N
    from gpz_amd import api, batch, datagen
    L = N.lib()
    specs = datagen.batch_mix(256, seed=3)
    snaps = []
    in_bytes = 0
    for sp in specs:
        x, y, z = datagen.clustered(sp["n"], dim=sp["dim"], bits=sp["bits"],
                                    clusters=sp["clusters"], seed=sp["seed"])
        dev = [torch.from_numpy(a).cuda() for a in (x, y) + ((z,) if sp["dim"] == 3 else ())]
        snaps.append((dev[0], dev[1], dev[2] if sp["dim"] == 3 else None, "rel", sp["e"], 0))
        in_bytes += x.nbytes * sp["dim"]
    torch.cuda.synchronize()
    b = batch.BatchCompressor(snaps)
    ctune, _ = b.autotune(reps=2)
    b.run()
    streams = b.streams()
    out_bytes = sum(int(s.numel()) for s in streams)
    d = batch.BatchDecompressor(streams)
    dtune, _ = d.autotune(reps=2)

    # per-call resources, allocated outside the timed region
    s = api._stream()
    calls = []
    for p, (x, y, z) in zip(b.params, b.planes):
        out = torch.empty(api.stream_bound(p), dtype=torch.uint8, device="cuda")
        ws = torch.empty(api.workspace_bytes(p), dtype=torch.uint8, device="cuda")
        calls.append((p, x, y, z, out, ws))
    douts = []
    for st, h in zip(streams, d.headers):
        dt = torch.float64 if h.float_bits == 64 else torch.float32
        douts.append([torch.empty(h.n, dtype=dt, device="cuda") for _ in range(h.dim)])

    def per_call_compress():
        for p, x, y, z, out, ws in calls:
            L.gpz_compress(ctypes.byref(p), api._vp(x), api._vp(y), api._vp(z), api._vp(out),
                           out.numel(), api._vp(ws), ws.numel(), s)

    def per_call_decompress():
        for st, h, o in zip(streams, d.headers, douts):
            L.gpz_decompress(ctypes.byref(h), api._vp(st), st.numel(), api._vp(o[0]),
                             api._vp(o[1]), api._vp(o[2]) if len(o) == 3 else None, s)

    gc, gd = b.graph(), d.graph()

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_pc = timed(per_call_compress)
    t_pd = timed(per_call_decompress)
    t_bc = timed(gc.launch)
    t_bd = timed(gd.launch)
    # the batched and per-call streams agree
    for i in (0, 100, 255):
        r = (N.Result)()
        L.gpz_compress_result(api._vp(calls[i][5]), ctypes.byref(r), s)
        assert torch.equal(calls[i][4][: r.stream_bytes], streams[i]), i
    alg = in_bytes + out_bytes
    return dict(snapshots=len(specs), points=int(sum(sp["n"] for sp in specs)),
                in_bytes=in_bytes, stream_bytes=out_bytes,
                compress_tuning=list(ctune), decompress_tuning=list(dtune),
                per_call_compress_GBps=round(in_bytes / t_pc / 1e9, 2),
                per_call_decompress_GBps=round(in_bytes / t_pd / 1e9, 2),
                batched_compress_GBps=round(in_bytes / t_bc / 1e9, 2),
                batched_decompress_GBps=round(in_bytes / t_bd / 1e9, 2),
                compress_speedup=round(t_pc / t_bc, 2),
                decompress_speedup=round(t_pd / t_bd, 2),
                batched_compress_roofline_frac=round(alg / t_bc / 8e12, 4),
                batched_decompress_roofline_frac=round(alg / t_bd / 8e12, 4),
                ms=dict(per_call_compress=round(t_pc * 1e3, 3),
                        per_call_decompress=round(t_pd * 1e3, 3),
                        batched_compress=round(t_bc * 1e3, 3),
                        batched_decompress=round(t_bd * 1e3, 3)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="large_f32_3d")
    ap.add_argument("--n", type=int, default=None, help="override point count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-snapshot batched-vs-per-call leg")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-sample", type=int, default=1 << 62,
                    help="points of the snapshot the CPU baseline runs on")
    ap.add_argument("--profile-pass", action="store_true",
                    help="counter-collection run: a 1 GiB device copy (the "
                         "calibration kernel) then two compress and two "
                         "decompress calls; no timing output")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py runs on one GPU only: there is no multi-GPU path "
                 "(the compressor would shard by contiguous ranges of "
                 "1024-point groups, exchanging one 8-byte total per GPU)")

    import numpy as np
    import torch
    from gpz_amd import _native as N
    from gpz_amd import api, datagen

    dev = torch.device("cuda:0")
    t0 = time.time()
    x, y, z, spec = datagen.workload(args.workload, n=args.n)
    spec["name"] = args.workload
    log(f"[bench] generated {args.workload}: {x.size} points in "
        f"{time.time() - t0:.1f}s")
    dim, bits, n = spec["dim"], spec["bits"], x.size
    in_bytes = x.nbytes * dim
    planes = [torch.from_numpy(a).to(dev) for a in (x, y, z) if a is not None]
    dx, dy = planes[0], planes[1]
    dz = planes[2] if dim == 3 else None
    torch.cuda.synchronize()

    eb = api.resolve_eb(dx, dy, dz, spec["mode"], spec["e"])
    rel = float(np.float32(spec["e"])) if bits == 32 else float(spec["e"])
    p = api.make_params(dim, bits, n, eb, rel, spec["mode"], spec["seg"])
    out = torch.empty(api.stream_bound(p), dtype=torch.uint8, device=dev)
    ws = torch.empty(api.workspace_bytes(p), dtype=torch.uint8, device=dev)
    dt = torch.float64 if bits == 64 else torch.float32
    ox = torch.empty(n, dtype=dt, device=dev)
    oy = torch.empty(n, dtype=dt, device=dev)
    oz = torch.empty(n, dtype=dt, device=dev) if dim == 3 else None

    api.compress_into(p, dx, dy, dz, out, ws)
    res = api.compress_result(ws)
    N.check(res.status, "compress")
    nbytes = int(res.stream_bytes)
    hdr = api.read_header(out)
    log(f"[bench] eb={eb:.9g} seg={res.seg_size} stream={nbytes} B "
        f"ratio={in_bytes / nbytes:.3f}")

    if args.profile_pass:
        a = torch.empty(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        for _ in range(2):
            api.compress_into(p, dx, dy, dz, out, ws)
            torch.cuda.synchronize()
        for _ in range(2):
            api.decompress_into(hdr, out, nbytes, ox, oy, oz)
            torch.cuda.synchronize()
        log(f"[bench] profile pass done: in_bytes={in_bytes} stream={nbytes}")
        print(json.dumps({"profile_pass": args.workload, "in_bytes": in_bytes,
                          "stream_bytes": nbytes, "copy_bytes": 1 << 30}))
        return

    def step(ev=None):
        if ev:
            ev[0].record()
        api.compress_into(p, dx, dy, dz, out, ws)
        if ev:
            ev[1].record()
        api.decompress_into(hdr, out, nbytes, ox, oy, oz)
        if ev:
            ev[2].record()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"[bench] warmup {args.warmup} steps done")

    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)]
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tc = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps
    td = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps
    ms_step = elapsed * 1e3 / args.steps
    log(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms/step "
        f"(compress {tc:.3f} ms, decompress {td:.3f} ms)")

    # Per-kernel durations with hipEvents on the launch stream.
    L = N.lib()
    L.gpz_timing_reset()
    L.gpz_timing_enable(1)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    L.gpz_timing_enable(0)
    G = (n + 1023) // 1024
    P, M = (32, 64) if bits == 64 else (24, 36)
    payload = nbytes - P - (8 + M) * G
    kernels = {}
    import ctypes
    for name in ("seg_sample", "seg_decide", "encode", "scan", "compact",
                 "decode"):
        ms = ctypes.c_double(0)
        cnt = ctypes.c_uint64(0)
        L.gpz_timing_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value:
            avg = ms.value / cnt.value
            b = kernel_bytes(name, in_bytes, nbytes, payload, G, M, P)
            kernels[name] = dict(avg_ms=round(avg, 4),
                                 bytes=int(b),
                                 GBps=round(b / (avg * 1e-3) / 1e9, 1))
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    kd = kernels[dom]
    roof = dict(bound="hbm", achieved=kd["GBps"], peak=HBM_PEAK_GBPS,
                unit="GB/s", frac=round(kd["GBps"] / HBM_PEAK_GBPS, 4),
                traffic=traffic_from_profiles(dom, args.workload),
                kernel=dom, avg_ms=kd["avg_ms"], algorithmic_bytes=kd["bytes"])
    log(f"[bench] kernels: {json.dumps(kernels)}")

    # correctness spot check of the timed output (decode of the last step)
    assert torch.isfinite(ox).all().item()

    bleg = None
    if not args.no_batch:
        log("[bench] batch leg ...")
        bleg = batch_leg(args.steps, args.warmup)
        log(f"[bench] batch {json.dumps(bleg)}")

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(x, y, z, spec, eb, res.seg_size, args.cpu_threads,
                           args.cpu_sample)
        log(f"[bench] cpu baseline {cpu}")

    value = in_bytes / (ms_step * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if bits == 64 else "f32",
        "data": ("synthetic: deterministic clustered generator "
                 f"({spec['clusters']} Gaussian clusters, {int(spec['frac'] * 100)}% "
                 f"clustered, box {spec['box']:g}, seed {spec['seed']})"),
        "config": {"workload": args.workload, "n_points": n, "dim": dim,
                   "float_bits": bits, "mode": spec["mode"], "e": spec["e"],
                   "eb_abs": eb, "seg_size": int(res.seg_size),
                   "step": "compress + decompress",
                   "value_definition": "uncompressed bytes / step time"},
        "compress_GBps": round(in_bytes / (tc * 1e-3) / 1e9, 3),
        "decompress_GBps": round(in_bytes / (td * 1e-3) / 1e9, 3),
        "compress_ms": round(tc, 4),
        "decompress_ms": round(td, 4),
        "ratio": round(in_bytes / nbytes, 4),
        "stream_bytes": nbytes,
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "batch": bleg,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
