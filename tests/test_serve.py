"""Serving loop (gpz_amd/serve.py): continuous batching over the batched
kernels, admission against a device-memory budget, and every sampled
result byte-identical to the per-call path run on the same request alone."""
import pytest


def test_arena_first_fit_and_merge():
    import torch
    from gpz_amd.serve import Arena
    a = Arena.__new__(Arena)            # allocator bookkeeping only, no device
    a.size, a.free, a.used = 4096, [(0, 4096)], 0
    x = a.alloc(100)
    y = a.alloc(300)
    z = a.alloc(256)
    assert x == (0, 256) and y == (256, 512) and z == (768, 256)
    assert a.alloc(4096) is None
    a.release(*y)
    assert a.alloc(512) == (256, 512)
    a.release(256, 512)
    a.release(*x)
    a.release(*z)
    assert a.free == [(0, 4096)] and a.used == 0


@pytest.mark.gpu
def test_serving_mix_matches_alone():
    from gpz_amd.serve import Server
    srv = Server(seed=4, bases=2, codes=6, max_points=3_000_000, budget_gb=4.0)
    reqs = srv.requests(120, rate_gbps=30.0)
    stats, done = srv.run(reqs, check_every=4)
    assert stats["requests"] == len(done) == 120
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    assert stats["max_device_bytes"] <= stats["budget_bytes"]
    assert stats["mean_batch"] >= 1
    assert srv.verify(done) == 30


@pytest.mark.gpu
def test_serving_admission_waits_for_memory():
    """A budget that holds one large request at a time still serves every
    request (they queue until memory frees up)."""
    from gpz_amd.serve import Server
    srv = Server(seed=5, bases=1, codes=2, max_points=2_000_000, budget_gb=0.12)
    reqs = srv.requests(24, rate_gbps=1000.0)
    stats, done = srv.run(reqs, check_every=6)
    assert len(done) == 24
    assert stats["max_device_bytes"] <= stats["budget_bytes"]
    assert srv.verify(done) == 4
