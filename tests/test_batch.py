"""Batched path (gpz_amd/batch.py over gpz_batch_*): every snapshot's
stream is byte-identical to the oracle's (and to the per-call path's), for
every tuning, on eager runs and on hipGraph replays; batched decodes are
bit-identical to the oracle's."""
import ctypes

import numpy as np
import pytest
import torch

from gpz_amd import _native as N

# (bits, dim, n, mode, e, seg)
MIX = [
    (32, 3, 70001, "rel", 1e-4, 0),
    (32, 2, 3 * 1024, "rel", 1e-5, 0),
    (64, 3, 5000, "abs", 0.05, 64),
    (64, 2, 1025, "rel", 1e-4, 0),
    (32, 3, 1, "rel", 1e-4, 0),
    (32, 3, 300, "abs", 0.01, 16),
    (64, 3, 20000, "rel", 1e-3, 0),
    (32, 2, 1023, "abs", 1e-3, 8),
    (32, 3, 250000, "rel", 1e-3, 0),
]


def _snapshots(mix, seed=7):
    from gpz_amd import datagen
    snaps, host = [], []
    for i, (bits, dim, n, mode, e, seg) in enumerate(mix):
        x, y, z = datagen.clustered(n, dim=dim, bits=bits, clusters=16, seed=seed + i)
        zz = z if dim == 3 else None
        host.append((x, y, zz, mode, e, seg))
        dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() if a is not None else None
               for a in (x, y, zz)]
        snaps.append((dev[0], dev[1], dev[2], mode, e, seg))
    return snaps, host


def test_batch_plan_rejects_bad_arguments_without_gpu_work():
    L = N.lib()
    plan = N.BatchPlan()
    assert L.gpz_batch_plan_compress(None, 0, None, None, 0, ctypes.byref(plan), None) \
        == N.GPZ_ERR_INVALID_ARG
    items = (N.BatchItem * 1)()
    items[0].p.dim = 4
    assert L.gpz_batch_plan_compress(items, 1, None, None, 0, ctypes.byref(plan), None) \
        == N.GPZ_ERR_INVALID_ARG
    bad = N.BatchTuning()
    bad.seg_groups, bad.waves_per_wg = 3, 1
    assert L.gpz_batch_plan_compress(items, 1, ctypes.byref(bad), None, 0,
                                     ctypes.byref(plan), None) == N.GPZ_ERR_INVALID_ARG
    assert L.gpz_batch_compress_workspace_bytes(None, 0) == 0
    assert L.gpz_batch_run(None, None, None) == N.GPZ_ERR_INVALID_ARG
    assert L.gpz_batch_graph_launch(None, None) == N.GPZ_ERR_INVALID_ARG


@pytest.mark.gpu
def test_batch_streams_match_oracle_and_per_call(oracle):
    from gpz_amd import api, batch
    snaps, host = _snapshots(MIX)
    b = batch.BatchCompressor(snaps)
    b.run()
    got = [s.cpu().numpy() for s in b.streams()]
    for i, ((x, y, z, mode, e, seg), g) in enumerate(zip(host, got)):
        s_o, eb, seg_o, adj = oracle.compress_auto(x, y, z, mode, e, seg)
        assert g.size == s_o.size and np.array_equal(g, s_o), f"snapshot {i} {MIX[i]}"
        r = b.results()[i]
        assert (r.seg_size, bool(r.adjusted)) == (seg_o, bool(adj))
        dx, dy, dz = snaps[i][:3]
        one = api.compress(dx, dy, dz, mode=mode, e=e, seg_size=seg).to_numpy()
        assert np.array_equal(one, g)


@pytest.mark.gpu
@pytest.mark.parametrize("seg_groups", [1, 2, 4, 8])
@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_batch_every_tuning_same_streams(oracle, seg_groups, waves):
    from gpz_amd import batch
    snaps, host = _snapshots(MIX[:5], seed=11)
    b = batch.BatchCompressor(snaps, tuning=(seg_groups, waves))
    b.run()
    for i, (g, (x, y, z, mode, e, seg)) in enumerate(zip(b.streams(), host)):
        s_o = oracle.compress_auto(x, y, z, mode, e, seg)[0]
        assert np.array_equal(g.cpu().numpy(), s_o), f"snapshot {i}"


@pytest.mark.gpu
def test_batch_graph_replays_match_eager():
    from gpz_amd import batch
    snaps, _ = _snapshots(MIX, seed=21)
    b = batch.BatchCompressor(snaps, tuning=(2, 4))
    b.run()
    eager = [s.clone() for s in b.streams()]
    g = b.graph()
    for rep in range(3):
        b.out.fill_(0xA5)                   # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, (s, e) in enumerate(zip(b.streams(), eager)):
            assert torch.equal(s, e), f"replay {rep} snapshot {i}"
        _ = sum(float(t[0].sum()) for t in snaps)  # more eager work


@pytest.mark.gpu
def test_batch_decompress_matches_oracle_and_graph(oracle):
    from gpz_amd import batch
    snaps, host = _snapshots(MIX, seed=31)
    b = batch.BatchCompressor(snaps)
    b.run()
    streams = b.streams()
    d = batch.BatchDecompressor(streams, tuning=(4, 2))
    d.run()
    torch.cuda.synchronize()
    want = [oracle.decompress(s.cpu().numpy()) for s in streams]
    for i, (out, w) in enumerate(zip(d.outputs, want)):
        for a in range(len(out)):
            if out[a] is not None:
                np.testing.assert_array_equal(out[a].cpu().numpy(), w[a],
                                              err_msg=f"snapshot {i} axis {a}")
    g = d.graph()
    for rep in range(3):
        for out in d.outputs:
            for t in out:
                if t is not None:
                    t.fill_(-1.0)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, (out, w) in enumerate(zip(d.outputs, want)):
            for a in range(len(out)):
                if out[a] is not None:
                    np.testing.assert_array_equal(out[a].cpu().numpy(), w[a],
                                                  err_msg=f"replay {rep} snapshot {i}")


@pytest.mark.gpu
def test_batch_autotune_picks_a_valid_tuning(oracle):
    from gpz_amd import batch
    snaps, host = _snapshots(MIX[:4], seed=41)
    b = batch.BatchCompressor(snaps)
    best, table = b.autotune(reps=1)
    assert best[0] in (1, 2, 4, 8) and best[1] in (1, 2, 4, 8)
    assert all(t > 0 for row in table for t in row)
    b.run()
    for g, (x, y, z, mode, e, seg) in zip(b.streams(), host):
        assert np.array_equal(g.cpu().numpy(), oracle.compress_auto(x, y, z, mode, e, seg)[0])
