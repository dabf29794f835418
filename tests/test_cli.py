"""The `gpz` CLI (gpz_amd/lib/gpz): flag validation on the CPU, file round
trips on the GPU (stream file byte-identical to the oracle's, reconstructed
AoS file equal to the oracle's decode, -v reporting zero errors)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPZ = os.path.join(ROOT, "gpz_amd", "lib", "gpz")


def run(*args, check=False):
    p = subprocess.run([GPZ, *map(str, args)], capture_output=True, text=True,
                       timeout=300)
    if check and p.returncode != 0:
        raise AssertionError(p.stdout + p.stderr)
    return p


@pytest.mark.parametrize("args,msg", [
    (["-x", "-d"], "Cannot specify both"),
    ([], "Must specify either"),
    (["-x", "-i", "a", "-o", "b"], "Missing required arguments"),
    (["-x", "-i", "a", "-o", "b", "-n", 10, "-m", "rel", "-e", 1e-4, "-f", 16],
     "Float bits must be 32 or 64"),
    (["-x", "-i", "a", "-o", "b", "-n", 10, "-m", "rel", "-e", 1e-4, "-f", 32,
      "-p", 4], "Dimension must be 2 or 3"),
    (["-x", "-i", "a", "-o", "b", "-n", 10, "-m", "abs", "-e", 1e-4, "-f", 32],
     "explicit segment size"),
    (["-x", "-i", "a", "-o", "b", "-n", 10, "-m", "rel", "-e", 1e-4, "-f", 32,
      "-t", "lz4"], "nvcomp"),
    (["-d", "-o", "x"], "required for decompression"),
])
def test_flag_validation(args, msg):
    if not os.path.exists(GPZ):
        pytest.skip("gpz CLI not built")
    p = run(*args)
    assert p.returncode == 1
    assert msg in p.stderr


@pytest.mark.gpu
# Every case keeps block and rel spaces inside the format's 31-bit masks,
# so -v must report zero errors (abs 1e-3 with seg 64 in a 1000-wide box
# would not: uniform groups need 7812^3 block ids).
@pytest.mark.parametrize("bits,dim,mode,e,seg", [
    (32, 3, "rel", 1e-4, 0), (32, 2, "rel", 1e-5, 0), (64, 3, "abs", 0.01, 128),
    (64, 2, "rel", 1e-4, 0)])
def test_round_trip_files(tmp_path, oracle, bits, dim, mode, e, seg):
    from gpz_amd import datagen
    n = 200_000 + 77
    x, y, z = datagen.clustered(n, dim=dim, bits=bits, clusters=64, seed=31)
    planes = [x, y] + ([z] if dim == 3 else [])
    aos = np.stack(planes, axis=1)
    src = tmp_path / "in.bin"
    aos.tofile(src)
    comp = tmp_path / "out.gpz"
    rec = tmp_path / "rec.bin"
    args = ["-x", "-i", src, "-o", comp, "-n", n, "-p", dim, "-f", bits, "-m",
            mode, "-e", e]
    if seg:
        args += ["-s", seg]
    p = run(*args, check=True)
    assert "PCompression ratio" in p.stdout
    s_o, eb, seg_o, _ = oracle.compress_auto(x, y, z if dim == 3 else None,
                                             mode, e, seg)
    got = np.fromfile(comp, np.uint8)
    assert got.size == s_o.size and np.array_equal(got, s_o)
    assert f"Segment size: {seg_o}" in p.stdout
    p = run("-d", "-o", comp, "-r", rec, "-v", src, check=True)
    assert f"Decompressed num points: {n}" in p.stdout
    d = oracle.decompress(s_o)
    exp = np.stack([d[0], d[1]] + ([d[2]] if dim == 3 else []), axis=1)
    np.testing.assert_array_equal(np.fromfile(rec, x.dtype).reshape(n, dim), exp)
    assert re.search(r"^Error Count: 0$", p.stdout, re.M), p.stdout
