"""Synthetic particle snapshots (deterministic; see csrc/datagen.cpp).

Workloads, named by what they are:
  small_f32_3d   1,048,576 3D FP32 points, 512 clusters in [0,1000)^3, seed 0
  large_f32_3d   143,735,721 3D FP32 points, 65,536 clusters in [0,5000)^3,
                 seed 1
  f64_3d         67,108,864 3D FP64 points, 32,768 clusters in [0,1e4)^3,
                 seed 2
  f32_2d         100,000,000 2D FP32 points uniform in [0,1)^2, 20% in 4,096
                 clusters of sigma 1e-3, seed 2
Clustered sets put 70% of the points in clusters with sigma ~
lognormal(ln 5, 0.5); the rest are uniform in the box.
"""
import math

import numpy as np

from . import _native as N

LN5 = math.log(5.0)

WORKLOADS = {
    "small_f32_3d": dict(n=1_048_576, dim=3, bits=32, clusters=512, box=1000.0,
                         frac=0.7, sigma=0.0, mu=LN5, s=0.5, seed=0,
                         mode="rel", e=1e-4, seg=0),
    "large_f32_3d": dict(n=143_735_721, dim=3, bits=32, clusters=65_536,
                         box=5000.0, frac=0.7, sigma=0.0, mu=LN5, s=0.5,
                         seed=1, mode="rel", e=1e-4, seg=0),
    "large_f32_3d_abs": dict(n=143_735_721, dim=3, bits=32, clusters=65_536,
                             box=5000.0, frac=0.7, sigma=0.0, mu=LN5, s=0.5,
                             seed=1, mode="abs", e=0.05, seg=64),
    "f64_3d": dict(n=67_108_864, dim=3, bits=64, clusters=32_768, box=1e4,
                   frac=0.7, sigma=0.0, mu=LN5, s=0.5, seed=2, mode="abs",
                   e=1e-3, seg=64),
    "f32_2d": dict(n=100_000_000, dim=2, bits=32, clusters=4096, box=1.0,
                   frac=0.2, sigma=1e-3, mu=0.0, s=0.0, seed=2, mode="rel",
                   e=1e-5, seg=0),
}


def clustered(n, dim=3, bits=32, clusters=512, box=1000.0, frac=0.7,
              sigma=0.0, mu=LN5, s=0.5, seed=0, **_):
    """Host SoA planes (numpy) of a clustered snapshot."""
    dt = np.float64 if bits == 64 else np.float32
    x = np.empty(n, dt)
    y = np.empty(n, dt)
    z = np.empty(n, dt) if dim == 3 else None
    rc = N.datagen().gpz_gen_clustered(
        bits, dim, n, clusters, box, frac, sigma, mu, s, seed,
        x.ctypes.data, y.ctypes.data, z.ctypes.data if z is not None else None)
    if rc != 0:
        raise ValueError("bad generator arguments")
    return x, y, z


def workload(name, n=None):
    """(x, y, z, spec) for a named workload; n overrides the point count."""
    spec = dict(WORKLOADS[name])
    if n is not None:
        spec["n"] = int(n)
    x, y, z = clustered(**spec)
    return x, y, z, spec


def uniform01(seed, stream, idx):
    return N.datagen().gpz_gen_uniform(seed, stream, idx)


def batch_mix(count=256, seed=3):
    """Snapshot specs of the uneven batch: N log-uniform in [1e4, 4e6],
    80% 3D, 75% FP32, rel bound 10^U(-5,-3)."""
    out = []
    for i in range(count):
        u = [uniform01(seed, 10 + k, i) for k in range(5)]
        n = int(round(math.exp(math.log(1e4) + u[0] * (math.log(4e6) - math.log(1e4)))))
        dim = 3 if u[1] < 0.8 else 2
        bits = 32 if u[2] < 0.75 else 64
        e = 10.0 ** (-5.0 + 2.0 * u[3])
        out.append(dict(n=n, dim=dim, bits=bits, clusters=max(1, n // 2048),
                        box=1000.0, frac=0.7, sigma=0.0, mu=LN5, s=0.5,
                        seed=seed * 1_000_003 + i, mode="rel", e=e, seg=0))
    return out
