"""Build recipes for the native parts of gpz_amd (hipcc / g++, in-tree).

    python -m gpz_amd.build            # everything
    python -m gpz_amd.build --only hip

Outputs land in gpz_amd/lib/.  Device code is compiled for gfx950 only.
"""
import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off : no implicit FMA; every fma in the kernels is explicit,
#                     as the reference's __fmaf_rn / __fma_rn are.
# correctly rounded fp32 divide/sqrt and no denormal flushing keep float
# arithmetic IEEE, identical to the host oracle.
# -packed-fp32-ops  : keep the compiler from emitting packed fp32 VALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]

TARGETS = {
    "hip": ("libgpz_amd.so", ["gpz_kernels.hip"]),
    "datagen": ("libgpz_datagen.so", ["datagen.cpp"]),
}


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    out = os.path.join(LIB, "libgpz_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp,
          os.path.join(CSRC, "gpz_kernels.hip")])
    os.replace(tmp, out)
    return out


def build_datagen():
    out = os.path.join(LIB, "libgpz_datagen.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-o",
          tmp, os.path.join(CSRC, "datagen.cpp")])
    os.replace(tmp, out)
    return out


def build_cli():
    """The `gpz` drop-in CLI and the C++ ICudaCompressor classes."""
    out = os.path.join(LIB, "gpz")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "gpz_cxx.cpp"),
          os.path.join(CSRC, "gpz_cli.hip"), f"-L{LIB}", "-lgpz_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_oracle():
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"),
                    "liboracle.so"], check=True)


def build_all(only=None):
    os.makedirs(LIB, exist_ok=True)
    steps = {"hip": build_hip, "datagen": build_datagen, "cli": build_cli,
             "oracle": build_oracle}
    for name, fn in steps.items():
        if only and name not in only:
            continue
        fn()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=None)
    args = ap.parse_args()
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        sys.exit(f"hipcc not found at {HIPCC}")
    build_all(args.only)
