"""Build recipe for libace_hip.so (gfx950) and the oracle's liboracle.so."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # no packed FP32 (v_pk_*_f32) code generation
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_library(out=None):
    out = out or os.path.join(HERE, "libace_hip.so")
    src = os.path.join(HERE, "csrc", "api.hip")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, out)
    return out


def build_oracle():
    """The oracle is test infrastructure; it is built here so the tests and the CPU baseline find it."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


if __name__ == "__main__":
    print(build_library())
    print(build_oracle())
