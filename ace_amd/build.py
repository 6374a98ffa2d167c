"""Build recipe for libace_hip.so (gfx950) and the oracle's liboracle.so.

The library is built by ace_amd/Makefile: csrc/api.hip plus one object per basis-kernel instantiation set
(csrc/basis_inst.hip), compiled in parallel into ace_amd/_build/.  build_library(force=True) rebuilds every
object; force=False rebuilds what changed.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def _jobs():
    n = os.cpu_count() or 4
    return str(max(1, min(n, 16)))


def build_library(force=True):
    cmd = ["make", "-s", "-j", _jobs(), f"HIPCC={HIPCC}"]
    if force:
        cmd.insert(1, "-B")
    _run(cmd, cwd=HERE)
    return os.path.join(HERE, "libace_hip.so")


def build_oracle():
    """The oracle is test infrastructure; it is built here so the tests and the CPU baseline find it."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_adapter_test(force=True):
    """The integration guide's C++ adapters compiled into a test program (tests/adapter), linked with the
    library; the GPU tests run it."""
    cmd = ["make", "-s", "-C", os.path.join(ROOT, "tests", "adapter")]
    if force:
        cmd.insert(1, "-B")
    _run(cmd)
    return os.path.join(ROOT, "tests", "adapter", "adapter_test")


if __name__ == "__main__":
    import sys
    print(build_library(force="-B" in sys.argv))
    print(build_oracle())
    print(build_adapter_test())
