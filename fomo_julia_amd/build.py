"""Builds libfomo_mi355x.so in-tree with hipcc for gfx950, and beside it the
bandwidth-probe library libfomo_probe.so that bench.py's --pattern-probe loads
(measurement only: its failure to build does not fail the product build).

    python -m fomo_julia_amd.build

Flags: -ffp-contract=off keeps the kernels' arithmetic exactly as written
(explicit fmaf only); packed-FP32 instructions are disabled; float division
stays IEEE (hipcc's default correctly rounded divide).  The pragma unroll
threshold is raised so that the fused kernel's z-march is fully unrolled at
every order and tile width (its ring slots must be compile-time indices; at
fd_order >= 10 with 2 or 4 columns per lane the default threshold left them
in scratch memory).  Kernels the default threshold already unrolled compile
to the same code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libfomo_mi355x.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = [os.path.join(HERE, "csrc", "fomo_mi355x.hip")]
PROBE_NAME = "libfomo_probe.so"
PROBE_PATH = os.path.join(HERE, PROBE_NAME)
PROBE_SOURCES = [os.path.join(HERE, "csrc", "pattern_probe.hip")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f) for f in ("kernels.h", "fomo_device.h", "setup.h")] + [
    os.path.join(os.path.dirname(HERE), "include", "fomo_mi355x.h")]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: str, sources: list[str] | None = None) -> list[str]:
    return [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-pragma-unroll-threshold=1000000",
            "-Wno-unused-command-line-argument", "-o", out] + (SOURCES if sources is None else sources)


def _compile(path: str, sources: list[str], verbose: bool) -> None:
    tmp = path + ".tmp"
    cmd = command(tmp, sources)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, path)


def build(verbose: bool = False) -> str:
    """Compiles the product library.  The bandwidth probe is measurement code:
    it is compiled afterwards, and a failure there is reported, not raised."""
    _compile(LIB_PATH, SOURCES, verbose)
    try:
        build_probe(verbose)
    except (subprocess.CalledProcessError, OSError) as e:
        print(f"warning: {PROBE_NAME} (bench.py --pattern-probe only) did not build: {e}",
              file=sys.stderr, flush=True)
    return LIB_PATH


def build_probe(verbose: bool = False) -> str:
    _compile(PROBE_PATH, PROBE_SOURCES, verbose)
    return PROBE_PATH


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB_PATH)
