"""Build libmfcfd_hip.so (gfx950) in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 only, with packed FP32
instructions disabled; the FP64 parity kernels additionally use
-ffp-contract=off so no multiply-add is fused that the reference does not
fuse.  The sync-free sweep's unit is the exception: its sweep form of the
split fluxes already rounds differently from the reference, and its
flux evaluations sit on the sweep's dependency chain, so it fuses the
multiply-adds within an expression (-ffp-contract=on), the same way at every
call site.  Objects go to 3d-mfcfd_amd/_build/, the library to
3d-mfcfd_amd/libmfcfd_hip.so.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmfcfd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-function", "-I", CSRC]
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# translation unit -> extra flags
UNITS = {
    "capi.hip": ["-ffp-contract=off"],
    "kernels_fp64.hip": ["-ffp-contract=off"],
    "kernels_flux_fused.hip": ["-ffp-contract=off"],
    "kernels_sweep.hip": ["-ffp-contract=on"],
    "preprocess.cpp": ["-ffp-contract=off", "-x", "hip"],
}


def _compile(src: str, extra: list, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *NO_PACKED_FP32, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1], verbose), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
