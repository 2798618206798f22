"""Build the native library ``ksim_amd/libksim_hip.so`` for gfx950 in-tree.

    python -m ksim_amd.build            # or __graft_entry__.build()

hipcc compiles the kernels and the C ABI (``csrc/api.hip``) and the host
packer (``csrc/pack.cpp``) into one shared object.  Flags that matter:

* ``--offload-arch=gfx950`` — MI355X only, no other targets;
* ``-ffp-contract=off`` — the float64 magnitude and Bernoulli test must be the
  same IEEE operation sequence as the CPU oracle (no fused multiply-add);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` — no packed FP32 ops.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libksim_hip.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    return [os.path.join(CSRC, "api.hip"), os.path.join(CSRC, "pack.cpp"),
            os.path.join(CSRC, "wire.cpp"), os.path.join(CSRC, "noise.cpp")]


def build(verbose: bool = False, out: str = LIB, extra: list | None = None) -> str:
    """Compile every source; always recompiles (this is the build check)."""
    tmp = out + ".tmp"
    cmd = [HIPCC, *DEVICE, *COMMON, "-shared", *sources(), "-o", tmp, *(extra or [])]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode})")
    os.replace(tmp, out)
    return out


def asm(out_dir: str) -> None:
    """Dump the gfx950 assembly of the kernels (for inspection)."""
    os.makedirs(out_dir, exist_ok=True)
    cmd = [HIPCC, *DEVICE, *COMMON, "--cuda-device-only", "-S",
           os.path.join(CSRC, "api.hip"), "-o", os.path.join(out_dir, "api-gfx950.s")]
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    print(build(verbose=True))
