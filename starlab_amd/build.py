"""Build libg6.so (HIP, gfx950) in-tree with hipcc.

    python -m starlab_amd.build [--asm DIR]

The library is compiled for gfx950 only.  All HIP code is built with packed
FP32 instructions disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``);
the host pass warns that it does not know the feature, which is harmless.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libg6.so")

SOURCES = ["g6_kernels.hip", "g6_engine.hip", "g6_ens.hip"]
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libg6")


def build(verbose: bool = False) -> str:
    """Compile every source into starlab_amd/libg6.so (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-shared", "-o", tmp,
           *[os.path.join(CSRC, s) for s in SOURCES],
           f"-I{os.path.join(ROOT, 'include')}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libg6.so")
    os.replace(tmp, LIB)
    return LIB


def dump_asm(outdir: str) -> None:
    """Write the gfx950 assembly of the kernels (for inspection only)."""
    os.makedirs(outdir, exist_ok=True)
    src = os.path.join(CSRC, "g6_kernels.hip")
    out = os.path.join(outdir, "g6_kernels.s")
    subprocess.run([hipcc(), *HIP_FLAGS, "--cuda-device-only", "-S", "-o", out, src,
                    f"-I{os.path.join(ROOT, 'include')}"], check=True)
    print(out)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--asm", default=None, help="also write kernel assembly into DIR")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(verbose=a.v))
    if a.asm:
        dump_asm(a.asm)
