"""Build libg6.so and libg4x.so (HIP, gfx950) in-tree with hipcc.

    python -m starlab_amd.build [--asm DIR]

libg6.so  exports the GRAPE-6 ABI (include/grape6.h), its extensions
          (include/g6x.h) and the ensemble engine (include/g6ens.h);
libg4x.so exports the GRAPE-4 ABI (include/grape4.h) and include/g4x.h.
abi_bench is a plain C program (csrc/abi_bench.c) linked against libg6.so
          that times kira's per-block call sequence.
libkira_blockstep.so (csrc/kira_blockstep.cc, include/kira_blockstep.h) is
          kira's block-step integration as a native caller of any GRAPE-6 ABI
          library, opened by path.
Both link the same force kernels and engine core; each has its own engine
state.  Symbols other than the declared C entry points are hidden.

The code is compiled for gfx950 only, with packed FP32 instructions disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); the host pass warns
that it does not know the feature, which is harmless.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libg6.so")
LIBG4 = os.path.join(HERE, "libg4x.so")
ABI_BENCH = os.path.join(HERE, "abi_bench")
LIBKB = os.path.join(HERE, "libkira_blockstep.so")

LIBS = {
    LIB: ["g6_kernels.hip", "g6_core.hip", "g6_engine.hip", "g6_ens.hip"],
    LIBG4: ["g6_kernels.hip", "g6_core.hip", "g4_engine.hip"],
}
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libg6")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}")


def build(verbose: bool = False) -> str:
    """Compile every source and link starlab_amd/libg6.so and libg4x.so
    (always recompiles)."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted({s for v in LIBS.values() for s in v})
    inc = f"-I{os.path.join(ROOT, 'include')}"

    def obj(s):
        return os.path.join(OBJ, s.replace(".hip", ".o"))

    def compile_one(s):
        _run([hipcc(), *HIP_FLAGS, "-c", "-o", obj(s), os.path.join(CSRC, s), inc], verbose)

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        list(ex.map(compile_one, srcs))
    for lib, parts in LIBS.items():
        tmp = lib + ".tmp"
        _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
              *[obj(s) for s in parts]], verbose)
        os.replace(tmp, lib)
    # the per-block ABI timing program: a plain C caller of libg6.so
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        raise RuntimeError("no C compiler (gcc) for abi_bench")
    tmp = ABI_BENCH + ".tmp"
    _run([cc, "-O2", "-std=c99", inc, os.path.join(CSRC, "abi_bench.c"), f"-L{HERE}",
          "-Wl,-rpath,$ORIGIN", "-lg6", "-lm", "-o", tmp], verbose)
    os.replace(tmp, ABI_BENCH)
    # kira's block-step caller in native code; it opens a GRAPE-6 library by path
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("no C++ compiler (g++) for libkira_blockstep.so")
    tmp = LIBKB + ".tmp"
    _run([cxx, "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", inc,
          os.path.join(CSRC, "kira_blockstep.cc"), "-ldl", "-o", tmp], verbose)
    os.replace(tmp, LIBKB)
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
