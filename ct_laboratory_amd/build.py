"""Build libctlab_hip.so for gfx950 with hipcc (in-tree, no torch extension).

    python -m ct_laboratory_amd.build          # compile
    python -m ct_laboratory_amd.build --isa    # also dump the gfx950 assembly

Flags that matter for correctness:
  -ffp-contract=off                  every float op rounded on its own (the
                                     reference's torch op order); fused
                                     multiply-adds only where the source
                                     writes __builtin_fmaf
  -fhip-fp32-correctly-rounded-divide-sqrt
                                     IEEE division and sqrt, so plane
                                     crossings equal the reference's bits
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                     no packed FP32 instructions
  -munsafe-fp-atomics                hardware float atomic add
"""
from __future__ import annotations

import argparse
import glob
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
OUT_DIR = os.path.join(PKG_DIR, "_lib")
LIB_NAME = "libctlab_hip.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-munsafe-fp-atomics",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found; the HIP extension cannot be built")
    return cand


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def lib_path() -> str:
    return os.path.join(OUT_DIR, LIB_NAME)


def build(verbose: bool = False, isa: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-shared", "-o", tmp] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    if isa:
        dump_isa(verbose=verbose)
    return out


def dump_isa(verbose: bool = False) -> list[str]:
    """Writes gfx950 assembly of each source next to the library (_lib/*.s)."""
    outs = []
    for src in sources():
        s_out = os.path.join(OUT_DIR, os.path.basename(src).replace(".hip", f"-{ARCH}.s"))
        cmd = [hipcc()] + HIPCC_FLAGS + ["--offload-device-only", "-S", "-o", s_out, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        outs.append(s_out)
    return outs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--isa", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(verbose=a.verbose, isa=a.isa))
    return 0


if __name__ == "__main__":
    sys.exit(main())
