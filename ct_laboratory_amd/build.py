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
from concurrent.futures import ThreadPoolExecutor

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
    """Compiles every source to an object in parallel (one hipcc per source),
    then links the shared library."""
    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    out = lib_path()
    tmp = out + ".tmp"
    objs = [os.path.join(obj_dir, os.path.basename(src).replace(".hip", ".o")) for src in sources()]
    cmds = [[hipcc()] + HIPCC_FLAGS + ["-c", "-o", obj, src] for src, obj in zip(sources(), objs)]
    if verbose:
        for c in cmds:
            print(" ".join(c), flush=True)
    jobs = max(1, min(len(cmds), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        for r in pool.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds):
            if r.returncode != 0:
                sys.stderr.write(r.stderr)
                raise subprocess.CalledProcessError(r.returncode, r.args)
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
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
