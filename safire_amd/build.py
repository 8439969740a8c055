"""Builds libsafire_afq.so (HIP for gfx950 + the C ABI) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsafire_afq.so")
SOURCES = ["afq_api.hip", "afq_phmsd.hip", "afq_rng.hip", "k_green.hip", "k_gemm.hip", "k_gemm_bf16.hip", "k_energy.hip", "k_exchange_bf16.hip",
           "k_exchange_i8.hip", "k_phmsd.hip", "k_propagate.hip", "k_qr.hip", "k_walker.hip",
           "k_setup.hip"]
# k_green.hip unrolls the 32 Gauss-Jordan steps of its N <= 32 kernel in full (each step indexes
# the register-resident rows with the step number); the default budget of a pragma-driven unroll
# would leave the loop rolled and the rows in scratch memory.
FILE_FLAGS = {"k_green.hip": ["-mllvm", "-pragma-unroll-threshold=1000000"]}
# Packed FP32 VALU ops are switched off for every HIP translation unit.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result"]


def build(verbose: bool = False, jobs: int = 8) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = ["hipcc", *HIPFLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o",
               obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + ".tmp"
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-shared", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    return LIB


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(out.decode(errors="replace"))
        raise RuntimeError(f"hipcc failed on {src}")
    text = out.decode(errors="replace")
    if text.strip():
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    print(build(verbose=True))
