"""Builds libsgmc.so (all HIP sources under csrc/) for gfx950 with hipcc.

Usage: python jax-sgmc_amd/build.py   (or __graft_entry__.build()).
The library lands next to this file, in-tree, so it travels with the repository.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsgmc.so")
ARCH = os.environ.get("SGMC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # Keep the operation order of every float expression as written (no fused
    # multiply-add contraction), and IEEE-rounded float32 division and sqrt.
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # No packed-FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources():
  return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False):
  hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
  tmp = LIB + ".tmp"
  cmd = [hipcc] + FLAGS + ["-I", os.path.join(HERE, "..", "include")] + sources() + ["-o", tmp]
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
    raise RuntimeError("hipcc failed building libsgmc.so")
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
