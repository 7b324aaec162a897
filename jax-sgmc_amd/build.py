"""Builds libsgmc.so (all HIP sources under csrc/) for gfx950 with hipcc.

Usage: python jax-sgmc_amd/build.py   (or __graft_entry__.build()).
Each source compiles to its own object (in parallel, under _build/), then one link.
The library lands next to this file, in-tree, so it travels with the repository.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsgmc.so")
ARCH = os.environ.get("SGMC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
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


def _run(cmd, verbose):
  if verbose:
    print(" ".join(cmd), flush=True)
  res = subprocess.run(cmd, capture_output=True, text=True)
  if res.returncode != 0:
    sys.stderr.write(res.stdout + res.stderr)
  return res.returncode


def build(verbose=False, jobs=None):
  hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
  os.makedirs(OBJ, exist_ok=True)
  inc = ["-I", os.path.join(HERE, "..", "include")]
  srcs = sources()
  objs = [os.path.join(OBJ, os.path.basename(s)[:-4] + ".o") for s in srcs]
  cmds = [[hipcc] + FLAGS + inc + ["-c", s, "-o", o] for s, o in zip(srcs, objs)]
  jobs = jobs or min(len(cmds), os.cpu_count() or 1, 16)
  with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
    rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
  if any(rcs):
    raise RuntimeError("hipcc failed compiling " +
                       ", ".join(os.path.basename(s) for s, r in zip(srcs, rcs) if r))
  tmp = LIB + ".tmp"
  if _run([hipcc, "-shared", f"--offload-arch={ARCH}"] + objs + ["-o", tmp], verbose):
    raise RuntimeError("hipcc failed linking libsgmc.so")
  os.replace(tmp, LIB)
  return LIB


if __name__ == "__main__":
  print(build(verbose=True))
