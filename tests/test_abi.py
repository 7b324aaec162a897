"""The C-ABI library (CPU-only checks): it loads, exports every function that
include/sgmc.h declares, the ctypes binding declares each of them, and the ctypes
mirrors of the header's structs have the C layout (sizes and field offsets from gcc's
offsetof). No kernel is launched."""
import ctypes
import os
import re
import subprocess

import pytest

import __graft_entry__

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sgmc.h")

def _binding():
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L
  if not os.path.exists(L.LIB_PATH):
    __graft_entry__.build()
  return L


def header_functions():
  text = open(HEADER).read()
  text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
  names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(sgmc_\w+)\s*\(", text, flags=re.M)
  return sorted(set(names))


def test_header_declares_functions():
  names = header_functions()
  assert len(names) > 30
  assert "sgmc_glm_fused_step" in names and "sgmc_mlp_step" in names


def test_library_exports_every_header_function():
  L = _binding()
  lib = ctypes.CDLL(L.LIB_PATH)
  missing = [n for n in header_functions() if not hasattr(lib, n)]
  assert not missing, f"not exported: {missing}"


def test_binding_declares_every_header_function():
  L = _binding()
  names = set(L.exported_names())
  missing = [n for n in header_functions() if n not in names]
  assert not missing, f"no ctypes signature: {missing}"
  lib = L.load()
  assert lib.sgmc_abi_version() >= 1


STRUCTS = {
    "sgmc_glm_chain": "GlmChain",
    "sgmc_swap_pair": "SwapPair",
    "sgmc_mlp_state": "MlpState",
}


def test_struct_layouts_match_the_header(tmp_path):
  L = _binding()
  lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sgmc.h"', "int main(void) {"]
  for cname, pyname in STRUCTS.items():
    cls = getattr(L, pyname)
    lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
    for field, _ in cls._fields_:
      lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
  lines.append("return 0; }")
  src = tmp_path / "layout.c"
  src.write_text("\n".join(lines))
  exe = tmp_path / "layout"
  subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                 check=True)
  out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
  c = {}
  for line in out:
    if line:
      name, field, value = line.split()
      c[(name, field)] = int(value)
  for cname, pyname in STRUCTS.items():
    cls = getattr(L, pyname)
    assert c[(cname, "size")] == ctypes.sizeof(cls), cname
    for field, _ in cls._fields_:
      assert c[(cname, field)] == getattr(cls, field).offset, (cname, field)
