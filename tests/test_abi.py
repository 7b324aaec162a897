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
  assert lib.sgmc_abi_version() == L.ABI_VERSION


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


def _prototypes():
  """name -> (return type, [parameter types]) parsed from the header."""
  text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
  out = {}
  for m in re.finditer(r"^\s*((?:const\s+)?[A-Za-z_][\w\s\*]*?)\b(sgmc_\w+)\s*\(([^;]*?)\)\s*;",
                       text, flags=re.M | re.S):
    ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
    types = []
    if params and params != "void":
      for p in params.split(","):
        p = " ".join(p.split())
        # drop the parameter name (the last identifier), keep the type and its stars
        types.append(re.sub(r"\b\w+$", "", p).strip())
    out[name] = (ret, types)
  return out


def _kind_c(t):
  if "*" in t:
    return "ptr"
  t = t.replace("const", "").strip()
  return {"int": "i32", "int32_t": "i32", "unsigned": "u32", "uint32_t": "u32",
          "int64_t": "i64", "float": "f32", "double": "f64"}[t]


def _kind_ctypes(c):
  if c in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(c, ctypes._Pointer):
    return "ptr"
  return {ctypes.c_int: "i32", ctypes.c_int32: "i32", ctypes.c_uint32: "u32",
          ctypes.c_int64: "i64", ctypes.c_float: "f32", ctypes.c_double: "f64"}[c]


def test_binding_argument_types_match_the_header():
  """Every ctypes signature has the header prototype's arity and, parameter by
  parameter, the same kind (pointer, 32/64-bit integer, float, double): a mismatch
  would pass arguments in the wrong registers without any error."""
  L = _binding()
  sigs = dict(L._SIGNATURES)
  sigs.update(L.EXTRA_SIGNATURES)
  protos = _prototypes()
  assert set(protos) == set(header_functions())
  bad = []
  for name, (ret, params) in protos.items():
    restype, argtypes = sigs[name]
    if [_kind_c(p) for p in params] != [_kind_ctypes(a) for a in argtypes]:
      bad.append((name, params, [a.__name__ for a in argtypes]))
    if ret != "void" and _kind_c(ret) != _kind_ctypes(restype):
      bad.append((name, ret, restype.__name__))
  assert not bad, bad
