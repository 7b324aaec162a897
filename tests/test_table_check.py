"""Host-side check of fused chain tables (sgmc_glm_check_table, CPU only): the fused
kernels read the block map without bounds checks, so every table FusedGLM uploads is
checked first. A well-formed table passes; each kind of malformed table (a block
mapped to the wrong chain, a range outside the grid, ranges that do not cover the
grid, a loop table with a multi-block chain, too many rows per block, a wrong D class)
is refused with SGMC_E_ARG before any launch. No kernel is launched."""
import ctypes

import numpy as np
import pytest

import __graft_entry__

__graft_entry__.load()
from jax_sgmc_amd import _lib as L  # noqa: E402

# 16-byte aligned host buffer whose address stands in for every pointer of a chain
# (the check reads no memory through them)
_BUF = (ctypes.c_float * 64)()
_ADDR = (ctypes.addressof(_BUF) + 15) // 16 * 16


def table(nblocks_chain, D=32, n=256, N=1000, begins=None):
  C = len(nblocks_chain)
  arr = (L.GlmChain * C)()
  begin = np.concatenate([[0], np.cumsum(nblocks_chain)[:-1]]) if begins is None else begins
  for i in range(C):
    e = arr[i]
    for f in ("X", "y", "theta", "key", "dkey", "U", "partial", "counter"):
      setattr(e, f, _ADDR)
    e.N, e.ldx, e.D, e.n = N, 4 * ((D + 3) // 4), D, n
    e.block_begin, e.nblocks = int(begin[i]), int(nblocks_chain[i])
    e.dclass = L.query("sgmc_glm_dclass", D)
    e.temperature = 1.0
  return arr


def bmap(values):
  return (ctypes.c_int32 * len(values))(*values)


def check(arr, C, bc, nblocks, loop=0):
  return L.call("sgmc_glm_check_table", arr, C, bc, nblocks, loop)


def test_well_formed_tables_pass():
  assert check(table([2, 2, 2]), 3, None, 6) == 0
  assert check(table([1, 3, 2]), 3, bmap([0, 1, 1, 1, 2, 2]), 6) == 0
  # blocks of the heaviest chain first (FusedGLM's grid order)
  arr = table([1, 3], begins=[3, 0])
  assert check(arr, 2, bmap([1, 1, 1, 0]), 4) == 0
  assert check(table([1, 1, 1, 1]), 4, None, 4, loop=1) == 0
  # a loop launch ignores block_begin (FusedGLM orders blocks by bytes per block)
  assert check(table([1, 1, 1], begins=[2, 0, 1]), 3, None, 3, loop=1) == 0


@pytest.mark.parametrize("case", ["wrong_owner", "outside_grid", "gap", "uniform_begin",
                                  "loop_multiblock", "rows_per_block", "dclass", "null",
                                  "misaligned"])
def test_malformed_tables_are_refused(case):
  bc, nb, loop, C = None, 6, 0, 3
  if case == "wrong_owner":
    arr, bc = table([1, 3, 2]), bmap([0, 1, 2, 1, 2, 2])
  elif case == "outside_grid":
    arr, bc = table([2, 2, 3]), bmap([0, 0, 1, 1, 2, 2])
  elif case == "gap":
    arr, bc, nb = table([1, 2, 2]), bmap([0, 1, 1, 2, 2, 2]), 6
  elif case == "uniform_begin":
    arr = table([2, 2, 2], begins=[0, 3, 2])
  elif case == "loop_multiblock":
    arr, nb, loop = table([1, 2, 1]), 4, 1
  elif case == "rows_per_block":
    arr, nb, C = table([1, 1], n=4097), 2, 2
  elif case == "dclass":
    arr = table([2, 2, 2])
    arr[1].dclass = 4
  elif case == "null":
    arr = table([2, 2, 2])
    arr[2].partial = None
  else:
    arr = table([2, 2, 2])
    arr[0].X = _ADDR + 4
  with pytest.raises(ValueError, match="sgmc_glm_check_table"):
    check(arr, C, bc, nb, loop)


def test_blocks_for_respects_the_lds_rows():
  from jax_sgmc_amd import fused
  # one chain of 20000 rows with a huge byte target still gets >= 5 blocks
  assert fused.blocks_for(20000, 1, 1 << 40) == 5
  assert fused.blocks_for(4096, 512, 1 << 40) == 1
