"""GEMM tiling plan of the MLP engine (sgmc_mlp_gemm_plan, host only): the built-in
k-tile depth and split-K choice, tuned overrides, and the rejection of tunings that do
not fit the shapes or the precision (CPU; the library is loaded, no kernel runs)."""
import ctypes

import pytest

from jax_sgmc_amd import _lib as L

C3 = dict(D_in=3072, H1=512, H2=512, n=256)
PREC = {"bf16": 0, "fp8": 1, "fp32": 2}


def plan(shape, C, prec, update=True, tune=None):
  out = (ctypes.c_int32 * 5)()
  t = None if tune is None else (ctypes.c_int32 * 5)(*tune)
  L.call("sgmc_mlp_gemm_plan", shape["D_in"], shape["H1"], shape["H2"], shape["n"], C,
         PREC[prec], 1 if update else 0, t, out)
  return [(v >> 4, v & 15) for v in out]


def test_builtin_choice():
  # K = 3072, 512, 512, 256, 256: deep tiles everywhere except the bf16 update GEMMs;
  # the forward grids of 16 tiles x 16 chains split K in two (512 blocks <= 1024).
  assert plan(C3, 16, "bf16") == [(128, 2), (128, 2), (128, 2), (32, 1), (32, 1)]
  assert plan(C3, 16, "bf16", update=False) == [(128, 2)] * 3 + [(128, 1)] * 2
  assert plan(C3, 16, "fp8") == [(128, 2)] * 3 + [(128, 1)] * 2
  assert plan(C3, 16, "fp32") == [(32, 2)] * 3 + [(32, 1)] * 2
  # 64 chains: 1024 forward blocks already, no split
  assert plan(C3, 64, "bf16")[:3] == [(128, 1)] * 3
  # K = 192: not a multiple of 128 -> 32 (bf16) / 64 (e4m3); 6 or 3 k-tiles
  odd = dict(D_in=192, H1=128, H2=128, n=64)
  assert plan(odd, 1, "bf16")[0] == (32, 2)
  assert plan(odd, 1, "fp8")[0] == (64, 1)


def test_tuned_entries_override_only_their_gemm():
  tune = [128 * 16 + 4, 0, 32 * 16 + 1, 128 * 16 + 1, 0]
  got = plan(C3, 16, "bf16", tune=tune)
  assert got == [(128, 4), (128, 2), (32, 1), (128, 1), (32, 1)]


@pytest.mark.parametrize("prec,tune,what", [
    ("bf16", [64 * 16 + 1, 0, 0, 0, 0], "k-tile 64"),        # bf16 takes 32 or 128
    ("fp8", [32 * 16 + 1, 0, 0, 0, 0], "k-tile 32"),         # e4m3 takes 64 or 128
    ("fp32", [128 * 16 + 1, 0, 0, 0, 0], "k-tile 128"),      # float32 takes 32
    ("bf16", [0, 0, 0, 32 * 16 + 2, 0], "split 2"),          # parameter GEMMs never split
    ("bf16", [32 * 16 + 5, 0, 0, 0, 0], "split 5"),
    ("bf16", [0, 128 * 16 + 8, 0, 0, 0], "split 8"),
])
def test_rejects_tunings_that_do_not_fit(prec, tune, what):
  with pytest.raises(ValueError, match="tune"):
    plan(C3, 16, prec, tune=tune)


def test_rejects_k_not_a_multiple_of_depth_times_split():
  shape = dict(D_in=192, H1=128, H2=128, n=64)
  with pytest.raises(ValueError, match="does not fit K = 192"):
    plan(shape, 1, "bf16", tune=[128 * 16 + 1, 0, 0, 0, 0])
  with pytest.raises(ValueError, match="does not fit K = 64"):
    plan(shape, 1, "fp8", tune=[0, 0, 0, 128 * 16 + 1, 0])  # gW2's K is the batch, 64
  plan(shape, 1, "bf16", tune=[32 * 16 + 2, 0, 0, 0, 0])     # 192 = 32 x 2 x 3
