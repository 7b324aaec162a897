"""GEMM tiling of the MLP engine: tuned k-tile depths and split-K factors (GPU).

- The k-tile depth leaves the order of the matrix instructions unchanged: for every
  GEMM, all depths with the same split-K factor give bit-equal potentials and
  gradients (bf16, e4m3, float32 operands).
- Split-K adds the partial products of the K slices in order, a different float32
  summation order: the gradient stays within 1e-5 (float32 operands) / 2e-3 (bf16) /
  2e-2 (e4m3, whose activation rounding amplifies last-bit changes) of the unsplit one
  in norm.
- autotune() restores every state tensor bit-exactly, its choice validates, and
  iterations with it follow iterations with the built-in plan.
Shapes: MLP 256-256-256-10, n = 128, 2 chains, 2048 rows."""
import numpy as np
import pytest
import torch

from jax_sgmc_amd import mlp as M

pytestmark = pytest.mark.gpu
DEV = "cuda"
DIMS, N, n, C = (256, 256, 256), 2048, 128, 2
SPLIT_TOL = {"fp32": 1e-5, "bf16": 2e-3, "fp8": 2e-2}


def engine(precision, integrator="sgld", rms=True, dims=DIMS):
  rng = np.random.default_rng(11)
  X = torch.as_tensor(rng.random((N, dims[0]), dtype=np.float32), device=DEV)
  labels = torch.as_tensor(rng.integers(0, 10, N).astype(np.int32), device=DEV)
  th0 = np.stack([M.init_params(40 + c, *dims) for c in range(C)])
  keys = np.array([[0, 3 + c] for c in range(C)], np.uint32)
  dkeys = np.array([[1, 8 + c] for c in range(C)], np.uint32)
  eps = np.full(16, 1e-4, np.float32)
  kw = dict(rms=rms) if integrator == "sgld" else dict(integrator=integrator, steps=2)
  return M.FusedMLP(X, labels, dims, n, th0, eps, keys=keys, dkeys=dkeys,
                    precision=precision, save_slots=np.arange(16), **kw)


def snapshot(eng):
  return [t.clone() for t in eng._mutable()]


def restore(eng, snap):
  for t, s in zip(eng._mutable(), snap):
    t.copy_(s)


@pytest.mark.parametrize("dims", [DIMS, (384, 256, 256)])  # 384 = 3 x 128: a three-way split
@pytest.mark.parametrize("precision", ["bf16", "fp8", "fp32"])
def test_tile_depth_bit_equal_and_split_within_rounding(precision, dims):
  print("MLP GEMM tilings, gradient mode:", precision, dims, flush=True)
  eng = engine(precision, dims=dims)
  eng.grad = torch.empty_like(eng.theta)  # before the snapshot: part of the state list
  snap = snapshot(eng)
  for slot in range(5):
    res = {}
    for code in eng.tile_candidates(slot):
      eng.tune = [0] * 5
      eng.tune[slot] = code
      restore(eng, snap)
      U, g = eng.potential_and_grad()
      torch.cuda.synchronize()
      res[code] = (U.cpu().numpy(), g.cpu().numpy())
    by_split = {}
    for code, r in res.items():
      by_split.setdefault(code & 15, []).append((code, r))
    for sp, group in by_split.items():
      code0, (U0, g0) = group[0]
      for code, (U, g) in group[1:]:
        assert np.array_equal(U, U0) and np.array_equal(g, g0), \
            f"slot {slot}: tilings {code0} and {code} differ at split {sp}"
    U1, g1 = by_split[1][0][1]
    for sp, group in by_split.items():
      U, g = group[0][1]
      rel = np.linalg.norm(g - g1) / np.linalg.norm(g1)
      assert rel <= SPLIT_TOL[precision], (slot, sp, rel)
      assert np.all(np.abs(U - U1) <= SPLIT_TOL[precision] * np.abs(U1)), (slot, sp, U, U1)


@pytest.mark.parametrize("precision,integrator", [("bf16", "sgld"), ("fp32", "sggmc")])
def test_autotune_restores_state_and_follows_the_builtin_plan(precision, integrator):
  print("MLP autotune:", precision, integrator, flush=True)
  eng = engine(precision, integrator)
  ref = engine(precision, integrator)
  eng.iteration()
  ref.iteration()
  before = snapshot(eng)
  plan = eng.autotune(reps=2)
  torch.cuda.synchronize()
  for t, s in zip(eng._mutable(), before):
    if t is not eng.work:
      assert torch.equal(t, s)
  assert plan == eng.gemm_plan() and all(k * 16 + s == t for (k, s), t in zip(plan, eng.tune))
  assert set(eng.tune_ms) >= {(i, eng.tune[i]) for i in range(3)} | {(3, eng.tune[3], eng.tune[4])}
  for _ in range(3):
    eng.iteration()
    ref.iteration()
  torch.cuda.synchronize()
  assert torch.equal(eng.keys, ref.keys) and torch.equal(eng.dkeys, ref.dkeys)
  assert torch.equal(eng.iter, ref.iter)
  d = (eng.theta - ref.theta).norm() / (ref.theta - torch.as_tensor(
      np.stack([M.init_params(40 + c, *DIMS) for c in range(C)]), device=DEV)).norm()
  # relative to the distance travelled from the initial point
  assert d.item() <= SPLIT_TOL[precision] * 10, d.item()


@pytest.mark.parametrize("precision,integrator,k_one,k_other", [
    ("bf16", "sgld", 32, 128), ("bf16", "sgld", 128, 32), ("fp8", "sggmc", 128, 64)])
def test_paired_parameter_gemms_equal_separate_launches(precision, integrator, k_one, k_other):
  """gW2 and gW1 at the same k-tile run in one launch (gemm_pair_kernel); at different
  depths in two. The k-tile does not change the results, so both give the same
  iterations bit for bit."""
  print("MLP paired parameter GEMMs:", precision, integrator, k_one, k_other, flush=True)
  a = engine(precision, integrator)
  b = engine(precision, integrator)
  a.tune = [0, 0, 0, k_one * 16 + 1, k_one * 16 + 1]
  b.tune = [0, 0, 0, k_one * 16 + 1, k_other * 16 + 1]
  for _ in range(3):
    a.iteration()
    b.iteration()
  torch.cuda.synchronize()
  for ta, tb in zip(a._mutable(), b._mutable()):
    if ta is not a.work:
      assert torch.equal(ta, tb)
  U1, g1 = a.potential_and_grad()
  U2, g2 = b.potential_and_grad()
  torch.cuda.synchronize()
  assert torch.equal(U1, U2) and torch.equal(g1, g2)
