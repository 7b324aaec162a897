"""Fused multi-chain path against the oracle (GPU): every saved iteration of every
chain within 1e-5 relative (floor 1e-5 max|theta|) of the float64 oracle run with the
same keys; keys bit-exact; graph replay identical to eager launches."""
import numpy as np
import pytest
import torch

from oracle import model as om
from oracle import steps as ost

from jax_sgmc_amd import fused
from jax_sgmc_amd import random as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _data(N, D, kind, seed):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  if kind == "logistic":
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  else:
    y = (X @ w + 0.5 * rng.standard_normal(N)).astype(np.float32)
  return X, y


def _dev(X):
  N, D = X.shape
  ldx = 4 * ((D + 3) // 4)
  Xp = np.zeros((N, ldx), np.float32)
  Xp[:, :D] = X
  return torch.as_tensor(Xp, device=DEV)


def _check(got, ref, what):
  ref = np.asarray(ref, np.float64)
  tol = 1e-5 * np.abs(ref) + 1e-5 * max(1e-3, np.max(np.abs(ref)))
  err = np.abs(np.asarray(got, np.float64) - ref)
  assert np.all(err <= tol), f"{what}: max rel {np.max(err / (np.abs(ref) + 1e-30))}"


def _run_case(kind, mode, specs, iters, eps, steps=1, friction=0.5, sigma=0.5, pscale=1.0,
              graph=0, alpha=0.9, lmbd=1e-5):
  chains, fams = [], []
  for i, s in enumerate(specs):
    X, y = _data(s["N"], s["D"], kind, seed=10 + i)
    rng = np.random.default_rng(100 + i)
    th0 = (0.1 * rng.standard_normal(s["D"])).astype(np.float32)
    chains.append(dict(X=_dev(X), y=torch.as_tensor(y, device=DEV), D=s["D"], n=s["n"],
                       theta0=th0, key=(0, i), dkey=(1, 7 * i + 3),
                       temperature=s.get("T", 1.0)))
    fams.append((X, y, th0))
  eng = fused.FusedGLM(chains, kind, mode, eps, steps=steps, friction=friction, sigma=sigma,
                       prior_scale=pscale, save_slots=np.arange(iters), alpha=alpha, lmbd=lmbd)
  eng.run(iters, graph_iterations=graph)
  torch.cuda.synchronize()
  for i, s in enumerate(specs):
    X, y, th0 = fams[i]
    fam = om.Family(kind, X, y, sigma=sigma)
    pot = lambda th, idx, fam=fam: om.stochastic_potential(fam, th, idx, pscale)
    key = np.array([0, i], np.uint32)
    dkey = np.array([1, 7 * i + 3], np.uint32)
    if mode == "sghmc":
      r = ost.sghmc(pot, th0, iters, eps, steps, friction, s["N"], s["n"], [s["D"]], key=key,
                    dkey=dkey)
      _check(eng.momentum(i).cpu().numpy(), r["p"][-1], f"chain {i} momentum")
    else:
      r = ost.sgld(pot, th0, iters, eps, s.get("T", 1.0), s["N"], s["n"], [s["D"]], key=key,
                   dkey=dkey, rms=(alpha, lmbd) if mode == "sgld_rms" else None)
    tr = eng.sample_trace(i).cpu().numpy()
    for t in range(iters):
      _check(tr[t], r["theta"][t], f"chain {i} iteration {t}")
    np.testing.assert_array_equal(R.key_words(eng.keys[i]), r["key"])
    np.testing.assert_array_equal(R.key_words(eng.dkeys[i]), r["dkey"])
    Uref = r["U"][-1]
    assert abs(eng.U[i].item() - Uref) <= 1e-5 * abs(Uref) + 1e-3, (eng.U[i].item(), Uref)
  assert eng.iteration_count() == iters
  return eng


def test_fused_sgld_linear_regression():
  print("fused sgld linear", flush=True)
  eps = ost.polynomial_step_size_first_last(2e-4, 5e-5, 0.33, 30)
  _run_case("linear_gaussian", "sgld", [dict(N=2000, D=32, n=128)] * 2, 30, eps,
            pscale=np.sqrt(10.0))


def test_fused_sgld_rms_logistic():
  print("fused sgld rms logistic", flush=True)
  eps = np.full(20, 1e-3, np.float32)
  _run_case("logistic", "sgld_rms", [dict(N=3000, D=64, n=100)], 20, eps)


def test_fused_sghmc_logistic_multiblock():
  print("fused sghmc logistic", flush=True)
  eps = np.full(6, 2e-3, np.float32)
  _run_case("logistic", "sghmc", [dict(N=20000, D=512, n=1024)] * 3, 6, eps, steps=5)


def test_fused_heterogeneous_table():
  print("fused heterogeneous", flush=True)
  specs = [dict(N=2000, D=16, n=64, T=1.0), dict(N=5000, D=64, n=200, T=2.5),
           dict(N=3000, D=256, n=96, T=1.0), dict(N=4000, D=1000, n=300, T=7.0)]
  eps = np.full(8, 1e-4, np.float32)
  _run_case("logistic", "sgld", specs, 8, eps)


def test_fused_graph_replay_equals_eager():
  print("fused graph", flush=True)
  eps = np.full(12, 2e-3, np.float32)
  spec = [dict(N=5000, D=128, n=256)] * 2
  e1 = _run_case("logistic", "sghmc", spec, 12, eps, steps=3, graph=4)
  e2 = _run_case("logistic", "sghmc", spec, 12, eps, steps=3, graph=0)
  assert torch.equal(e1.theta_flat, e2.theta_flat)
  assert torch.equal(e1.samples, e2.samples)


def test_fused_autotune_preserves_state():
  eps = np.full(4, 1e-3, np.float32)
  X, y = _data(4000, 512, "logistic", 0)
  ch = [dict(X=_dev(X), y=torch.as_tensor(y, device=DEV), D=512, n=1024, theta0=np.zeros(512),
             key=(0, c), dkey=(1, c)) for c in range(4)]
  a = fused.FusedGLM(ch, "logistic", "sghmc", eps, steps=2)
  b = fused.FusedGLM(ch, "logistic", "sghmc", eps, steps=2)
  a.autotune(reps=1)
  a.run(4)
  b.run(4)
  torch.cuda.synchronize()
  # block counts differ, so sums are reordered: equal to float32 rounding only
  np.testing.assert_allclose(a.theta_flat.cpu().numpy(), b.theta_flat.cpu().numpy(),
                             rtol=1e-4, atol=1e-5)
