"""Samplers through the reference's alias API on the GPU: trajectory parity with the
oracle (first 100 steps within 1e-5 relative), fused multi-chain path == sequential
modular path, the reference's statistical fixtures (tests/test_alias.py:32-65), and
the analytic posterior of Bayesian linear regression."""
import math

import numpy as np
import pytest
import torch
from scipy import stats

from oracle import jaxrandom as jr
from oracle import model as om
from oracle import posterior as op
from oracle import steps as ost

from jax_sgmc_amd import alias
from jax_sgmc_amd import data as DT
from jax_sgmc_amd import fused as FU
from jax_sgmc_amd import likelihoods as LK
from jax_sgmc_amd import potential as PT
from jax_sgmc_amd import random as R
from jax_sgmc_amd import scheduler as SC

pytestmark = pytest.mark.gpu
DEV = "cuda"


def linreg(N=1000, D=8, seed=0, sigma=0.5):
  rng = np.random.default_rng(seed)
  X = rng.standard_normal((N, D)).astype(np.float32)
  w = rng.standard_normal(D)
  y = (X @ w + sigma * rng.standard_normal(N)).astype(np.float32)
  return X, y


def logreg(N=4000, D=32, seed=1):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  return X, y


def _close(got, ref, what, rtol=1e-5):
  ref = np.asarray(ref, np.float64)
  got = np.asarray(got, np.float64)
  tol = rtol * np.abs(ref) + rtol * max(1e-3, float(np.max(np.abs(ref))))
  bad = np.abs(got - ref) > tol
  assert not np.any(bad), f"{what}: max rel {np.max(np.abs(got - ref) / (np.abs(ref) + 1e-30))}"


@pytest.mark.parametrize("fused", [False, True])
def test_sgld_linear_regression_matches_oracle(fused):
  """First 100 SGLD steps on linear regression (config-1 shape, smaller N): every
  saved sample and potential within 1e-5 of the oracle with the same keys."""
  print("sgld parity fused=", fused, flush=True)
  X, y = linreg(N=2000, D=32)
  sigma, tau = 0.5, math.sqrt(10.0)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(tau), LK.LinearGaussian(sigma), strategy="vmap")
  its, first, last = 100, 1e-4, 2e-5
  run = alias.sgld(pot, loader, cache_size=1, batch_size=128, first_step_size=first,
                   last_step_size=last, burn_in=0, accepted_samples=its)
  res = run(torch.zeros(32), iterations=its, fused=fused)[0]
  assert res["sample_count"] == its
  eps = SC.polynomial_table(first, last, its)
  fam = om.Family("linear_gaussian", X, y, sigma=sigma)
  potf = lambda th, idx: om.stochastic_potential(fam, th, idx, tau)
  r = ost.sgld(potf, np.zeros(32), its, eps, 1.0, 2000, 128, [32])
  _close(res["samples"]["variables"], r["theta"], "theta")
  _close(-np.asarray(res["samples"]["likelihood"]), r["U"], "potential")


@pytest.mark.parametrize("fused", [False, True])
def test_re_sgld_matches_oracle(fused):
  """alias.re_sgld: the normal chain's saved samples and potentials over 40 iterations
  within 1e-5 of the oracle's parallel_tempering composition with the same keys, the
  exchange decisions included."""
  print("re_sgld parity fused=", fused, flush=True)
  N, D, n, its, tau = 500, 8, 16, 40, 3.0
  X, y = linreg(N=N, D=D, seed=5)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.LinearGaussian(0.5), strategy="vmap")
  first, last = 2e-3, 5e-4
  run = alias.re_sgld(pot, loader, cache_size=1, batch_size=n, first_step_size=first,
                      last_step_size=last, burn_in=0, accepted_samples=its, temperature=tau)
  xn = np.zeros(D, np.float32)
  xh = np.full(D, 0.3, np.float32)
  res = run((torch.as_tensor(xn, device=DEV), torch.as_tensor(xh, device=DEV)),
            iterations=its, fused=fused)[0]
  eps = SC.polynomial_table(first, last, its)
  fam = om.Family("linear_gaussian", X, y, sigma=0.5)
  potf = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  r = ost.resgld(potf, xn, xh, its, eps, tau, N, n, [D])
  print("oracle exchanges", int(r["swaps"].sum()), "of", its, flush=True)
  assert 0 < r["swaps"].sum() < its  # both outcomes occur in this run
  _close(res["samples"]["variables"], r["samples"], "theta")
  _close(-np.asarray(res["samples"]["likelihood"]), r["samples_U"], "potential")


@pytest.mark.parametrize("fused", [False, True])
def test_sgld_rmsprop_matches_oracle(fused):
  print("rms parity fused=", fused, flush=True)
  X, y = logreg(N=3000, D=16)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  its = 60
  run = alias.sgld(pot, loader, cache_size=1, batch_size=64, first_step_size=1e-3,
                   last_step_size=1e-4, accepted_samples=its, rms_prop=True)
  res = run(torch.zeros(16), iterations=its, fused=fused)[0]
  eps = SC.polynomial_table(1e-3, 1e-4, its)
  fam = om.Family("logistic", X, y)
  r = ost.sgld(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0), np.zeros(16), its,
               eps, 1.0, 3000, 64, [16], rms=(0.9, 1e-5))
  _close(res["samples"]["variables"], r["theta"], "theta")


@pytest.mark.parametrize("fused", [False, True])
def test_sghmc_logistic_matches_oracle(fused):
  print("sghmc parity fused=", fused, flush=True)
  X, y = logreg(N=5000, D=64)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  its = 20
  run = alias.sghmc(pot, loader, cache_size=1, batch_size=256, integration_steps=5,
                    friction=0.5, first_step_size=5e-3, last_step_size=1e-3,
                    accepted_samples=its)
  keys = [R.PRNGKey(3), R.PRNGKey(4)]
  dkeys = [R.PRNGKey(5), R.PRNGKey(6)]
  res = run(torch.zeros(64), torch.full((64,), 0.1), iterations=its, keys=keys,
            data_keys=dkeys, fused=fused)
  eps = SC.polynomial_table(5e-3, 1e-3, its)
  fam = om.Family("logistic", X, y)
  for c, th0 in enumerate([np.zeros(64), np.full(64, 0.1)]):
    r = ost.sghmc(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0), th0, its, eps, 5,
                  0.5, 5000, 256, [64], key=jr.PRNGKey(3 + c), dkey=jr.PRNGKey(5 + c))
    _close(res[c]["samples"]["variables"], r["theta"], f"chain {c} theta")
    _close(res[c]["samples"]["energy"], r["U"], f"chain {c} energy")


@pytest.mark.parametrize("fused", [True, False])
def test_obabo_matches_oracle(fused):
  print("obabo parity fused", fused, flush=True)
  X, y = logreg(N=3000, D=24)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  its = 15
  run = alias.obabo(pot, loader, cache_size=1, batch_size=128, integration_steps=3,
                    friction=2.0, first_step_size=5e-3, last_step_size=1e-3,
                    accepted_samples=its)
  res = run(torch.zeros(24), iterations=its, fused=fused)[0]
  eps = SC.polynomial_table(5e-3, 1e-3, its)
  fam = om.Family("logistic", X, y)
  r = ost.obabo(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0), np.zeros(24), its,
                eps, 3, 2.0, 3000, 128, [24])
  _close(res["samples"]["variables"], r["theta"], "theta")
  _close(res["samples"]["energy"], r["U"], "energy")


def test_generic_torch_likelihood_matches_family_kernel():
  """An arbitrary torch likelihood (generic path) and the recognized family (kernel
  path) give the same chain."""
  X, y = linreg(N=1500, D=6)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)

  def lik(sample, obs):
    r = obs["y"] - obs["x"] @ sample["w"]
    return (math.log(2 * math.pi * 0.25) + r * r / 0.25) / -2.0

  def prior(sample):
    return torch.sum((math.log(2 * math.pi * 10.0) + sample["w"] ** 2 / 10.0) / -2.0)

  p_gen = PT.minibatch_potential(prior, lik, strategy="vmap")
  p_fam = PT.minibatch_potential(LK.GaussianPrior(math.sqrt(10.0)), LK.LinearGaussian(0.5),
                                 strategy="vmap")
  a = alias.sgld(p_gen, loader, cache_size=1, batch_size=50, first_step_size=1e-4,
                 last_step_size=1e-5, accepted_samples=30)({"w": torch.zeros(6)},
                                                           iterations=30)[0]
  b = alias.sgld(p_fam, loader, cache_size=1, batch_size=50, first_step_size=1e-4,
                 last_step_size=1e-5, accepted_samples=30)({"w": torch.zeros(6)},
                                                           iterations=30, fused=False)[0]
  _close(a["samples"]["variables"]["w"], b["samples"]["variables"]["w"], "generic vs kernel",
         rtol=2e-5)


def _ks_setup():
  """tests/test_alias.py:32-65: 1-D Gaussian target, sigma 0.5, 2-row dummy data."""
  sigma = 0.5
  x = sigma * jr.normal(jr.PRNGKey(11), (100,))
  loader = DT.DeviceNumpyDataLoader(x=np.zeros((2, 1), np.float32))

  def lik(sample, _):
    return -0.5 * (sample / sigma) ** 2

  pot = PT.minibatch_potential(lambda s: torch.zeros((), device=DEV), lik, strategy="vmap")
  full = PT.full_potential(lambda s: torch.zeros((), device=DEV), lik, strategy="vmap")

  def check(sampled):
    p = stats.kstest(np.asarray(sampled).reshape(-1), x).pvalue
    assert p > 0.05, f"KS p-value {p}"

  return loader, pot, full, torch.tensor(float(x[0]), device=DEV), check


def test_ks_sgld_rms():
  loader, pot, _, x0, check = _ks_setup()
  run = alias.sgld(pot, loader, cache_size=1, batch_size=1, first_step_size=0.5,
                   last_step_size=0.1, burn_in=100, accepted_samples=100, rms_prop=True)
  check(run(x0, iterations=500)[0]["samples"]["variables"])


def test_ks_re_sgld():
  loader, pot, _, x0, check = _ks_setup()
  run = alias.re_sgld(pot, loader, cache_size=1, batch_size=1, first_step_size=0.5,
                      last_step_size=0.1, burn_in=100, accepted_samples=100, temperature=1.5)
  check(run((x0, x0), iterations=500)[0]["samples"]["variables"])


def test_ks_obabo():
  loader, pot, _, x0, check = _ks_setup()
  run = alias.obabo(pot, loader, cache_size=1, batch_size=1, first_step_size=0.5,
                    last_step_size=0.1, friction=10, burn_in=100, accepted_samples=100)
  check(run(x0, iterations=500)[0]["samples"]["variables"])


def test_ks_sghmc():
  loader, pot, _, x0, check = _ks_setup()
  run = alias.sghmc(pot, loader, cache_size=1, batch_size=1, integration_steps=5,
                    first_step_size=0.5, last_step_size=0.1, friction=0.995, burn_in=100,
                    accepted_samples=100)
  check(run(x0, iterations=500)[0]["samples"]["variables"])


def test_ks_sggmc():
  loader, pot, full, x0, check = _ks_setup()
  run = alias.sggmc(pot, full, loader, cache_size=1, batch_size=1, first_step_size=0.5,
                    last_step_size=0.1, burn_in=100)
  check(run(x0, iterations=200)[0]["samples"]["variables"])


def test_sgld_posterior_moments_linear_regression():
  """Config-1 model at reduced N: SGLD at a fixed step size samples a Gaussian whose
  mean is the posterior mean and whose covariance is the exact Lyapunov stationary
  covariance; 256 fused chains, Monte-Carlo error bounds stated in the asserts."""
  print("posterior", flush=True)
  N, D, n, sigma, tau2 = 2000, 4, 128, 0.5, 10.0
  X, y = linreg(N=N, D=D, seed=3, sigma=sigma)
  mu, S = op.sgld_stationary(X, y, sigma, tau2, eps=2e-4, n=n)
  Xd = torch.as_tensor(X, device=DEV)
  yd = torch.as_tensor(y, device=DEV)
  C, burn, keep = 256, 300, 200
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(1), C))
  chains = [dict(X=Xd, y=yd, D=D, n=n, theta0=mu.astype(np.float32), key=tuple(keys[c]),
                 dkey=tuple(dkeys[c])) for c in range(C)]
  its = burn + keep
  slots = np.full(its, -1, np.int32)
  slots[burn:] = np.arange(keep)
  eng = FU.FusedGLM(chains, "linear_gaussian", "sgld", np.full(its, 2e-4, np.float32),
                    sigma=sigma, prior_scale=math.sqrt(tau2), save_slots=slots)
  eng.run(its, graph_iterations=50)
  torch.cuda.synchronize()
  draws = np.concatenate([eng.sample_trace(c).cpu().numpy()[::20] for c in range(C)])
  m = draws.mean(0)
  sd = np.sqrt(np.diag(S))
  # 2560 nearly independent draws (thinned by 20 >> autocorrelation time)
  se = sd / np.sqrt(draws.shape[0])
  assert np.all(np.abs(m - mu) < 5 * se), (m - mu) / se
  emp = np.cov(draws.T)
  rel = np.abs(np.diag(emp) / np.diag(S) - 1)
  # variance estimate of 2560 draws: relative standard error sqrt(2/2560) = 2.8%
  assert np.all(rel < 0.15), rel


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp8"])
def test_sgld_mlp_engine_matches_generic_path(precision):
  """alias.sgld(rms_prop=True) on an MLPSoftmax potential: the MLP engine (the default)
  against the generic float32 torch path (fused=False) with the same keys. The noise
  and minibatch indices are bit-identical, so the samples differ only through the GEMM
  gradients: per leaf, within 1e-3 (fp32), 5e-2 (bf16) or 0.2 (fp8) of the generic
  path's whole displacement from the start. Without ``fused=`` and ``precision=`` the
  alias runs the fp32 engine (bit for bit the explicit call)."""
  from jax_sgmc_amd import mlp as ML
  print("alias mlp engine", precision, flush=True)
  dims = (128, 128, 128)
  rng = np.random.default_rng(0)
  N = 1000
  X = rng.random((N, dims[0]), dtype=np.float32)
  y = rng.integers(0, 10, N).astype(np.int32)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), ML.MLPSoftmax(), strategy="vmap")
  run = alias.sgld(pot, loader, cache_size=1, batch_size=64, first_step_size=2e-4,
                   last_step_size=1e-4, accepted_samples=3, rms_prop=True)
  flat = ML.init_params(7, *dims)
  shapes = ML.leaf_shapes(*dims)
  p0, off = {}, 0
  for k in ("W1", "W2", "W3", "b1", "b2", "b3"):
    size = int(np.prod(shapes[k]))
    p0[k] = torch.as_tensor(flat[off:off + size].reshape(shapes[k]), device=DEV)
    off += size
  a = run(p0, iterations=3, fused=True, precision=precision)[0]["samples"]["variables"]
  b = run(p0, iterations=3, fused=False)[0]["samples"]["variables"]
  if precision == "fp32":
    dflt = run(p0, iterations=3)[0]["samples"]["variables"]
    for k in p0:
      np.testing.assert_array_equal(dflt[k], a[k])
  tol = {"fp32": 1e-3, "bf16": 5e-2, "fp8": 0.2}[precision]
  for k in p0:
    start = p0[k].cpu().numpy()[None]
    d_ref = np.linalg.norm(b[k] - start)
    d = np.linalg.norm(a[k] - b[k])
    print(k, d, d_ref, flush=True)
    assert d <= tol * d_ref + 1e-6, (k, d, d_ref)


def test_sgld_mlp_alias_default_runs_the_engine():
  """alias.sgld(rms_prop=True) on an MLPSoftmax potential with neither ``fused=`` nor
  ``precision=``: the float32 MLP engine runs (its result equals the explicit engine
  call bit for bit, and the samples move). Profiled alone, this test launches the
  engine's kernels and no library GEMM."""
  from jax_sgmc_amd import mlp as ML
  print("alias mlp default", flush=True)
  dims = (128, 128, 128)
  rng = np.random.default_rng(3)
  N = 1000
  X = rng.random((N, dims[0]), dtype=np.float32)
  y = rng.integers(0, 10, N).astype(np.int32)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), ML.MLPSoftmax(), strategy="vmap")
  run = alias.sgld(pot, loader, cache_size=1, batch_size=64, first_step_size=2e-4,
                   last_step_size=1e-4, accepted_samples=3, rms_prop=True)
  flat = ML.init_params(11, *dims)
  shapes = ML.leaf_shapes(*dims)
  p0, off = {}, 0
  for k in ("W1", "W2", "W3", "b1", "b2", "b3"):
    size = int(np.prod(shapes[k]))
    p0[k] = torch.as_tensor(flat[off:off + size].reshape(shapes[k]), device=DEV)
    off += size
  dflt = run(p0, iterations=3)[0]["samples"]["variables"]
  expl = run(p0, iterations=3, fused=True, precision="fp32")[0]["samples"]["variables"]
  for k in p0:
    np.testing.assert_array_equal(dflt[k], expl[k])
    assert np.all(np.isfinite(dflt[k]))
    assert not np.array_equal(dflt[k][-1], p0[k].cpu().numpy())


def test_sghmc_many_chains_two_tables_match_generic_path():
  """alias.sghmc with 16 multi-block chains runs as two chain tables on two streams;
  every chain equals the same chain through the generic kernels."""
  print("sghmc 16 chains, two tables", flush=True)
  X, y = logreg(N=4000, D=64, seed=11)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  run = alias.sghmc(pot, loader, cache_size=1, batch_size=1024, integration_steps=3,
                    friction=0.5, first_step_size=2e-3, last_step_size=1e-3, accepted_samples=5)
  C = 16
  # the case takes the two-table path: enough chains, several blocks per chain
  tb = max(64 * 1024, C * 1024 * 4 * 64 // 2048)
  assert C >= alias._TABLE_SPLIT_CHAINS and FU.blocks_for(1024, 64, tb) > 1
  keys = [R.PRNGKey(100 + c, DEV) for c in range(C)]
  dkeys = [R.PRNGKey(200 + c, DEV) for c in range(C)]
  inits = [torch.as_tensor(0.02 * np.random.default_rng(c).standard_normal(64).astype(np.float32))
           for c in range(C)]
  a = run(*inits, iterations=6, keys=keys, data_keys=dkeys, fused=True)
  b = run(*inits, iterations=6, keys=keys, data_keys=dkeys, fused=False)
  for c in range(C):
    _close(a[c]["samples"]["variables"], b[c]["samples"]["variables"], f"chain {c}")
    _close(a[c]["samples"]["energy"], b[c]["samples"]["energy"], f"energy {c}")
