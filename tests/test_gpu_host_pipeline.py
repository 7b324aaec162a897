"""Host data pipeline into the HIP kernels, and the GLM potential as a differentiable
torch value (GPU).

- NumpyDataLoader (data/numpy_loader.py:147-460): host-drawn minibatches (numpy PCG64,
  with replacement, shuffled, shuffled in epochs) copied to the device and evaluated by
  the sgmc_glm_potential kernel give the oracle's potential and gradient on the rows of
  the oracle's index stream (oracle/loader.py); alias.sgld over the host loader
  (cache of 4 batches, data/core.py:664-835) follows the oracle's SGLD on that stream.
- StochasticPotential.differentiable: U.backward() returns the kernel's gradient
  (bit-equal to value_and_grad), which matches float64 central differences of the
  oracle's potential."""
import math

import numpy as np
import pytest
import torch

from oracle import loader as ol
from oracle import model as om
from oracle import steps as ost

from jax_sgmc_amd import alias, data as DT, likelihoods as LK, potential as PT
from jax_sgmc_amd import scheduler as SC

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _data(N, D, kind, seed=0):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  if kind == "logistic":
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  else:
    y = (X @ w + 0.5 * rng.standard_normal(N)).astype(np.float32)
  return X, y


def _close(got, ref, what, rtol=1e-5):
  ref = np.asarray(ref, np.float64)
  got = np.asarray(got, np.float64)
  tol = rtol * np.abs(ref) + rtol * max(1e-3, float(np.max(np.abs(ref))))
  assert np.all(np.abs(got - ref) <= tol), \
      f"{what}: max rel {np.max(np.abs(got - ref) / (np.abs(ref) + 1e-30))}"


@pytest.mark.parametrize("mode", ["replace", "shuffle", "epochs"])
def test_host_batches_through_the_kernel(mode):
  print("host batches through sgmc_glm_potential:", mode, flush=True)
  N, D, n, cache, batches = 1000, 16, 96, 3, 14
  X, y = _data(N, D, "logistic", seed=3)
  loader = DT.NumpyDataLoader(x=X, y=y)
  init_fn, batch_fn, _ = DT.random_reference_data(loader, cache, n)
  kw = {"replace": {}, "shuffle": {"shuffle": True},
        "epochs": {"shuffle": True, "in_epochs": True}}[mode]
  state = init_fn(**kw)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.5), LK.Logistic(), strategy="vmap")
  idx, _ = ol.index_stream(N, n, batches, mode, seed=0)
  fam = om.Family("logistic", X, y)
  rng = np.random.default_rng(9)
  for b in range(batches):
    state, rd = batch_fn(state, information=True)
    assert rd[0]["x"].is_cuda and pot._fast_ok(torch.zeros(D, device=DEV), rd, None, None)
    w = (0.5 * rng.standard_normal(D)).astype(np.float32)
    (U, _), g = pot.value_and_grad(torch.as_tensor(w, device=DEV), rd)
    Ur, gr, _ = om.stochastic_potential(fam, w, idx[b], 1.5)
    _close(U.item(), Ur, f"batch {b} U")
    _close(g.cpu().numpy(), gr, f"batch {b} grad")


def test_sgld_over_host_loader_matches_oracle():
  """alias.sgld over a NumpyDataLoader (cache of 4 batches, refilled from the host):
  every saved sample within 1e-5 of the oracle's SGLD on the host index stream."""
  print("alias.sgld over NumpyDataLoader", flush=True)
  N, D, n, its = 2000, 16, 64, 30
  X, y = _data(N, D, "linear")
  sigma, tau = 0.5, math.sqrt(10.0)
  loader = DT.NumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(tau), LK.LinearGaussian(sigma), strategy="vmap")
  first, last = 2e-4, 5e-5
  run = alias.sgld(pot, loader, cache_size=4, batch_size=n, first_step_size=first,
                   last_step_size=last, accepted_samples=its)
  res = run(torch.zeros(D), iterations=its)[0]
  idx, _ = ol.index_stream(N, n, its, "replace", seed=0)
  stream = iter(idx)
  fam = om.Family("linear_gaussian", X, y, sigma=sigma)
  potf = lambda th, _idx: om.stochastic_potential(fam, th, next(stream), tau)
  r = ost.sgld(potf, np.zeros(D), its, SC.polynomial_table(first, last, its), 1.0, N, n, [D])
  _close(res["samples"]["variables"], r["theta"], "theta")
  _close(-np.asarray(res["samples"]["likelihood"]), r["U"], "potential")


@pytest.mark.parametrize("kind", ["logistic", "linear"])
def test_glm_potential_autograd_function(kind):
  print("GLM potential autograd:", kind, flush=True)
  N, D, n = 3000, 64, 256
  X, y = _data(N, D, kind, seed=4)
  rng = np.random.default_rng(5)
  idx = rng.integers(0, N, n)
  Xb = torch.as_tensor(X[idx], device=DEV)
  yb = torch.as_tensor(y[idx], device=DEV)
  info = DT.MiniBatchInformation(observation_count=N, mask=None, batch_size=n)
  lik = LK.Logistic() if kind == "logistic" else LK.LinearGaussian(0.5)
  pot = PT.minibatch_potential(LK.GaussianPrior(2.0), lik, strategy="vmap", temperature=1.5)
  w0 = (0.3 * rng.standard_normal(D)).astype(np.float32)
  w = torch.as_tensor(w0, device=DEV).requires_grad_(True)
  U = pot.differentiable(w, ({"x": Xb, "y": yb}, info))
  (3.0 * U).backward()
  (U2, _), g = pot.value_and_grad(torch.as_tensor(w0, device=DEV), ({"x": Xb, "y": yb}, info))
  assert U.item() == U2.item()
  assert torch.equal(w.grad, 3.0 * g)
  fam = om.Family("logistic" if kind == "logistic" else "linear_gaussian", X, y, sigma=0.5)
  f = lambda th: om.stochastic_potential(fam, th, idx, 2.0, 1.5)[0]
  h = 1e-4
  gk = g.cpu().numpy().astype(np.float64)
  for j in (0, 7, 31, 63):
    e = np.zeros(D)
    e[j] = h
    fd = (f(w0 + e) - f(w0 - e)) / (2 * h)
    assert abs(gk[j] - fd) <= 1e-4 * max(abs(fd), np.abs(gk).max() * 1e-2), (j, gk[j], fd)
