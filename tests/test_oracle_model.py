"""Pins the oracle's potentials, gradients, schedules, integrators and posterior
targets (CPU): closed forms, finite differences, invariants, and the reference's own
unit tests restated on the oracle."""
import itertools

import numpy as np
import pytest

from oracle import jaxrandom as jr
from oracle import model as om
from oracle import posterior as op
from oracle import steps as ost


def _fd_grad(f, x, h=1e-6):
  g = np.zeros_like(x)
  for i in range(x.size):
    e = np.zeros_like(x)
    e[i] = h
    g[i] = (f(x + e) - f(x - e)) / (2 * h)
  return g


@pytest.mark.parametrize("kind", ["linear_gaussian", "logistic"])
def test_glm_gradient_finite_differences(kind):
  rng = np.random.default_rng(0)
  N, D, n = 50, 6, 9
  X = rng.standard_normal((N, D))
  y = (rng.random(N) < 0.5).astype(float) if kind == "logistic" else rng.standard_normal(N)
  fam = om.Family(kind, X, y, sigma=0.7)
  idx = rng.integers(0, N, n)
  w = rng.standard_normal(D)
  U, g, _ = om.stochastic_potential(fam, w, idx, prior_scale=2.0, temperature=1.3)
  fd = _fd_grad(lambda v: om.stochastic_potential(fam, v, idx, 2.0, 1.3)[0], w)
  np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-6)


def test_glm_closed_form_gradients():
  """linear-Gaussian: -(N/n) X^T (y - Xw)/s^2 + w/tau2; logistic: -(N/n) X^T (y - s(Xw)) + w."""
  rng = np.random.default_rng(1)
  N, D, n = 40, 5, 40
  X = rng.standard_normal((N, D))
  y = rng.standard_normal(N)
  w = rng.standard_normal(D)
  idx = np.arange(N)
  _, g, _ = om.stochastic_potential(om.Family("linear_gaussian", X, y, sigma=0.5), w, idx,
                                    np.sqrt(10.0))
  np.testing.assert_allclose(g, -X.T @ (y - X @ w) / 0.25 + w / 10.0, rtol=1e-12)
  yb = (rng.random(N) < 0.5).astype(float)
  _, g, _ = om.stochastic_potential(om.Family("logistic", X, yb), w, idx, 1.0)
  np.testing.assert_allclose(g, -X.T @ (yb - 1 / (1 + np.exp(-X @ w))) + w, rtol=1e-12)


def test_mlp_gradient_finite_differences():
  rng = np.random.default_rng(2)
  dims = (7, 5, 4, 3)
  P = sum(int(np.prod(s)) for s in om.mlp_shapes(*dims))
  X = rng.random((20, 7))
  lab = rng.integers(0, 3, 20)
  fam = om.Family("mlp", X, lab, dims=dims)
  theta = 0.5 * rng.standard_normal(P)
  idx = rng.integers(0, 20, 6)
  _, g, _ = om.stochastic_potential(fam, theta, idx, 1.0)
  fd = _fd_grad(lambda v: om.stochastic_potential(fam, v, idx, 1.0)[0], theta)
  np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("obs,dim", itertools.product([7, 11], [3, 5]))
def test_reference_linear_potential_closed_form(obs, dim):
  """tests/test_potential.py:65-92: lik = sum(sample * obs), prior = sum(sample),
  observations tile(arange(4)) over `dim` rows, N = obs: U = -sum(arange(4)) obs - 4."""
  observations = np.tile(np.arange(4), (dim, 1)).astype(np.float64)
  sample = np.ones(4)
  lik = observations @ sample
  U = -obs * np.mean(lik) - np.sum(sample)
  assert np.isclose(U, -np.sum(np.arange(4)) * obs - 4)


@pytest.mark.parametrize("N,mb,kind", [(7, 2, "logistic"), (11, 3, "linear_gaussian"),
                                       (11, 2, "logistic"), (100, 3, "linear_gaussian")])
def test_full_potential_equals_minibatch_over_all_rows(N, mb, kind):
  """tests/test_potential.py:298-333: the masked, wrapped full-data batching equals
  the minibatch potential evaluated on every row once."""
  rng = np.random.default_rng(3)
  X = rng.standard_normal((N, 4))
  y = (rng.random(N) < 0.5).astype(float) if kind == "logistic" else rng.standard_normal(N)
  fam = om.Family(kind, X, y, sigma=1.0)
  w = rng.standard_normal(4)
  full = om.full_potential(fam, w, 1.5, mb)
  U, _, _ = om.stochastic_potential(fam, w, np.arange(N), 1.5)
  assert np.isclose(full, U, rtol=1e-12)
  assert np.isclose(om.full_potential(fam, w, 1.5, mb, offset=N - 1), U, rtol=1e-12)


@pytest.mark.parametrize("first,last,gamma,its", itertools.product(
    [1.0, 0.05], [0.01, 0.0009], [0.33, 0.55], [100, 14723]))
def test_polynomial_first_last(first, last, gamma, its):
  """tests/test_scheduler.py:83-102."""
  s = ost.polynomial_step_size_first_last(first, last, gamma, its)
  np.testing.assert_allclose(s[0], first, rtol=1e-5)
  np.testing.assert_allclose(s[-1], last, rtol=1e-5)
  assert np.all(np.diff(s) <= 0)


@pytest.mark.parametrize("n", [123, 243])
def test_initial_burn_in(n):
  """tests/test_scheduler.py:105-121."""
  b = ost.initial_burn_in(n, 1000)
  assert np.all(b[:n] == 0.0) and np.all(b[n:] == 1.0)


def _quad_family(N, D, sigma=1.0):
  """All rows identical: every minibatch potential equals the full potential."""
  X = np.tile(np.linspace(-1, 1, D), (N, 1))
  y = np.full(N, 0.3)
  return om.Family("linear_gaussian", X, y, sigma=sigma)


def test_sgld_zero_step_is_identity_and_one_step_closed_form():
  fam = _quad_family(10, 3)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  th0 = np.array([0.1, -0.2, 0.3])
  r = ost.sgld(pot, th0, 3, 0.0, 1.0, 10, 4, [3])
  np.testing.assert_array_equal(r["theta"][-1], th0)
  eps = np.float32(0.01)
  r = ost.sgld(pot, th0, 1, eps, 1.0, 10, 4, [3])
  key, sub = jr.split(jr.PRNGKey(0))
  xi = jr.normal(jr.split(sub, 1)[0], (3,)).astype(np.float64)
  _, g, _ = pot(th0, np.zeros(4, int))
  exp = th0 + (float(-eps) * g + float(np.sqrt(np.float32(2) * eps)) * xi)
  np.testing.assert_allclose(r["theta"][0], exp, rtol=1e-14)
  np.testing.assert_array_equal(r["noise"][0], xi)


def test_obabo_without_friction_is_velocity_verlet_and_conserves_energy():
  """friction 0: permanence 1, no noise; with exact gradients the leapfrog conserves
  H = U + KE to O(eps^2)."""
  fam = _quad_family(20, 4)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  th = np.array([0.5, -0.4, 0.3, 0.2])
  p = np.array([0.3, 0.1, -0.2, 0.05])
  U0 = pot(th, np.zeros(1, int))[0]
  H0 = U0 + 0.5 * p @ p
  eps = 1e-3
  th1, p1, *_ = ost.obabo_integrate(pot, th, p, jr.PRNGKey(0), jr.PRNGKey(0), eps, 1.0, 0.0, 50,
                                    20, 5, [4], np.ones(4), np.ones(4))
  H1 = pot(th1, np.zeros(1, int))[0] + 0.5 * p1 @ p1
  assert abs(H1 - H0) < 1e-4 * abs(H0)


def test_sghmc_small_friction_no_noise_limit():
  """With the noise scale forced to zero (eps -> the friction term vanishes) the
  leapfrog body reduces to symplectic Euler: p -= eps g(theta + eps p)."""
  fam = _quad_family(10, 2)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  r = ost.sghmc(pot, np.zeros(2), 1, 1e-3, 3, 0.0, 10, 2, [2])
  assert np.all(np.isfinite(r["theta"]))


def test_sgld_stationary_covariance_matches_simulation():
  """The Lyapunov target (oracle/posterior.py) against a long oracle chain on a tiny
  problem; and S -> Lambda^-1 as eps -> 0."""
  rng = np.random.default_rng(4)
  N, D, n = 60, 2, 6
  X = rng.standard_normal((N, D))
  y = X @ np.array([0.5, -1.0]) + 0.5 * rng.standard_normal(N)
  mu, Lam = op.linreg_posterior(X, y, 0.5, 10.0)
  _, S_small = op.sgld_stationary(X, y, 0.5, 10.0, 1e-7, n)
  np.testing.assert_allclose(S_small, np.linalg.inv(Lam), rtol=1e-3)
  eps = 2e-3
  mu2, S = op.sgld_stationary(X, y, 0.5, 10.0, eps, n)
  np.testing.assert_allclose(mu2, mu)
  # direct simulation in float64 with numpy randomness (fast, not jax keys)
  H = Lam
  b = X.T @ y / 0.25
  th = mu.copy()
  draws = []
  for t in range(60000):
    idx = rng.integers(0, N, n)
    g = -(N / n) * X[idx].T @ (y[idx] - X[idx] @ th) / 0.25 + th / 10.0
    th = th - eps * g + np.sqrt(2 * eps) * rng.standard_normal(D)
    if t > 1000:
      draws.append(th)
  draws = np.array(draws)
  emp = np.cov(draws.T)
  np.testing.assert_allclose(emp, S, rtol=0.1, atol=0.1 * np.max(np.abs(S)))
  assert not np.allclose(S, np.linalg.inv(H), rtol=0.02)  # finite-eps effect is visible
