"""SGHMC noise model and mass adaption (adaption.py:296-457, integrator.py:626-650):
the oracle's Fisher noise pinned by a hand-computed correction and by finite
differences of each observation's likelihood; adaption.fisher_information and
adaption.mass_matrix (torch, run here on the CPU) against the oracle."""
import numpy as np
import torch

from oracle import model as om
from oracle import steps as ost

import __graft_entry__

__graft_entry__.load()
from jax_sgmc_amd import adaption as AD  # noqa: E402
from jax_sgmc_amd import data as DT  # noqa: E402
from jax_sgmc_amd import likelihoods as LK  # noqa: E402
from jax_sgmc_amd import potential as PT  # noqa: E402


def test_fisher_correction_by_hand():
  # v = ssq / (n - 1) with n = 3 and diff rows chosen so that b = 0.5 eps v is
  # (0.5, 2.0, 0.9) at eps = 1: friction 1 gives corrections (0.5, -1.0, 0.1); the
  # non-positive one is replaced by the smallest positive, 0.1.
  v = np.array([1.0, 4.0, 1.8])
  ssq = v * 2
  diff = np.zeros((3, 3))
  diff[0] = np.sqrt(ssq)
  cb, b = ost.fisher_noise(diff, np.zeros(3), 1, 3, 1.0, 1.0)
  np.testing.assert_allclose(cb ** 2, [0.5, 0.1, 0.1], rtol=1e-12)
  np.testing.assert_allclose(b ** 2, [0.5, 0.9, 0.9], rtol=1e-12)


def test_observation_gradients_against_finite_differences():
  rng = np.random.default_rng(0)
  X = rng.standard_normal((5, 4))
  y = (rng.random(5) < 0.5).astype(float)
  w = rng.standard_normal(4)
  for fn in (lambda w_: om.logistic(w_, X, y), lambda w_: om.linear_gaussian(w_, X, y, 0.7)):
    _, G = fn(w)
    h = 1e-6
    for j in range(4):
      e = np.zeros(4)
      e[j] = h
      fd = (fn(w + e)[0] - fn(w - e)[0]) / (2 * h)
      np.testing.assert_allclose(G[:, j], fd, rtol=1e-6, atol=1e-8)


def _batch(N=200, n=16, D=5, seed=1):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  y = (rng.random(N) < 0.5).astype(np.float32)
  idx = rng.integers(0, N, n)
  batch = {"x": torch.as_tensor(X[idx]), "y": torch.as_tensor(y[idx])}
  info = DT.MiniBatchInformation(observation_count=N, batch_size=n,
                                 mask=torch.ones(n, dtype=torch.bool))
  return X, y, idx, (batch, info)


def test_fisher_information_matches_oracle_on_cpu():
  X, y, idx, mb = _batch()
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  w = np.random.default_rng(2).standard_normal(5).astype(np.float32) * 0.3
  fam = om.Family("logistic", X, y)
  _, g, _ = om.stochastic_potential(fam, w.astype(np.float64), idx, 1.0)
  nm = AD.fisher_information(minibatch_potential=pot)
  st = nm.init(torch.zeros(5))
  for eps in (1e-3, 0.05):
    out = nm.get(st, torch.as_tensor(w), torch.as_tensor(g, dtype=torch.float32),
                 torch.full((5,), 0.8), mini_batch=mb, step_size=eps)
    _, G = om.logistic(w, X[idx], y[idx])
    cb, bs = ost.fisher_noise(G, g, 200, 16, 0.8, eps)
    np.testing.assert_allclose(out.cb_diff_sqrt.tensor.numpy(), cb, rtol=1e-5)
    # b_sqrt = sqrt(friction - (friction - b)) cancels in float32 (as in the reference):
    # absolute error ~ ulp(friction) / (2 sqrt(b))
    np.testing.assert_allclose(out.b_sqrt.tensor.numpy(), bs, rtol=1e-4, atol=2e-5)


def test_mass_matrix_matches_welford_closed_form():
  rng = np.random.default_rng(3)
  S = rng.standard_normal((12, 4)) * np.array([0.5, 1.0, 2.0, 3.0])
  mm = AD.mass_matrix(diagonal=True, burn_in=8)
  st = mm.init(torch.zeros(4))
  ref = ost.mass_adaption(S, 8)
  for t in range(12):
    st = mm.update(st, torch.as_tensor(S[t], dtype=torch.float64))
    m = mm.get(st)
    np.testing.assert_allclose(m.inv.tensor.numpy(), ref[t][0], rtol=1e-10)
    np.testing.assert_allclose(m.sqrt.tensor.numpy(), ref[t][1], rtol=1e-10)


def _signs(v, M):
  """Coefficients of v in the eigenbasis of the symmetric matrix M, each divided by the
  square root of its eigenvalue: +-1 each when v = U sqrt(s) for an SVD U s V^T of M."""
  lam, E = np.linalg.eigh(M)
  return (E.T @ v) / np.sqrt(lam)


def test_full_noise_model_up_to_singular_vector_signs():
  """fisher_information(diagonal=False) (adaption.py:441-452) returns U sqrt(s) of the SVDs
  of diag(friction) - b and of b: a vector whose value depends on the sign of each
  singular vector, which LAPACK (and the GPU solvers) leave unspecified. What is fixed
  is that every coefficient of that vector in the eigenbasis of the (symmetric,
  positive semi-definite) matrix is +-sqrt(s_j): checked for the oracle and for the
  library (torch, here on the CPU). A square root left out, or U transposed, fails."""
  X, y, idx, mb = _batch(D=5, n=16)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  w = np.random.default_rng(4).standard_normal(5).astype(np.float32) * 0.3
  fam = om.Family("logistic", X, y)
  _, g, _ = om.stochastic_potential(fam, w.astype(np.float64), idx, 1.0)
  _, G = om.logistic(w, X[idx], y[idx])
  eps, fric = 0.05, 0.8
  diff = G - g / 200
  b = 0.5 * np.float32(eps) * (diff.T @ diff / 15)
  Mcb = np.diag(np.full(5, fric)) - b
  assert np.linalg.eigvalsh(b).min() > 1e-9 and np.linalg.eigvalsh(Mcb).min() > 1e-3
  cb, bs = ost.fisher_noise(G, g, 200, 16, fric, eps, diagonal=False)
  np.testing.assert_allclose(np.abs(_signs(cb, Mcb)), 1.0, rtol=1e-8)
  np.testing.assert_allclose(np.abs(_signs(bs, b)), 1.0, rtol=1e-8)
  nm = AD.fisher_information(minibatch_potential=pot, diagonal=False)
  out = nm.get(nm.init(torch.zeros(5)), torch.as_tensor(w), torch.as_tensor(g, dtype=torch.float32),
               torch.full((5,), fric), mini_batch=mb, step_size=eps)
  np.testing.assert_allclose(np.abs(_signs(out.cb_diff_sqrt.tensor.double().numpy(), Mcb)), 1.0,
                             rtol=2e-3)
  np.testing.assert_allclose(np.abs(_signs(out.b_sqrt.tensor.double().numpy(), b)), 1.0,
                             rtol=2e-3)
  # planted mistakes: singular values without the square root, U transposed
  lam, E = np.linalg.svd(b)[1], np.linalg.svd(b)[0]
  assert not np.allclose(np.abs(_signs(E @ lam, b)), 1.0, rtol=1e-2)
  assert not np.allclose(np.abs(_signs(E.T @ np.sqrt(lam), b)), 1.0, rtol=1e-2)
