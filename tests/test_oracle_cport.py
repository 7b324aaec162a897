"""The oracle's C port (timed CPU baseline) computes what the numpy oracle computes
(CPU): identical normals bit for bit, identical trajectories to float32 rounding."""
import numpy as np
import pytest

from oracle import cport
from oracle import jaxrandom as jr
from oracle import model as om
from oracle import steps as ost


@pytest.fixture(scope="module", autouse=True)
def _built():
  try:
    cport.load()
  except ImportError:
    pytest.skip("oracle C port not built (make -C oracle)")


def test_cport_normals_bit_exact():
  for S in (1, 2, 7, 1000, 1001):
    k = np.array([11, 22], np.uint32)
    np.testing.assert_array_equal(cport.normal(k, S).view(np.uint32),
                                  jr.normal(k, (S,)).view(np.uint32))


def _data(N, D, kind, seed=0):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  if kind == "logistic":
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  else:
    y = (X @ w + 0.5 * rng.standard_normal(N)).astype(np.float32)
  return X, y


def test_cport_sghmc_matches_numpy_oracle():
  N, D, n, C = 2000, 12, 64, 3
  X, y = _data(N, D, "logistic")
  eps = np.full(4, 0.01, np.float32)
  theta = np.zeros((C, D), np.float32)
  keys = np.array([[0, c] for c in range(C)], np.uint32)
  dkeys = np.array([[1, c] for c in range(C)], np.uint32)
  cport.sghmc(1, X, y, n, theta, keys.copy(), dkeys.copy(), eps, steps=3, friction=0.5,
              threads=2)
  fam = om.Family("logistic", X, y)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  for c in range(C):
    r = ost.sghmc(pot, np.zeros(D), 4, eps, 3, 0.5, N, n, [D], key=keys[c], dkey=dkeys[c])
    np.testing.assert_allclose(theta[c], r["theta"][-1], rtol=1e-4,
                               atol=1e-5 * np.max(np.abs(r["theta"][-1])))


def test_cport_sgld_matches_numpy_oracle():
  N, D, n = 500, 8, 32
  X, y = _data(N, D, "linear")
  eps = ost.polynomial_step_size_first_last(1e-3, 1e-4, 0.33, 20)
  theta = np.zeros((1, D), np.float32)
  keys = np.array([[0, 0]], np.uint32)
  dkeys = np.array([[0, 0]], np.uint32)
  cport.sgld(0, X, y, n, theta, keys, dkeys, eps, sigma=0.5, prior_scale=np.sqrt(10.0))
  fam = om.Family("linear_gaussian", X, y, sigma=0.5)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, np.sqrt(10.0))
  r = ost.sgld(pot, np.zeros(D), 20, eps, 1.0, N, n, [D])
  np.testing.assert_allclose(theta[0], r["theta"][-1], rtol=1e-4,
                             atol=1e-5 * np.max(np.abs(r["theta"][-1])))
  np.testing.assert_array_equal(keys[0], r["key"])
  np.testing.assert_array_equal(dkeys[0], r["dkey"])


def test_cport_amagold_matches_numpy_oracle():
  """AMAGOLD (reversible leapfrog + full-data MH) in the C port against the numpy
  oracle: same accept/reject decisions, positions within float32 rounding."""
  N, D, n, its, steps, fr = 600, 6, 32, 25, 3, 0.3
  X, y = _data(N, D, "logistic", seed=3)
  fam = om.Family("logistic", X, y)
  eps = np.full(its, 0.05, np.float32)
  r = ost.amagold(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0),
                  lambda th: om.full_potential(fam, th, 1.0, N), np.zeros(D), its, eps, steps,
                  fr, N, n, [D], key=jr.PRNGKey(5), dkey=jr.PRNGKey(6))
  assert r["accept"].any() and (~r["accept"]).any()
  p0, k_int, k_mh = cport.amagold_init(np.zeros(D), jr.PRNGKey(5))
  theta = np.zeros((1, D), np.float32)
  mom = p0.reshape(1, D).copy()
  keys = np.array([k_int], np.uint32)
  mh_keys = np.array([k_mh], np.uint32)
  dkeys = np.array([jr.PRNGKey(6)], np.uint32)
  U = np.array([om.full_potential(fam, np.zeros(D), 1.0, N)])
  acc = cport.amagold(1, X, y, n, theta, mom, keys, mh_keys, dkeys, U, eps, steps, fr)
  np.testing.assert_array_equal(acc[0], r["accept"])
  np.testing.assert_allclose(theta[0], r["theta"][-1], rtol=1e-5, atol=1e-6)
  np.testing.assert_allclose(mom[0], r["p"][-1], rtol=1e-4, atol=1e-5)


def test_cport_sggmc_matches_numpy_oracle():
  """SGGMC (OBABO proposals + full-data MH) in the C port against the numpy oracle:
  same accept/reject decisions, positions and momenta within float32 rounding."""
  N, D, n, its, steps, fr = 600, 6, 32, 25, 3, 1.0
  X, y = _data(N, D, "logistic", seed=3)
  fam = om.Family("logistic", X, y)
  eps = np.full(its, 0.02, np.float32)
  r = ost.sggmc(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0),
                lambda th: om.full_potential(fam, th, 1.0, N), np.zeros(D), its, eps, steps,
                fr, N, n, [D], key=jr.PRNGKey(5), dkey=jr.PRNGKey(6))
  assert r["accept"].any() and (~r["accept"]).any()
  k_int, k_mh = jr.split(jr.PRNGKey(5))
  theta = np.zeros((1, D), np.float32)
  mom = np.zeros((1, D), np.float32)
  keys = np.array([k_int], np.uint32)
  mh_keys = np.array([k_mh], np.uint32)
  dkeys = np.array([jr.PRNGKey(6)], np.uint32)
  U = np.array([om.full_potential(fam, np.zeros(D), 1.0, N)])
  acc = cport.sggmc(1, X, y, n, theta, mom, keys, mh_keys, dkeys, U, eps, steps, fr)
  np.testing.assert_array_equal(acc[0], r["accept"])
  np.testing.assert_allclose(theta[0], r["theta"][-1], rtol=1e-5, atol=1e-6)
  # the C port's per-row likelihoods are float32 (summed in FP64)
  np.testing.assert_allclose(U[0], r["U_full"][-1], rtol=1e-7)
