"""Pins the oracle's AMAGOLD restatement (integrator.py:349-560, solver.py:301-432) and
its dual-averaging step size (scheduler.py:376-444) by what the mathematics fixes:

- without friction the accumulated energy telescopes to KE(p_0) - KE(p_T) for any
  gradient, stochastic ones included (pins the 0.5 eps scale, p + p', M^-1 g and the
  sign);
- without friction and with a deterministic gradient the integrator is reversible
  (pins the leading and trailing half position steps and the skipped step 0);
- the friction/noise map p -> ((1 - eps f) p + sqrt(4 eps f) M^{1/2} xi)/(1 + eps f)
  leaves N(0, M) invariant (pins the three coefficients together);
- on a quadratic target the energy error is second order in eps;
- the MH-corrected chain samples the analytic linear-Gaussian posterior;
- the dual averaging has mu = log(10 eps0) as its fixed point when the acceptance
  equals the target and freezes after burn-in.
"""
import numpy as np

from oracle import jaxrandom as jr
from oracle import model as om
from oracle import posterior as op
from oracle import steps as ost


def _linreg(N=60, D=3, seed=0):
  rng = np.random.default_rng(seed)
  X = rng.standard_normal((N, D)).astype(np.float32)
  y = (X @ np.array([0.5, -1.0, 2.0][:D]) + 0.5 * rng.standard_normal(N)).astype(np.float32)
  return X, y, om.Family("linear_gaussian", X, y, sigma=0.5)


def _ke(p, inv_mass):
  return 0.5 * np.dot(p, inv_mass * p)


def test_energy_telescopes_without_friction():
  X, y, fam = _linreg()
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  rng = np.random.default_rng(1)
  inv_m = rng.uniform(0.5, 2.0, 3)
  p0 = rng.standard_normal(3)
  th, pT, _, _, energy, _ = ost.reversible_leapfrog_integrate(
      pot, np.zeros(3), p0, jr.PRNGKey(2), jr.PRNGKey(3), 1e-3, 7, 0.0, 60, 8, [3],
      inv_mass=inv_m, sqrt_mass=1.0 / np.sqrt(inv_m))
  assert not np.allclose(th, 0.0)
  np.testing.assert_allclose(energy, _ke(p0, inv_m) - _ke(pT, inv_m), rtol=1e-9, atol=1e-12)


def test_reversible_without_friction():
  A = np.array([[2.0, 0.3, 0.0], [0.3, 1.0, -0.2], [0.0, -0.2, 0.5]])
  grad = lambda th: A @ th + np.sin(th)
  th0 = np.array([0.4, -1.2, 0.7])
  p0 = np.array([1.0, 0.2, -0.5])
  inv_m = np.array([1.0, 0.5, 2.0])
  args = dict(pot=None, eps=0.05, steps=9, friction=0.0, N=1, n=1, leaf_sizes=[3],
              inv_mass=inv_m, grad_fn=grad)
  thT, pT, k, _, _, _ = ost.reversible_leapfrog_integrate(theta=th0, p=p0, key=jr.PRNGKey(0),
                                                          dkey=jr.PRNGKey(0), **args)
  thB, pB, _, _, _, _ = ost.reversible_leapfrog_integrate(theta=thT, p=-pT, key=k,
                                                          dkey=jr.PRNGKey(0), **args)
  assert np.max(np.abs(thT - th0)) > 0.1
  np.testing.assert_allclose(thB, th0, atol=1e-12)
  np.testing.assert_allclose(pB, -p0, atol=1e-12)


def test_friction_map_leaves_gaussian_invariant():
  zero = lambda th: np.zeros_like(th)
  rng = np.random.default_rng(4)
  sqrt_m = rng.uniform(0.5, 2.0, 6)
  for eps, fr in ((0.05, 0.3), (0.01, 0.25), (0.2, 1.0)):
    key = jr.PRNGKey(7)
    run = lambda p: ost.reversible_leapfrog_integrate(
        None, np.zeros(6), p, key, jr.PRNGKey(0), eps, 1, fr, 1, 1, [6],
        inv_mass=1.0 / sqrt_m ** 2, sqrt_mass=sqrt_m, grad_fn=zero)
    x = rng.standard_normal(6)
    p_zero = run(np.zeros(6))[1]
    p_x = run(x)[1]
    a = (p_x - p_zero) / x
    _, sub = jr.split(key)
    xi = jr.random_tree_flat(sub, [6]).astype(np.float64)
    b = p_zero / (sqrt_m * xi)
    np.testing.assert_allclose(a * a + b * b, 1.0, rtol=2e-6)
    assert np.all(a < 1.0) and np.all(b > 0.0)


def test_energy_error_is_second_order():
  A = np.diag([1.0, 4.0])
  U = lambda th: 0.5 * th @ A @ th
  th0 = np.array([1.0, 0.5])
  p0 = np.array([0.3, -0.8])
  errs = []
  for eps, steps in ((0.08, 10), (0.04, 20), (0.02, 40)):
    thT, pT, _, _, energy, _ = ost.reversible_leapfrog_integrate(
        None, th0, p0, jr.PRNGKey(0), jr.PRNGKey(0), eps, steps, 0.0, 1, 1, [2],
        grad_fn=lambda th: A @ th)
    log_alpha = U(th0) - U(thT) + energy
    errs.append(abs(log_alpha))
  assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0, errs


def test_amagold_samples_linreg_posterior():
  N, D = 40, 2
  rng = np.random.default_rng(5)
  X = rng.standard_normal((N, D)).astype(np.float32)
  y = (X @ np.array([1.0, -0.5]) + 0.5 * rng.standard_normal(N)).astype(np.float32)
  fam = om.Family("linear_gaussian", X, y, sigma=0.5)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, np.sqrt(10.0))
  full = lambda th: om.full_potential(fam, th, np.sqrt(10.0), N)
  mu, lam = op.linreg_posterior(X, y, 0.5, 10.0)
  sd = np.sqrt(np.diag(np.linalg.inv(lam)))
  r = ost.amagold(pot, full, np.zeros(D), 2000, 0.01, 5, 0.5, N, 10, [D])
  s = r["theta"][200:]
  acc = np.mean(r["accept"])
  assert 0.3 < acc < 1.0, acc
  err = np.abs(s.mean(0) - mu) / sd
  assert np.all(err < 4.0 / np.sqrt(len(s) / 30)), (s.mean(0), mu, sd)
  np.testing.assert_allclose(s.std(0), sd, rtol=0.25)


def test_amagold_rejection_flips_momentum_and_keeps_theta():
  X, y, fam = _linreg()
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  full = lambda th: om.full_potential(fam, th, 1.0, 60)
  r = ost.amagold(pot, full, np.zeros(3), 30, 0.05, 5, 0.25, 60, 8, [3])
  acc = r["accept"]
  assert acc.any() and (~acc).any()
  for t in range(1, 30):
    if not acc[t]:
      np.testing.assert_array_equal(r["theta"][t], r["theta"][t - 1])
      assert r["U_full"][t] == r["U_full"][t - 1]
      np.testing.assert_array_equal(r["p"][t], -r["p"][t - 1])


def test_nan_log_ratio_follows_each_solver():
  """A NaN full potential at the proposal: sggmc's jnp.where(log_alpha <= 0, log_alpha,
  0.0) (solver.py:526) turns it into log a = 0 and accepts; amagold's
  jnp.where(log_alpha > 0.0, 0.0, log_alpha) (solver.py:386) keeps NaN and rejects."""
  X, y, fam = _linreg()
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  calls = []

  def full(th):
    calls.append(1)
    return om.full_potential(fam, th, 1.0, 60) if len(calls) == 1 else float("nan")

  r = ost.sggmc(pot, full, np.zeros(3), 4, np.full(4, 0.01, np.float32), 2, 1.0, 60, 8, [3])
  assert r["accept"].all() and np.all(r["log_alpha"] == 0.0)
  calls.clear()
  r = ost.amagold(pot, full, np.zeros(3), 4, 0.01, 2, 0.25, 60, 8, [3])
  assert not r["accept"].any() and np.all(np.isnan(r["log_alpha"]))


def test_dual_averaging_fixed_point_and_burn_in():
  eps0 = 0.05
  steps = ost.adaptive_step_size(np.full(200, 0.25, np.float32), 100, eps0, 10, 0.75, 0.05,
                                 0.25)
  np.testing.assert_allclose(steps[0], eps0, rtol=1e-6)
  np.testing.assert_allclose(steps[100], 10 * eps0, rtol=2e-2)
  assert np.all(steps[101:] == steps[101])
  lo = ost.adaptive_step_size(np.zeros(50, np.float32), 50, eps0, 10, 0.75, 0.05, 0.25)
  hi = ost.adaptive_step_size(np.ones(50, np.float32), 50, eps0, 10, 0.75, 0.05, 0.25)
  assert lo[-1] < eps0 < hi[-1]
