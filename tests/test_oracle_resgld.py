"""Pins the oracle's reSGLD swap step (solver.py:264-293) by what the arithmetic fixes:
equal temperatures never exchange; an overwhelming potential gap decides every test in
the reference's (inverted) direction; the ssq correction and F enter with the
published sign and scale, checked through the keep frequency exp(log_s) over many
keys; the two-chain composition reduces to two independent SGLD runs when nothing
is exchanged."""
import numpy as np

from oracle import jaxrandom as jr
from oracle import model as om
from oracle import steps as ost


def _keep_fraction(dU, ssq, F, T=(1.0, 2.0), trials=2000):
  key = jr.PRNGKey(11)
  keeps = 0
  for _ in range(trials):
    sw, key = ost.resgld_swap_decision(dU, 0.0, T[0], T[1], ssq, F, key)
    keeps += not sw
  return keeps / trials


def test_equal_temperatures_never_exchange():
  key = jr.PRNGKey(3)
  for dU in (-1e3, -1.0, 0.0, 5.0):
    for _ in range(50):
      sw, key = ost.resgld_swap_decision(dU, 0.0, 1.0, 1.0, 0.7, 1.0, key)
      assert not sw


def test_large_gap_decides_in_reference_direction():
  key = jr.PRNGKey(4)
  for _ in range(50):
    sw, key = ost.resgld_swap_decision(1e30, 0.0, 1.0, 1000.0, 0.0, 1.0, key)
    assert not sw  # log s = +inf side: kept (log u < log s)
    sw, key = ost.resgld_swap_decision(-1e30, 0.0, 1.0, 1000.0, 0.0, 1.0, key)
    assert sw      # log s = -inf side: exchanged


def test_correction_sign_and_scale():
  # temps = 1 - 1/2 = 0.5; log s = 0.5 (dU - 0.5 ssq / F); keep iff u < exp(log s)
  f = _keep_fraction(1.0, 4.0, 1.0)        # log s = -0.5
  assert abs(f - np.exp(-0.5)) < 0.04, f
  f = _keep_fraction(1.0, 4.0, 2.0)        # log s = 0: always kept
  assert f == 1.0
  f = _keep_fraction(3.0, 0.0, 1.0)        # log s = 1.5 > 0: always kept
  assert f == 1.0
  f = _keep_fraction(-2.0, 0.0, 1.0)       # log s = -1
  assert abs(f - np.exp(-1.0)) < 0.04, f


def test_two_chain_composition_without_exchange():
  rng = np.random.default_rng(0)
  N, D, n = 200, 3, 16
  X = rng.standard_normal((N, D)).astype(np.float32)
  y = (X @ np.ones(D) + 0.1 * rng.standard_normal(N)).astype(np.float32)
  fam = om.Family("linear_gaussian", X, y, sigma=0.5)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  th0 = np.zeros(D, np.float32)
  th1 = np.ones(D, np.float32)
  eps = np.full(6, 1e-3, np.float32)
  # equal slot temperatures: no exchange, so each slot is a plain SGLD chain
  r = ost.resgld(pot, th0, th1, 6, eps, 1.0, N, n, [D])
  assert not r["swaps"].any()
  ks = jr.split(jr.PRNGKey(0), 3)
  a = ost.sgld(pot, th0, 6, eps, 1.0, N, n, [D], key=ks[1])
  b = ost.sgld(pot, th1, 6, eps, 1.0, N, n, [D], key=ks[2])
  np.testing.assert_array_equal(r["normal"]["theta"], a["theta"][-1])
  np.testing.assert_array_equal(r["hot"]["theta"], b["theta"][-1])
  np.testing.assert_array_equal(r["samples"][-1], a["theta"][-1])
