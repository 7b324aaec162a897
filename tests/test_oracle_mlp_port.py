"""The float32 MLP port (oracle/mlp_port.py, the timed CPU baseline of the MLP bench
lines) against the float64 oracle (CPU): the noise tree and index draws bit-exact
(oracle build of the C port; the vectorized baseline build within float32 rounding),
potential and gradient within float32 rounding, and pSGLD / OBABO iterations and the
full-data potential following the oracle's."""
import numpy as np
import pytest

from oracle import cport
from oracle import jaxrandom as jr
from oracle import model as om
from oracle import mlp_port as mp
from oracle import steps as ost

DIMS = (64, 128, 128)


def _data(N=500, seed=0):
  rng = np.random.default_rng(seed)
  X = rng.random((N, DIMS[0]), dtype=np.float32)
  labels = rng.integers(0, 10, N)
  th = rng.standard_normal(sum(_sizes()))
  return X, labels, (0.1 * th).astype(np.float32)


def _sizes():
  return [int(np.prod(s)) for s in om.mlp_shapes(*DIMS, 10)]


@pytest.mark.parametrize("fast", [False, True])
def test_noise_tree_and_indices(fast):
  key = np.array([3, 17], np.uint32)
  ref = jr.random_tree_flat(key, _sizes())
  got = cport.tree_normals(key, _sizes(), threads=4, fast=fast)
  if fast:
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
  else:
    np.testing.assert_array_equal(got, ref)
  dk = np.array([1, 2], np.uint32)
  dk_ref, idx_ref = ost.draw_indices(dk.copy(), 500, 48)
  idx = cport.draw_indices(dk, 48, 500, fast=fast)
  np.testing.assert_array_equal(idx, idx_ref)
  np.testing.assert_array_equal(dk, dk_ref)


def test_potential_and_gradient():
  X, lab, th = _data()
  port = mp.MLPPort(X, lab, DIMS, 48, prior_scale=1.5, temperature=1.3)
  fam = om.Family("mlp", X, lab, dims=DIMS + (10,))
  idx = np.random.default_rng(1).integers(0, 500, 48)
  U, g = port.potential_grad(th, idx)
  Ur, gr, _ = om.stochastic_potential(fam, th.astype(np.float64), idx, 1.5, 1.3)
  assert abs(U - Ur) <= 1e-5 * abs(Ur)
  assert np.linalg.norm(g - gr) <= 1e-5 * np.linalg.norm(gr)


def test_psgld_iterations_follow_the_oracle():
  X, lab, th = _data()
  fam = om.Family("mlp", X, lab, dims=DIMS + (10,))
  pot = lambda t, idx: om.stochastic_potential(fam, t, idx, 1.0)
  eps = np.array([1e-3, 8e-4, 6e-4], np.float32)
  r = ost.sgld(pot, th.astype(np.float64), 3, eps, 1.0, 500, 48, _sizes(), key=np.array([0, 4]),
               dkey=np.array([0, 9]), rms=(0.9, 1e-5))
  port = mp.MLPPort(X, lab, DIMS, 48, threads=2)
  theta, v = th.copy(), np.ones_like(th)
  key, dkey = np.array([0, 4], np.uint32), np.array([0, 9], np.uint32)
  port.psgld(theta, v, key, dkey, eps)
  np.testing.assert_allclose(theta, r["theta"][-1], rtol=1e-4, atol=1e-6)
  np.testing.assert_array_equal(key, r["key"])
  np.testing.assert_array_equal(dkey, r["dkey"])


def test_obabo_steps_and_full_potential_follow_the_oracle():
  X, lab, th = _data()
  fam = om.Family("mlp", X, lab, dims=DIMS + (10,))
  pot = lambda t, idx: om.stochastic_potential(fam, t, idx, 1.0)
  P = th.size
  ones = np.ones(P)
  key, dkey = np.array([0, 5], np.uint32), np.array([0, 6], np.uint32)
  th_r, p_r, key_r, dkey_r, _, ks_r, ke_r = ost.obabo_integrate(
      pot, th.astype(np.float64), np.zeros(P), key.copy(), dkey.copy(), np.float32(1e-3), 1.0,
      0.7, 2, 500, 48, _sizes(), ones, ones)
  port = mp.MLPPort(X, lab, DIMS, 48, threads=2)
  theta, p = th.copy(), np.zeros_like(th)
  ks, ke = port.obabo_steps(theta, p, key, dkey, 1e-3, 0.7, 2)
  np.testing.assert_allclose(theta, th_r, rtol=1e-4, atol=1e-6)
  np.testing.assert_allclose(p, p_r, rtol=1e-3, atol=1e-5)
  assert abs(ks - ks_r) <= 1e-4 * abs(ks_r) and abs(ke - ke_r) <= 1e-4 * abs(ke_r)
  np.testing.assert_array_equal(key, key_r)
  np.testing.assert_array_equal(dkey, dkey_r)
  full = port.full_potential(theta, 64)
  ref = om.full_potential(fam, theta.astype(np.float64), 1.0, mb=64)
  assert abs(full - ref) <= 1e-5 * abs(ref)
