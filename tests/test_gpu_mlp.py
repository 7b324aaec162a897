"""MLP likelihood on MFMA (bf16 operands, fp32 accumulation) against the float64 oracle
(GPU). The bf16 path is checked to its precision: potential within 1e-2 relative and
the gradient within 3e-2 relative in norm of the float64 oracle on the same
minibatch; the update reproduces U, the gradient and the reference's update rule with
the exact noise words; the autograd wrapper returns the kernel gradient."""
import numpy as np
import pytest
import torch

from oracle import jaxrandom as jr
from oracle import model as om
from oracle import steps as ost

from jax_sgmc_amd import mlp as M

pytestmark = pytest.mark.gpu
DEV = "cuda"
DIMS = (128, 128, 128)


def data(N=1000, seed=0):
  rng = np.random.default_rng(seed)
  X = rng.random((N, DIMS[0]), dtype=np.float32)
  teacher = M.init_params(1, *DIMS)
  fam = om.Family("mlp", X, np.zeros(N, np.int64), dims=DIMS + (10,))
  W1, W2, W3, b1, b2, b3 = om.mlp_unflatten(teacher, DIMS + (10,))
  h = np.maximum(np.maximum(X @ W1 + b1, 0) @ W2 + b2, 0)
  labels = np.argmax(h @ W3 + b3, axis=1).astype(np.int32)
  return X, labels


def engine(X, labels, C=2, n=64, eps=1e-5, keys=None, dkeys=None, rms=False):
  th = np.stack([M.init_params(10 + c, *DIMS) for c in range(C)])
  eng = M.FusedMLP(torch.as_tensor(X, device=DEV), torch.as_tensor(labels, device=DEV), DIMS,
                   n, th, np.full(8, eps, np.float32), keys=keys, dkeys=dkeys, rms=rms,
                   prior_scale=1.0)
  return eng, th


def oracle_potential(X, labels, theta, dkey, n):
  fam = om.Family("mlp", X, labels, dims=DIMS + (10,))
  _, idx = ost.draw_indices(np.asarray(dkey, np.uint32), X.shape[0], n)
  return om.stochastic_potential(fam, theta.astype(np.float64), idx, 1.0)


def test_mlp_gradient_against_oracle():
  print("mlp gradient", flush=True)
  X, lab = data()
  keys = [(0, 1), (0, 2)]
  dkeys = [(3, 4), (5, 6)]
  eng, th = engine(X, lab, keys=keys, dkeys=dkeys)
  U, g = eng.potential_and_grad()
  torch.cuda.synchronize()
  for c in range(2):
    Ur, gr, _ = oracle_potential(X, lab, th[c], dkeys[c], 64)
    assert abs(U[c].item() - Ur) <= 1e-2 * abs(Ur), (U[c].item(), Ur)
    gk = g[c].cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(gk - gr) / np.linalg.norm(gr)
    assert rel < 3e-2, rel
  # gradient mode leaves keys and the iteration counter alone
  assert eng.iter.item() == 0
  np.testing.assert_array_equal(eng.keys.cpu().numpy().view(np.uint32), np.asarray(keys, np.uint32))


@pytest.mark.parametrize("rms", [False, True])
def test_mlp_update_follows_reference_rule(rms):
  print("mlp update rms=", rms, flush=True)
  X, lab = data()
  keys = [(0, 7), (0, 8)]
  dkeys = [(1, 1), (1, 2)]
  eps = 1e-5
  eng, th = engine(X, lab, eps=eps, keys=keys, dkeys=dkeys, rms=rms)
  _, g = eng.potential_and_grad()
  g = g.cpu().numpy()
  eng.iteration()
  torch.cuda.synchronize()
  th1 = eng.theta.cpu().numpy()
  sizes = [128 * 128, 128 * 128, 1280, 128, 128, 10]
  f = np.float32
  for c in range(2):
    nk, sub = jr.split(np.asarray(keys[c], np.uint32))
    xi = jr.random_tree_flat(sub, sizes)
    sg = f(-eps) * g[c]
    sn = np.sqrt(f(f(2) * f(1)) * f(eps)) * xi
    if rms:
      v = f(0.9) * np.ones_like(g[c]) + (f(1) - f(0.9)) * (g[c] * g[c])
      G = f(1) / (f(1e-5) + np.sqrt(v))
      ref = th[c] + ((f(0) + G * sg) + np.sqrt(G) * sn)
    else:
      ref = th[c] + (sg + sn)
    np.testing.assert_allclose(th1[c], ref, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(eng.keys[c].cpu().numpy().view(np.uint32), nk)
  assert eng.iter.item() == 1


def test_mlp_autograd_function():
  X, lab = data()
  eng, th = engine(X, lab, dkeys=[(2, 2), (2, 3)])
  theta = torch.as_tensor(th, device=DEV).requires_grad_(True)
  U = M.MLPPotential.apply(theta, eng)
  (U * torch.tensor([1.0, 2.0], device=DEV)).sum().backward()
  _, g = eng.potential_and_grad()
  torch.testing.assert_close(theta.grad[0], g[0])
  torch.testing.assert_close(theta.grad[1], 2 * g[1])


def test_mlp_graph_replay_matches_eager():
  X, lab = data()
  a, _ = engine(X, lab, keys=[(0, 1), (0, 2)], dkeys=[(1, 1), (1, 2)])
  b, _ = engine(X, lab, keys=[(0, 1), (0, 2)], dkeys=[(1, 1), (1, 2)])
  a.run(6, graph_iterations=3)
  b.run(6)
  torch.cuda.synchronize()
  assert torch.equal(a.theta, b.theta)
  assert a.iter.item() == 6
