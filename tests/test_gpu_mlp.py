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


def _obabo_engine(X, lab, th, integrator, keys, dkeys, eps, steps, friction=1.0, slots=None):
  return M.FusedMLP(torch.as_tensor(X, device=DEV), torch.as_tensor(lab, device=DEV), DIMS, 64,
                    th, eps, keys=keys, dkeys=dkeys, prior_scale=1.0, integrator=integrator,
                    friction=friction, steps=steps, save_slots=slots)


def test_mlp_obabo_matches_oracle():
  """One iteration of 3 OBABO steps (integrator.py:203-273) for 2 chains: the noise
  words are exact, so momentum and position differ from the float64 oracle only
  through the bf16 gradients: the difference stays within 5e-2 of the gradient's
  whole contribution (oracle with and without gradient); the kinetic energies within
  1e-3 and the step potential within 1e-2."""
  print("mlp obabo", flush=True)
  X, lab = data()
  rng = np.random.default_rng(3)
  th = np.stack([M.init_params(20 + c, *DIMS) for c in range(2)])
  keys, dkeys = [(0, 11), (0, 12)], [(4, 1), (4, 2)]
  eps = np.full(4, 2e-3, np.float32)
  eng = _obabo_engine(X, lab, th, "obabo", keys, dkeys, eps, steps=3, friction=0.7)
  eng.iteration()
  torch.cuda.synchronize()
  fam = om.Family("mlp", X, lab, dims=DIMS + (10,))
  pot = lambda t, idx: om.stochastic_potential(fam, t, idx, 1.0)
  zero = lambda t, idx: (0.0, np.zeros_like(t), np.zeros(len(idx)))
  P = th.shape[1]
  ones = np.ones(P)
  for c in range(2):
    args = (np.asarray(keys[c], np.uint32), np.asarray(dkeys[c], np.uint32), eps[0], 1.0,
            0.7, 3, X.shape[0], 64, [128 * 128, 128 * 128, 1280, 128, 128, 10], ones, ones)
    th_r, p_r, key_r, dkey_r, U_r, ks_r, ke_r = ost.obabo_integrate(
        pot, th[c].astype(np.float64), np.zeros(P), *args)
    th_0, p_0, *_ = ost.obabo_integrate(zero, th[c].astype(np.float64), np.zeros(P), *args)
    p_g = eng.mom[c].cpu().numpy().astype(np.float64)
    t_g = eng.theta[c].cpu().numpy().astype(np.float64)
    assert np.linalg.norm(p_g - p_r) <= 5e-2 * np.linalg.norm(p_r - p_0) + 1e-6
    assert np.linalg.norm(t_g - th_r) <= 5e-2 * np.linalg.norm(th_r - th_0) + 1e-6
    ks, ke = eng.ke[c].tolist()
    assert abs(ks - ks_r) <= 1e-3 * abs(ks_r) and abs(ke - ke_r) <= 1e-3 * abs(ke_r), (ks, ks_r)
    assert abs(eng.U[c].item() - U_r) <= 1e-2 * abs(U_r)
    np.testing.assert_array_equal(eng.keys[c].cpu().numpy().view(np.uint32), key_r)
    np.testing.assert_array_equal(eng.dkeys[c].cpu().numpy().view(np.uint32), dkey_r)


def test_mlp_full_potential_f64():
  """float64 full-data potential of the bf16 forward pass within 1e-2 of the float64
  oracle's (every row once, N not a multiple of the pass size)."""
  print("mlp full potential", flush=True)
  X, lab = data(N=1000)
  th = np.stack([M.init_params(30 + c, *DIMS) for c in range(2)])
  eng = _obabo_engine(X, lab, th, "obabo", None, None, np.full(2, 1e-3, np.float32), steps=1)
  out = eng.full_potential()
  fam = om.Family("mlp", X, lab, dims=DIMS + (10,))
  for c in range(2):
    ref = om.full_potential(fam, th[c].astype(np.float64), 1.0, mb=64)
    assert abs(out[c].item() - ref) <= 1e-2 * abs(ref), (out[c].item(), ref)


def test_mlp_sggmc_mh_decision_and_revert():
  """sggmc's MH step (solver.py:500-567) on chosen full potentials: the decision
  follows log u < min(0, -(dU + dKE)/T) with u from the solver key, a rejection
  restores theta, momentum and the step potential, an acceptance keeps the proposal;
  the kinetic energies reset and the iteration counter advances."""
  print("mlp sggmc mh", flush=True)
  X, lab = data()
  th = np.stack([M.init_params(40 + c, *DIMS) for c in range(2)])
  keys = [(0, 5), (0, 6)]
  eng = _obabo_engine(X, lab, th, "sggmc", keys, [(1, 1), (1, 2)],
                      np.full(4, 2e-3, np.float32), steps=2, slots=np.arange(4))
  mh_keys = eng.mh_key.cpu().numpy().view(np.uint32).copy()
  for _ in range(eng.steps):
    eng.obabo_step()
  torch.cuda.synchronize()
  th_prop = eng.theta.clone()
  mom_prop = eng.mom.clone()
  U_step = eng.U.clone()
  ke = eng.ke.cpu().numpy().copy()
  U_full = eng.U_full.cpu().numpy().copy()
  # chain 0: a proposal far uphill (rejected for any u); chain 1: dE = -1 (accepted)
  dKE = ke[:, 1] - ke[:, 0]
  U_prop = np.array([U_full[0] + 1e6, U_full[1] - dKE[1] - 1.0])
  eng.U_prop.copy_(torch.as_tensor(U_prop, device=DEV))
  st = eng._state()
  import ctypes
  from jax_sgmc_amd import _lib
  _lib.call("sgmc_mlp_mh", ctypes.byref(st), _lib.stream())
  torch.cuda.synchronize()
  for c in range(2):
    _, sub = jr.split(mh_keys[c])
    log_u = jr.logf_f32(jr.uniform(sub).reshape(1))[0]
    log_a = min(0.0, -(U_prop[c] - U_full[c] + dKE[c]))
    acc = bool(log_u < log_a)
    assert acc == (c == 1)
  assert torch.equal(eng.theta[0], torch.as_tensor(th[0], device=DEV))
  assert torch.equal(eng.mom[0], torch.zeros_like(eng.mom[0]))
  assert eng.U[0].item() == 0.0  # the initial ObaboState potential
  assert torch.equal(eng.theta[1], th_prop[1]) and torch.equal(eng.mom[1], mom_prop[1])
  assert torch.equal(eng.theta_old[1], th_prop[1])
  assert eng.U_full[1].item() == U_prop[1] and eng.U_full[0].item() == U_full[0]
  assert torch.all(eng.ke == 0)
  assert eng.iter.item() == 1 and eng.accepted.tolist() == [0, 1]
  np.testing.assert_array_equal(eng.samples[1, 0].cpu().numpy(), th_prop[1].cpu().numpy())
  assert eng.samples_U[1, 0].item() == U_step[1].item()


def test_mlp_sggmc_graph_replay_matches_eager():
  """Two SGGMC iterations (OBABO proposals, full-data potential, MH) replayed from a
  captured graph give the eager result bit for bit."""
  print("mlp sggmc graph", flush=True)
  X, lab = data()
  th = np.stack([M.init_params(50 + c, *DIMS) for c in range(2)])
  mk = lambda: _obabo_engine(X, lab, th, "sggmc", [(0, 1), (0, 2)], [(1, 3), (1, 4)],
                             np.full(4, 1e-3, np.float32), steps=2, slots=np.arange(4))
  a, b = mk(), mk()
  a.run(2, graph_iterations=1)
  b.run(2)
  torch.cuda.synchronize()
  assert torch.equal(a.theta, b.theta) and torch.equal(a.mom, b.mom)
  assert torch.equal(a.U_full, b.U_full) and torch.equal(a.samples, b.samples)
  assert a.iter.item() == 2 == b.iter.item()


def test_mlp_fp8_gradient_against_oracle():
  """e4m3 GEMM operands with per-tensor delayed scaling (after the engine's
  calibration): potential within 3e-2 and gradient within 0.15 (relative norm) of the
  float64 oracle on the same minibatch; the noise and update arithmetic stay float32."""
  print("mlp fp8 gradient", flush=True)
  X, lab = data()
  keys, dkeys = [(0, 1), (0, 2)], [(3, 4), (5, 6)]
  th = np.stack([M.init_params(10 + c, *DIMS) for c in range(2)])
  eng = M.FusedMLP(torch.as_tensor(X, device=DEV), torch.as_tensor(lab, device=DEV), DIMS, 64,
                   th, np.full(8, 1e-5, np.float32), keys=keys, dkeys=dkeys, prior_scale=1.0,
                   precision="fp8")
  U, g = eng.potential_and_grad()
  torch.cuda.synchronize()
  for c in range(2):
    Ur, gr, _ = oracle_potential(X, lab, th[c], dkeys[c], 64)
    gk = g[c].cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(gk - gr) / np.linalg.norm(gr)
    print("chain", c, "U", U[c].item(), Ur, "grad rel", rel, flush=True)
    assert abs(U[c].item() - Ur) <= 3e-2 * abs(Ur)
    assert rel < 0.15, rel
  q = eng.q.cpu().numpy()
  assert np.all(q[:, :6] > 0) and np.all(np.isfinite(q[:, :8]))


def test_mlp_fp8_sggmc_graph_replay_matches_eager():
  print("mlp fp8 sggmc graph", flush=True)
  X, lab = data()
  th = np.stack([M.init_params(60 + c, *DIMS) for c in range(2)])
  mk = lambda: M.FusedMLP(torch.as_tensor(X, device=DEV), torch.as_tensor(lab, device=DEV), DIMS,
                          64, th, np.full(4, 1e-3, np.float32), keys=[(0, 1), (0, 2)],
                          dkeys=[(1, 3), (1, 4)], prior_scale=1.0, integrator="sggmc", steps=2,
                          save_slots=np.arange(4), precision="fp8")
  a, b = mk(), mk()
  a.run(2, graph_iterations=1)
  b.run(2)
  torch.cuda.synchronize()
  assert torch.isfinite(a.theta).all()
  assert torch.equal(a.theta, b.theta) and torch.equal(a.U_full, b.U_full)
  assert torch.equal(a.q, b.q)


def test_mlp_head_rows_per_wavefront_bit_equal():
  """The output-layer kernel takes four rows per wavefront once a launch has 8192 rows
  and one row below that; every row's sums run in the same order either way. Chain 0 of
  a 16-chain engine (8192 rows: four per wavefront) equals the same chain alone (512
  rows: one per wavefront) bit for bit, potential and gradient, with the GEMM plan
  fixed; and it stays within the bf16 bounds of the float64 oracle."""
  print("mlp head rows per wavefront", flush=True)
  X, lab = data()
  n = 512
  dk = [(3, 100 + c) for c in range(16)]
  big, th = engine(X, lab, C=16, n=n, dkeys=dk)
  one, _ = engine(X, lab, C=1, n=n, dkeys=dk[:1])
  fixed = [128 * 16 + 1, 128 * 16 + 1, 128 * 16 + 1, 128 * 16 + 1, 128 * 16 + 1]
  big.tune = list(fixed)
  one.tune = list(fixed)
  Ub, gb = big.potential_and_grad()
  Uo, go = one.potential_and_grad()
  torch.cuda.synchronize()
  assert torch.equal(Ub[0], Uo[0]) and torch.equal(gb[0], go[0])
  for c in (0, 15):
    Ur, gr, _ = oracle_potential(X, lab, th[c], dk[c], n)
    assert abs(Ub[c].item() - Ur) <= 1e-2 * abs(Ur), (c, Ub[c].item(), Ur)
    gk = gb[c].cpu().numpy().astype(np.float64)
    assert np.linalg.norm(gk - gr) / np.linalg.norm(gr) < 3e-2
