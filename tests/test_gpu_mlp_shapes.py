"""MLP engine at the benchmark shapes against the float64 oracle (GPU).

pSGLD MLP 3072-512-512-10, n = 256, bf16 GEMMs, rms_prop, 2 chains; SGGMC MLP
3072-1024-1024-10, n = 512, e4m3 GEMMs, 2 chains. At these shapes the forward GEMMs
(K = 3072, and K = 512 / 1024 for the second layer and dA1) split K in two: the split-K
partial GEMM and its finish kernel, the multi-k-tile prefetch loop of the deep forward
pass, and (fp8) the finish kernel's block amax, all run here. Checks, each against the
oracle on the same minibatch (indices from the same data keys):
- potential within 1e-2 (bf16) / 3e-2 (e4m3) relative; gradient within 3e-2 (bf16) /
  0.15 (e4m3) relative in norm, overall and per leaf;
- the update follows the reference's rule given the kernel's gradient, with the exact
  noise words (integrator.py:882-912, adaption.py:254-291; OBABO's first half step,
  integrator.py:203-273);
- e4m3 delayed scaling: after an evaluation the activation scales are 224 / max|v| of
  that evaluation's tensors (inputs exactly; hidden activations and backward
  activations within the e4m3 forward's error of the float64 values);
- a full SGGMC iteration (OBABO proposal, float64 full-data potential, MH test): finite
  state and a kept full potential within 3e-2 (e4m3) / 1e-5 (float32) of the oracle's
  at the kept position.
The float32 GEMM path (precision "fp32") is checked at the same shapes to 1e-5 (potential)
and 1e-4 (gradient).
Dataset: 4096 rows (the minibatch draw spans N; the shapes are the benchmark's)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import jaxrandom as jr
from oracle import model as om
from oracle import steps as ost

from jax_sgmc_amd import _lib
from jax_sgmc_amd import mlp as M

pytestmark = pytest.mark.gpu
DEV = "cuda"
N = 4096
C3 = (3072, 512, 512)
C5 = (3072, 1024, 1024)
LEAVES = ("W1", "W2", "W3", "b1", "b2", "b3")
_DATA = {}


def data(dims):
  if dims not in _DATA:
    rng = np.random.default_rng(7)
    X = rng.random((N, dims[0]), dtype=np.float32)
    W1, W2, W3, b1, b2, b3 = om.mlp_unflatten(M.init_params(99, *dims), dims + (10,))
    h = np.maximum(np.maximum(X @ W1 + b1, 0) @ W2 + b2, 0)
    labels = np.argmax(h @ W3 + b3, axis=1).astype(np.int32)
    _DATA[dims] = (X, labels, torch.as_tensor(X, device=DEV),
                   torch.as_tensor(labels, device=DEV))
  return _DATA[dims]


def leaf_sizes(dims):
  D, H1, H2 = dims
  return [D * H1, H1 * H2, H2 * 10, H1, H2, 10]


def oracle_eval(X, labels, dims, theta, dkey, n):
  """U, grad and the intermediate tensors of the float64 oracle on the minibatch the
  data key draws (data/numpy_loader.py:132-134)."""
  fam = om.Family("mlp", X, labels, dims=dims + (10,))
  _, idx = ost.draw_indices(np.asarray(dkey, np.uint32), X.shape[0], n)
  U, g, _ = om.stochastic_potential(fam, theta.astype(np.float64), idx, 1.0)
  W1, W2, W3, b1, b2, b3 = om.mlp_unflatten(theta, dims + (10,))
  Xb = X[idx].astype(np.float64)
  a1 = Xb @ W1 + b1
  a2 = np.maximum(a1, 0) @ W2 + b2
  logits = np.maximum(a2, 0) @ W3 + b3
  p = np.exp(logits - logits.max(1, keepdims=True))
  p /= p.sum(1, keepdims=True)
  onehot = np.eye(10)[labels[idx]]
  dA2 = ((-X.shape[0] / n) * (onehot - p) @ W3.T) * (a2 > 0)
  dA1 = (dA2 @ W2.T) * (a1 > 0)
  return U, g, dict(x=np.abs(X[idx]).max(), h1=np.maximum(a1, 0).max(), da2=np.abs(dA2).max(),
                    da1=np.abs(dA1).max())


def check_gradient(U, g, Ur, gr, dims, u_tol, g_tol):
  assert abs(U - Ur) <= u_tol * abs(Ur), (U, Ur)
  rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
  assert rel < g_tol, rel
  off = 0
  for name, size in zip(LEAVES, leaf_sizes(dims)):
    a, b = g[off:off + size], gr[off:off + size]
    lrel = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  {name}: rel {lrel:.3e}", flush=True)
    assert lrel < g_tol, (name, lrel)
    off += size
  return rel


def c3_engine(keys, dkeys, eps=1e-5):
  X, lab, Xd, labd = data(C3)
  th = np.stack([M.init_params(10 + c, *C3) for c in range(2)])
  eng = M.FusedMLP(Xd, labd, C3, 256, th, np.full(4, eps, np.float32), keys=keys, dkeys=dkeys,
                   rms=True, prior_scale=1.0)
  return eng, th, X, lab


def test_c3_bf16_potential_and_gradient():
  print("C3 shape: bf16 potential and gradient, 2 chains, n=256", flush=True)
  dkeys = [(3, 4), (5, 6)]
  eng, th, X, lab = c3_engine([(0, 1), (0, 2)], dkeys)
  U, g = eng.potential_and_grad()
  torch.cuda.synchronize()
  for c in range(2):
    Ur, gr, _ = oracle_eval(X, lab, C3, th[c], dkeys[c], 256)
    rel = check_gradient(U[c].item(), g[c].cpu().numpy().astype(np.float64), Ur, gr, C3, 1e-2,
                         3e-2)
    print(f"chain {c}: U {U[c].item():.6e} oracle {Ur:.6e}, grad rel {rel:.3e}", flush=True)


def test_c3_bf16_rms_update_follows_reference_rule():
  print("C3 shape: rms_prop update with the kernel gradient", flush=True)
  keys = [(0, 7), (0, 8)]
  eps = 1e-5
  eng, th, X, lab = c3_engine(keys, [(1, 1), (1, 2)], eps)
  _, g = eng.potential_and_grad()
  g = g.cpu().numpy()
  eng.iteration()
  torch.cuda.synchronize()
  th1 = eng.theta.cpu().numpy()
  v1 = eng.v.cpu().numpy()
  f = np.float32
  for c in range(2):
    nk, sub = jr.split(np.asarray(keys[c], np.uint32))
    xi = jr.random_tree_flat(sub, leaf_sizes(C3))
    sg = f(-eps) * g[c]
    sn = np.sqrt(f(f(2) * f(1)) * f(eps)) * xi
    v = f(0.9) * np.ones_like(g[c]) + (f(1) - f(0.9)) * (g[c] * g[c])
    G = f(1) / (f(1e-5) + np.sqrt(v))
    ref = th[c] + ((f(0) + G * sg) + np.sqrt(G) * sn)
    np.testing.assert_allclose(v1[c], v, rtol=1e-6)
    np.testing.assert_allclose(th1[c], ref, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(eng.keys[c].cpu().numpy().view(np.uint32), nk)
  assert eng.iter.item() == 1


def test_c3_fp32_potential_and_gradient():
  """float32 GEMMs (precision "fp32", the plain path): potential within 1e-5 and the
  gradient within 1e-4 relative (overall and per leaf) of the float64 oracle."""
  print("C3 shape: fp32 potential and gradient, 2 chains, n=256", flush=True)
  X, lab, Xd, labd = data(C3)
  dkeys = [(3, 4), (5, 6)]
  th = np.stack([M.init_params(10 + c, *C3) for c in range(2)])
  eng = M.FusedMLP(Xd, labd, C3, 256, th, np.full(4, 1e-5, np.float32), keys=[(0, 1), (0, 2)],
                   dkeys=dkeys, rms=True, prior_scale=1.0, precision="fp32")
  assert eng.shadow is None
  U, g = eng.potential_and_grad()
  torch.cuda.synchronize()
  for c in range(2):
    Ur, gr, _ = oracle_eval(X, lab, C3, th[c], dkeys[c], 256)
    rel = check_gradient(U[c].item(), g[c].cpu().numpy().astype(np.float64), Ur, gr, C3, 1e-5,
                         1e-4)
    print(f"chain {c}: U {U[c].item():.8e} oracle {Ur:.8e}, grad rel {rel:.3e}", flush=True)


def test_c3_fp32_rms_update_follows_reference_rule():
  print("C3 shape: fp32 rms_prop update with the kernel gradient", flush=True)
  X, lab, Xd, labd = data(C3)
  keys, eps = [(0, 17), (0, 18)], 1e-5
  th = np.stack([M.init_params(30 + c, *C3) for c in range(2)])
  eng = M.FusedMLP(Xd, labd, C3, 256, th, np.full(4, eps, np.float32), keys=keys,
                   dkeys=[(2, 1), (2, 2)], rms=True, prior_scale=1.0, precision="fp32")
  _, g = eng.potential_and_grad()
  g = g.cpu().numpy()
  eng.iteration()
  torch.cuda.synchronize()
  f = np.float32
  for c in range(2):
    nk, sub = jr.split(np.asarray(keys[c], np.uint32))
    xi = jr.random_tree_flat(sub, leaf_sizes(C3))
    v = f(0.9) * np.ones_like(g[c]) + (f(1) - f(0.9)) * (g[c] * g[c])
    G = f(1) / (f(1e-5) + np.sqrt(v))
    ref = th[c] + ((f(0) + G * (f(-eps) * g[c])) + np.sqrt(G) * (np.sqrt(f(2) * f(eps)) * xi))
    np.testing.assert_allclose(eng.theta[c].cpu().numpy(), ref, rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(eng.keys[c].cpu().numpy().view(np.uint32), nk)


def c5_engine(keys, dkeys, eps=2e-5, steps=2, integrator="sggmc", precision="fp8"):
  X, lab, Xd, labd = data(C5)
  th = np.stack([M.init_params(20 + c, *C5) for c in range(2)])
  eng = M.FusedMLP(Xd, labd, C5, 512, th, np.full(8, eps, np.float32), keys=keys, dkeys=dkeys,
                   prior_scale=1.0, integrator=integrator, friction=1.0, steps=steps,
                   save_slots=np.arange(8), precision=precision)
  return eng, th, X, lab


def test_c5_fp8_potential_gradient_and_scales():
  print("C5 shape: e4m3 potential, gradient and delayed scales, 2 chains, n=512", flush=True)
  dk = [(3, 14), (5, 16)]
  # an engine without the full-data pass, so that the amax words hold only this
  # evaluation's maxima
  eng, th, X, lab = c5_engine([(0, 1), (0, 2)], dk, integrator="sgld")
  U, g = eng.potential_and_grad()
  torch.cuda.synchronize()
  q = eng.q.cpu().numpy()
  f = np.float32
  for c in range(2):
    Ur, gr, amax = oracle_eval(X, lab, C5, th[c], dk[c], 512)
    rel = check_gradient(U[c].item(), g[c].cpu().numpy().astype(np.float64), Ur, gr, C5, 3e-2,
                         0.15)
    print(f"chain {c}: U {U[c].item():.6e} oracle {Ur:.6e}, grad rel {rel:.3e}", flush=True)
    # inputs: exact (the gather's amax of float32 rows); the rest within the error of
    # the e4m3 forward / backward passes
    assert q[c, 0] == f(224) / f(amax["x"]), (q[c, 0], amax["x"])
    for t, name, tol in ((1, "h1", 0.05), (2, "da2", 0.10), (3, "da1", 0.15)):
      want = 224.0 / amax[name]
      print(f"  scale {name}: {q[c, t]:.6e} oracle {want:.6e}", flush=True)
      assert abs(q[c, t] - want) <= tol * want, (name, q[c, t], want)
    # the amax words were consumed
    assert np.all(q[c, 8:12].view(np.uint32) == 0)


@pytest.mark.parametrize("precision", ["fp8", "fp32"])
def test_c5_obabo_half_step_follows_reference_rule(precision):
  """The first half of an OBABO step (O B A: integrator.py:203-273) at the C5 shape
  given the kernel's gradient at theta: p1 = sqrt(pi) p + sqrt((1 - pi) T) xi(s1),
  p2 = -eps/2 g + p1, theta' = theta + eps p2, with (key', s1, s2) = split(key, 3)."""
  print("C5 shape: OBABO first half step with the kernel gradient,", precision, flush=True)
  eps = 2e-5
  eng, th, X, lab = c5_engine([(0, 3), (0, 4)], [(7, 1), (7, 2)], eps, precision=precision)
  keys = eng.keys.cpu().numpy().view(np.uint32).copy()
  q0 = eng.q.clone() if eng.q is not None else None
  _, g = eng.potential_and_grad()
  if q0 is not None:
    eng.q.copy_(q0)  # the step evaluates with the scales the gradient call used
  g = g.cpu().numpy()
  st = eng._state()
  st.mode = 1
  _lib.call("sgmc_mlp_step", ctypes.byref(st), 1, _lib.stream())
  torch.cuda.synchronize()
  f = np.float32
  perm = np.exp(f(-f(1.0)) * f(eps)).astype(f)
  decay, noise = np.sqrt(perm).astype(f), np.sqrt(((f(1) - perm) * f(1))).astype(f)
  half = f(0.5) * f(eps)
  for c in range(2):
    _, s1, _ = jr.split(keys[c], 3)
    xi = jr.random_tree_flat(s1, leaf_sizes(C5))
    p1 = decay * np.zeros_like(xi) + noise * xi
    p2 = (-half) * g[c] + p1
    ref = th[c] + f(eps) * p2
    # the kernel's update is written with fused multiply-adds and its normals are within
    # 3 ulp of the oracle's: agreement to 1e-6 relative, with an absolute floor of 1e-6
    # of the noise term where p2 = p1 - (eps/2) g nearly cancels
    floor = 1e-6 * float(np.abs(p1).max())
    np.testing.assert_allclose(eng.mom[c].cpu().numpy(), p2, rtol=1e-6, atol=floor)
    np.testing.assert_allclose(eng.theta[c].cpu().numpy(), ref, rtol=1e-6,
                               atol=f(eps) * floor)
    ks = eng.ke[c, 0].item()
    ks_ref = 0.5 * np.dot(p1.astype(np.float64), p1.astype(np.float64))
    assert abs(ks - ks_ref) <= 1e-5 * ks_ref, (ks, ks_ref)


@pytest.mark.parametrize("precision,tol", [("fp8", 3e-2), ("fp32", 1e-5)])
def test_c5_sggmc_iterations_against_oracle(precision, tol):
  print("C5 shape: two SGGMC iterations (2 OBABO steps each),", precision, flush=True)
  eng, th, X, lab = c5_engine([(0, 5), (0, 6)], [(9, 1), (9, 2)], precision=precision)
  eng.run(2)
  torch.cuda.synchronize()
  assert torch.isfinite(eng.theta).all() and torch.isfinite(eng.mom).all()
  assert torch.isfinite(eng.U_full).all() and eng.iter.item() == 2
  fam = om.Family("mlp", X, lab, dims=C5 + (10,))
  for c in range(2):
    ref = om.full_potential(fam, eng.theta[c].cpu().numpy().astype(np.float64), 1.0, mb=512)
    got = eng.U_full[c].item()
    print(f"chain {c}: kept U_full {got:.8e} oracle {ref:.8e}, accepted "
          f"{eng.accepted[c].item()}", flush=True)
    assert abs(got - ref) <= tol * abs(ref), (got, ref)
  if eng.q is not None:
    q = eng.q.cpu().numpy()
    assert np.all(np.isfinite(q[:, :8])) and np.all(q[:, :8] > 0)
