"""Kernel-level parity of libsgmc against the oracle (GPU).

RNG words, indices and normals must match bit for bit; potentials and gradients
within 1e-5 relative (atol floor 1e-5 * max|x|) of the float64 oracle.
"""
import math

import numpy as np
import pytest
import torch

from oracle import jaxrandom as jr
from oracle import model as om

import jax_sgmc_amd._lib as L
from jax_sgmc_amd import random as R
from jax_sgmc_amd import tree as T

pytestmark = pytest.mark.gpu

DEV = "cuda"


def words(t):
  return t.detach().cpu().numpy().view(np.uint32)


def close(a, b, rtol=1e-5, floor=1e-5):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  tol = rtol * np.abs(b) + floor * max(1.0, float(np.max(np.abs(b))) if b.size else 1.0)
  return np.all(np.abs(a - b) <= tol), float(np.max(np.abs(a - b) / (np.abs(b) + 1e-30)))


KEYS = [(0, 0), (0, 42), (0x12345678, 0x9abcdef0), (4146024105, 967050713)]


@pytest.mark.parametrize("num", [1, 2, 3, 5, 64])
@pytest.mark.parametrize("k", KEYS)
def test_split_exact(k, num):
  print("split", k, num, flush=True)
  key = R.key_from_words(k)
  got = words(R.split(key, num))
  np.testing.assert_array_equal(got, jr.split(np.asarray(k, np.uint32), num))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 1000, 1001, 65537])
@pytest.mark.parametrize("k", KEYS[:3])
def test_bits_uniform_normal_exact(k, n):
  print("draws", k, n, flush=True)
  key = R.key_from_words(k)
  kk = np.asarray(k, np.uint32)
  np.testing.assert_array_equal(words(R.bits(key, (n,))), jr.random_bits(kk, (n,)))
  u = R.uniform(key, (n,), -2.0, 3.0).cpu().numpy()
  np.testing.assert_array_equal(u.view(np.uint32),
                                jr.uniform(kk, (n,), -2.0, 3.0).view(np.uint32))
  z = R.normal(key, (n,)).cpu().numpy()
  ref = jr.normal(kk, (n,))
  mism = np.count_nonzero(z.view(np.uint32) != ref.view(np.uint32))
  assert mism == 0, f"{mism} of {n} normals differ in bits; max abs {np.max(np.abs(z - ref))}"


@pytest.mark.parametrize("N", [1, 10, 10000, 65536, 65537, 1000000, 2 ** 31 - 1])
@pytest.mark.parametrize("n", [1, 5, 128, 4097])
def test_randint_exact(N, n):
  print("randint", N, n, flush=True)
  k = (7, 11)
  key = R.key_from_words(k)
  got = R.randint(key, (n,), 0, N).cpu().numpy()
  np.testing.assert_array_equal(got, jr.randint(np.asarray(k, np.uint32), (n,), 0, N))


def test_randint_minval_and_empty_span():
  key = R.key_from_words((3, 4))
  kk = np.asarray((3, 4), np.uint32)
  np.testing.assert_array_equal(R.randint(key, (9,), -5, 17).cpu().numpy(),
                                jr.randint(kk, (9,), -5, 17))
  np.testing.assert_array_equal(R.randint(key, (9,), 4, 4).cpu().numpy(),
                                jr.randint(kk, (9,), 4, 4))


def test_tree_normal_exact():
  tree = {"b": torch.zeros(3), "a": torch.zeros(1), "c": torch.zeros(7, 300),
          "d": [torch.zeros(2048), torch.zeros(2501)]}
  lay = T.Layout(tree)
  k = (9, 99)
  got = R.random_tree(R.key_from_words(k), lay).cpu().numpy()
  ref = jr.random_tree_flat(np.asarray(k, np.uint32), lay.sizes)
  np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("N,n", [(10000, 128), (1000000, 4096), (7, 3)])
def test_draw_batch_and_gather(N, n):
  print("draw_batch", N, n, flush=True)
  dk = torch.tensor([0, 0], dtype=torch.int32, device=DEV)
  dk2 = torch.empty_like(dk)
  idx = torch.empty(n, dtype=torch.int32, device=DEV)
  L.call("sgmc_draw_batch", L.ptr(dk), L.ptr(dk2), L.ptr(idx), n, N, L.stream())
  ref_key, ref_idx = jr.split(jr.PRNGKey(0))[0], None
  sub = jr.split(jr.PRNGKey(0))[1]
  ref_idx = jr.randint(sub, (n,), 0, N)
  np.testing.assert_array_equal(words(dk2), ref_key)
  np.testing.assert_array_equal(idx.cpu().numpy(), ref_idx)
  D = 12
  src = torch.arange(N * D, dtype=torch.float32, device=DEV).reshape(N, D)
  dst = torch.full((n, D), -1.0, device=DEV)
  L.call("sgmc_gather_rows", L.ptr(src), N, L.ptr(idx), L.ptr(dst), n, D * 4, D * 4,
         L.stream())
  np.testing.assert_array_equal(dst.cpu().numpy(), src.cpu().numpy()[ref_idx])


def _glm_data(N, D, family, seed=0):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D).astype(np.float32)
  if family == L.LIK_LOGISTIC:
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  else:
    y = (X @ w + 0.5 * rng.standard_normal(N)).astype(np.float32)
  return X, y


def _padded(X):
  N, D = X.shape
  ldx = 4 * ((D + 3) // 4)
  Xp = np.zeros((N, ldx), np.float32)
  Xp[:, :D] = X
  return torch.as_tensor(Xp, device=DEV), ldx


@pytest.mark.parametrize("family", [L.LIK_LINEAR_GAUSSIAN, L.LIK_LOGISTIC])
@pytest.mark.parametrize("D,n", [(3, 7), (16, 128), (32, 128), (64, 300), (100, 33),
                                 (512, 4096), (1024, 513)])
@pytest.mark.parametrize("gathered", [False, True])
def test_glm_potential(family, D, n, gathered):
  print("glm_potential", family, D, n, gathered, flush=True)
  N = 5000
  X, y = _glm_data(N, D, family)
  rng = np.random.default_rng(1)
  theta = (0.3 * rng.standard_normal(D)).astype(np.float32)
  idx = rng.integers(0, N, n).astype(np.int32)
  sigma, pscale, temp = 0.5, 3.0, 1.7
  Xd, ldx = _padded(X)
  yd = torch.as_tensor(y, device=DEV)
  idx_d = torch.as_tensor(idx, device=DEV)
  if gathered:
    Xb, yb = Xd[idx_d].contiguous(), yd[idx_d].contiguous()
    src_idx = None
  else:
    Xb, yb, src_idx = Xd, yd, idx_d
  nb = L.query("sgmc_glm_blocks", n, D)
  for blocks in (nb, 3):
    ws = torch.zeros(int(L.query("sgmc_glm_workspace_floats", n, D, blocks)), device=DEV)
    U = torch.empty(1, device=DEV)
    g = torch.empty(D, device=DEV)
    lik = torch.empty(n, device=DEV)
    var = torch.empty(1, device=DEV)
    th = torch.as_tensor(theta, device=DEV)
    for _ in range(2):  # the second call reuses the workspace (counter reset)
      L.call("sgmc_glm_potential", family, L.ptr(Xb), ldx, L.ptr(yb), L.ptr(src_idx), n, N, D,
             L.ptr(th), sigma, 1, pscale, temp, L.ptr(U), L.ptr(g), L.ptr(lik), L.ptr(var),
             L.ptr(ws), ws.numel(), blocks, L.stream())
    kind = "linear_gaussian" if family == L.LIK_LINEAR_GAUSSIAN else "logistic"
    fam = om.Family(kind, X, y, sigma=sigma)
    Ur, gr, likr = om.stochastic_potential(fam, theta, idx, pscale, temp)
    ok, err = close(U.item(), Ur)
    assert ok, f"U {U.item()} vs {Ur} (rel {err})"
    ok, err = close(g.cpu().numpy(), gr)
    assert ok, f"grad rel err {err}"
    ok, err = close(lik.cpu().numpy(), likr)
    assert ok, f"lik rel err {err}"
    ok, err = close(var.item(), np.var(likr), rtol=1e-4)
    assert ok, f"var rel err {err}"


def test_glm_potential_block_handoff_repeatable():
  """Many blocks hand their partial records to the last block of the launch; every
  repetition (with unrelated device writes in between) gives bit-identical results that
  match the oracle."""
  print("glm_potential handoff", flush=True)
  N, D, n = 20000, 512, 4096
  X, y = _glm_data(N, D, L.LIK_LOGISTIC)
  rng = np.random.default_rng(7)
  theta = (0.3 * rng.standard_normal(D)).astype(np.float32)
  idx = rng.integers(0, N, n).astype(np.int32)
  Xd, ldx = _padded(X)
  yd = torch.as_tensor(y, device=DEV)
  idx_d = torch.as_tensor(idx, device=DEV)
  th = torch.as_tensor(theta, device=DEV)
  blocks = 1024
  ws = torch.zeros(int(L.query("sgmc_glm_workspace_floats", n, D, blocks)), device=DEV)
  U = torch.empty(1, device=DEV)
  g = torch.empty(D, device=DEV)
  var = torch.empty(1, device=DEV)
  junk = torch.empty(1 << 22, device=DEV)
  outs = []
  for r in range(24):
    junk.fill_(float(r))
    L.call("sgmc_glm_potential", L.LIK_LOGISTIC, L.ptr(Xd), ldx, L.ptr(yd), L.ptr(idx_d), n, N,
           D, L.ptr(th), 1.0, 1, 1.0, 1.0, L.ptr(U), L.ptr(g), None, L.ptr(var), L.ptr(ws),
           ws.numel(), blocks, L.stream())
    outs.append(torch.cat([U, g, var]).clone())
  for o in outs[1:]:
    assert torch.equal(o, outs[0])
  Ur, gr, likr = om.stochastic_potential(om.Family("logistic", X, y), theta, idx, 1.0, 1.0)
  assert close(U.item(), Ur)[0]
  assert close(g.cpu().numpy(), gr)[0]
  assert close(var.item(), np.var(likr), rtol=1e-4)[0]


@pytest.mark.parametrize("blocks,D", [(257, 64), (1024, 1024), (4093, 64), (2048, 512)])
def test_glm_block_handoff_stress(blocks, D):
  """The cross-block hand-off (records written with agent-scope atomic stores, completed
  before the arrival counter's increment, read by the last block with agent-scope atomic
  loads) under stress: 150 launches with block counts that spread the last-arriving
  block over all XCDs, L2-resident writes of unrelated data and a concurrent kernel on a
  second stream between launches. Every launch must be bit-identical to the first and
  the first must match the oracle."""
  print("glm handoff stress", blocks, D, flush=True)
  N, n = 30000, 4096
  X, y = _glm_data(N, D, L.LIK_LOGISTIC, seed=blocks)
  rng = np.random.default_rng(blocks)
  theta = (0.3 * rng.standard_normal(D)).astype(np.float32)
  idx = rng.integers(0, N, n).astype(np.int32)
  Xd, ldx = _padded(X)
  yd = torch.as_tensor(y, device=DEV)
  idx_d = torch.as_tensor(idx, device=DEV)
  th = torch.as_tensor(theta, device=DEV)
  ws = torch.zeros(int(L.query("sgmc_glm_workspace_floats", n, D, blocks)), device=DEV)
  U = torch.empty(1, device=DEV)
  g = torch.empty(D, device=DEV)
  var = torch.empty(1, device=DEV)
  junk = torch.empty(1 << 20, device=DEV)
  side = torch.cuda.Stream()
  other = torch.empty(1 << 22, device=DEV)
  first = None
  for r in range(150):
    junk.fill_(float(r))
    with torch.cuda.stream(side):
      other.mul_(0.5).add_(1.0)
    L.call("sgmc_glm_potential", L.LIK_LOGISTIC, L.ptr(Xd), ldx, L.ptr(yd), L.ptr(idx_d), n, N,
           D, L.ptr(th), 1.0, 1, 1.0, 1.0, L.ptr(U), L.ptr(g), None, L.ptr(var), L.ptr(ws),
           ws.numel(), blocks, L.stream())
    out = torch.cat([U, g, var])
    if first is None:
      first = out.clone()
    else:
      assert torch.equal(out, first), f"launch {r} differs"
  torch.cuda.synchronize()
  Ur, gr, likr = om.stochastic_potential(om.Family("logistic", X, y), theta, idx, 1.0, 1.0)
  assert close(first[0].item(), Ur)[0]
  assert close(first[1:1 + D].cpu().numpy(), gr)[0]


@pytest.mark.parametrize("family", [L.LIK_LINEAR_GAUSSIAN, L.LIK_LOGISTIC])
@pytest.mark.parametrize("N,D", [(1, 3), (1000, 32), (12345, 512)])
def test_glm_full_potential_f64(family, N, D):
  print("glm_full", family, N, D, flush=True)
  X, y = _glm_data(N, D, family, seed=3)
  theta = (0.2 * np.random.default_rng(4).standard_normal(D)).astype(np.float32)
  Xd, ldx = _padded(X)
  yd = torch.as_tensor(y, device=DEV)
  th = torch.as_tensor(theta, device=DEV)
  rpb = 512
  ws = torch.zeros(int(L.query("sgmc_glm_full_workspace_doubles", N, rpb)),
                   dtype=torch.float64, device=DEV)
  out = torch.empty(1, dtype=torch.float64, device=DEV)
  L.call("sgmc_glm_full_potential", family, L.ptr(Xd), ldx, L.ptr(yd), N, D, L.ptr(th), 0.5, 1,
         1.0, 1.0, L.ptr(out), L.ptr(ws), ws.numel(), rpb, L.stream())
  kind = "linear_gaussian" if family == L.LIK_LINEAR_GAUSSIAN else "logistic"
  ref = om.full_potential(om.Family(kind, X, y, sigma=0.5), theta, 1.0, mb=7)
  # float32 per-row likelihoods summed in float64
  assert abs(out.item() - ref) <= 1e-6 * (abs(ref) + N), (out.item(), ref)


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("family", [L.LIK_LINEAR_GAUSSIAN, L.LIK_LOGISTIC])
@pytest.mark.parametrize("N,D,C", [(1, 3, 1), (3001, 37, 33), (2500, 512, 64), (777, 1000, 70),
                                   (4097, 8, 2)])
def test_glm_full_potential_chains(family, N, D, C, split):
  """Many chains in one pass over X (fp32 matrix cores, or fp16 hi + lo halves): every
  chain's full potential against the FP64 oracle, bit-equal to a second call (fixed
  reduction order); one chain with a NaN entry (and, for the split kernel, one scaled
  far from the others) leaves the other chains' values untouched."""
  print("glm_full_chains", family, N, D, C, "split", split, flush=True)
  X, y = _glm_data(N, D, family, seed=5)
  thetas = (0.2 * np.random.default_rng(6).standard_normal((C, D))).astype(np.float32)
  if C > 2:
    thetas[1] *= 1e4    # its own power-of-two scale
  Xd, ldx = _padded(X)
  yd = torch.as_tensor(y, device=DEV)
  ld_theta = D + 3
  th = torch.zeros((C, ld_theta), device=DEV)
  th[:, :D] = torch.as_tensor(thetas, device=DEV)
  if C > 1:
    th[C - 1, D - 1] = float("nan")
  nblocks = 96
  ws = torch.zeros(int(L.query("sgmc_glm_full_chains_workspace_doubles", nblocks)),
                   dtype=torch.float64, device=DEV)
  x_exp = 14 - math.frexp(float(np.max(np.abs(X))))[1]
  outs = []
  for _ in range(2):
    out = torch.empty(C, dtype=torch.float64, device=DEV)
    head = (family, L.ptr(Xd), ldx, L.ptr(yd), N, D, L.ptr(th), ld_theta, C, 0.5, 1, 1.0, 2.0)
    tail = (L.ptr(out), L.ptr(ws), ws.numel(), nblocks, L.stream())
    if split:
      L.call("sgmc_glm_full_potential_chains_split", *head, x_exp, *tail)
    else:
      L.call("sgmc_glm_full_potential_chains", *head, *tail)
    outs.append(out.cpu().numpy())
  np.testing.assert_array_equal(outs[0], outs[1])
  kind = "linear_gaussian" if family == L.LIK_LINEAR_GAUSSIAN else "logistic"
  fam = om.Family(kind, X, y, sigma=0.5)
  for c in range(C):
    if C > 1 and c == C - 1:
      assert np.isnan(outs[0][c])
      continue
    ref = om.full_potential(fam, thetas[c], 1.0, mb=N, temperature=2.0)
    assert abs(outs[0][c] - ref) <= 1e-6 * (abs(ref) + N), (c, outs[0][c], ref)


def _layout(sizes):
  return T.Layout([torch.zeros(s) for s in sizes])


@pytest.mark.parametrize("rms", [False, True])
def test_sgld_step(rms):
  sizes = [5, 1, 3000]
  lay = _layout(sizes)
  P = lay.size
  rng = np.random.default_rng(5)
  theta = rng.standard_normal(P).astype(np.float32)
  grad = rng.standard_normal(P).astype(np.float32)
  v = rng.random(P).astype(np.float32)
  k = (1, 2)
  th, gd, vd = (torch.as_tensor(a, device=DEV) for a in (theta, grad, v))
  kin = R.key_from_words(k)
  kout = torch.empty_like(kin)
  table, nb = lay.leaf_table(th.device, L.PAIRS_PER_BLOCK)
  eps, T_, alpha, lmbd = 0.01, 1.3, 0.9, 1e-5
  L.call("sgmc_sgld_step", L.ptr(th), L.ptr(gd), L.ptr(vd) if rms else None, P, L.ptr(table),
         nb, L.ptr(kin), L.ptr(kout), eps, T_, alpha, lmbd, L.stream())
  nk, sub = jr.split(np.asarray(k, np.uint32))
  np.testing.assert_array_equal(words(kout), nk)
  xi = jr.random_tree_flat(sub, sizes)
  f = np.float32
  sg = f(-eps) * grad
  sn = np.sqrt(f(f(2) * f(T_)) * f(eps)) * xi
  if rms:
    vn = f(alpha) * v + (f(1) - f(alpha)) * (grad * grad)
    G = f(1) / (f(lmbd) + np.sqrt(vn))
    ref = theta + ((f(0) + G * sg) + np.sqrt(G) * sn)
    np.testing.assert_array_equal(vd.cpu().numpy(), vn)
  else:
    ref = theta + (sg + sn)
  np.testing.assert_array_equal(th.cpu().numpy(), ref)


def test_sghmc_kernels():
  sizes = [7, 513]
  lay = _layout(sizes)
  P = lay.size
  rng = np.random.default_rng(6)
  f = np.float32
  theta = rng.standard_normal(P).astype(f)
  grad = rng.standard_normal(P).astype(f)
  mass = (0.5 + rng.random(P)).astype(f)
  inv_m = (f(1) / mass).astype(f)
  sqrt_m = np.sqrt(mass).astype(f)
  table, nb = lay.leaf_table(torch.device(DEV), L.PAIRS_PER_BLOCK)
  th = torch.as_tensor(theta, device=DEV)
  p = torch.zeros(P, device=DEV)
  im, sm = torch.as_tensor(inv_m, device=DEV), torch.as_tensor(sqrt_m, device=DEV)
  k0 = R.key_from_words((5, 6))
  k1, k2 = torch.empty_like(k0), torch.empty_like(k0)
  L.call("sgmc_sghmc_resample", L.ptr(p), L.ptr(sm), L.ptr(table), nb, L.ptr(k0), L.ptr(k1),
         L.stream())
  a, sub = jr.split(np.asarray((5, 6), np.uint32))
  pr = sqrt_m * jr.random_tree_flat(sub, sizes)
  np.testing.assert_array_equal(p.cpu().numpy(), pr)
  eps, C = f(0.003), f(0.5)
  L.call("sgmc_sghmc_position", L.ptr(th), L.ptr(p), L.ptr(im), None, float(C), P, float(eps),
         L.stream())
  s = inv_m * pr
  thr = theta + eps * s
  pr = pr + (f(-eps) * C) * s
  np.testing.assert_array_equal(th.cpu().numpy(), thr)
  np.testing.assert_array_equal(p.cpu().numpy(), pr)
  gd = torch.as_tensor(grad, device=DEV)
  L.call("sgmc_sghmc_momentum", L.ptr(p), L.ptr(gd), None, float(C), L.ptr(table), nb,
         L.ptr(k1), L.ptr(k2), float(eps), L.stream())
  b, sub2 = jr.split(a)
  xi = jr.random_tree_flat(sub2, sizes)
  pr = pr + f(-eps) * grad
  pr = pr + C * (np.sqrt(f(2) * eps) * xi)
  np.testing.assert_array_equal(words(k2), b)
  np.testing.assert_array_equal(p.cpu().numpy(), pr)


def test_obabo_kernels():
  sizes = [3, 4000]
  lay = _layout(sizes)
  P = lay.size
  rng = np.random.default_rng(7)
  f = np.float32
  theta = rng.standard_normal(P).astype(f)
  p0 = rng.standard_normal(P).astype(f)
  grad = rng.standard_normal(P).astype(f)
  table, nb = lay.leaf_table(torch.device(DEV), L.PAIRS_PER_BLOCK)
  p = torch.as_tensor(p0, device=DEV)
  th = torch.as_tensor(theta, device=DEV)
  nkey = R.key_from_words((8, 9))
  ws = torch.zeros(int(L.query("sgmc_obabo_workspace_floats", nb)), device=DEV)
  ke_in = torch.tensor([0.25], device=DEV)
  ke_out = torch.empty(1, device=DEV)
  fr, eps, T_ = 0.7, 0.01, 1.0
  L.call("sgmc_obabo_o", L.ptr(p), None, None, L.ptr(table), nb, L.ptr(nkey), fr, eps, T_, 0,
         L.ptr(ke_in), L.ptr(ke_out), L.ptr(ws), L.stream())
  perm = np.exp(f(-f(fr) * f(eps)))
  xi = jr.random_tree_flat(np.asarray((8, 9), np.uint32), sizes)
  p1 = np.sqrt(perm) * p0 + np.sqrt(f((f(1) - perm) * f(T_))) * xi
  np.testing.assert_allclose(p.cpu().numpy(), p1, rtol=2e-6, atol=2e-6)
  ok, err = close(ke_out.item(), 0.25 + 0.5 * np.dot(p1.astype(np.float64), p1))
  assert ok, err
  L.call("sgmc_obabo_ba", L.ptr(th), L.ptr(p), L.ptr(torch.as_tensor(grad, device=DEV)), None,
         P, eps, L.stream())
  p1 = p.cpu().numpy()  # continue from the device state
  np.testing.assert_allclose(th.cpu().numpy(), theta + f(eps) * p1, rtol=1e-6, atol=1e-6)


def test_random_thinning_on_device_uniforms():
  """scheduler.random_thinning (scheduler.py:599-666) draws its Gumbel keys from the
  device generator: the selection equals the Gumbel top-k of the oracle's uniforms of
  the same key (bit-identical words), burn-in iterations excluded."""
  print("random_thinning", flush=True)
  from jax_sgmc_amd import scheduler as SC
  steps = SC.polynomial_step_size_first_last(first=1e-3, last=1e-4)
  burn = SC.initial_burn_in(20)
  th = SC.random_thinning(steps, burn, 15, key=R.PRNGKey(3, device=DEV))
  state, count = th.init(100)
  eps = steps.init(100)
  bst, _ = burn.init(100)
  probs = np.array([eps[i] * burn.get(bst, i) for i in range(100)], np.float32)
  u = jr.uniform(jr.PRNGKey(3), (100,), np.finfo(np.float32).tiny, 1.0)
  assert count == 15
  np.testing.assert_array_equal(state, SC.gumbel_topk(u, probs, 15))
  assert state.min() >= 20
  assert th.get(state, int(state[0])) and not th.get(state, 0)
