"""Fused multi-chain path against the oracle (GPU): every saved iteration of every
chain within 1e-5 relative (floor 1e-5 max|theta|) of the float64 oracle run with the
same keys; keys bit-exact; graph replay identical to eager launches."""
import numpy as np
import pytest
import torch

from oracle import model as om
from oracle import steps as ost

from jax_sgmc_amd import fused
from jax_sgmc_amd import random as R

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _data(N, D, kind, seed):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  if kind == "logistic":
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  else:
    y = (X @ w + 0.5 * rng.standard_normal(N)).astype(np.float32)
  return X, y


def _dev(X):
  N, D = X.shape
  ldx = 4 * ((D + 3) // 4)
  Xp = np.zeros((N, ldx), np.float32)
  Xp[:, :D] = X
  return torch.as_tensor(Xp, device=DEV)


def _check(got, ref, what):
  ref = np.asarray(ref, np.float64)
  tol = 1e-5 * np.abs(ref) + 1e-5 * max(1e-3, np.max(np.abs(ref)))
  err = np.abs(np.asarray(got, np.float64) - ref)
  assert np.all(err <= tol), f"{what}: max rel {np.max(err / (np.abs(ref) + 1e-30))}"


def _run_case(kind, mode, specs, iters, eps, steps=1, friction=0.5, sigma=0.5, pscale=1.0,
              graph=0, alpha=0.9, lmbd=1e-5):
  chains, fams = [], []
  for i, s in enumerate(specs):
    X, y = _data(s["N"], s["D"], kind, seed=10 + i)
    rng = np.random.default_rng(100 + i)
    th0 = (0.1 * rng.standard_normal(s["D"])).astype(np.float32)
    chains.append(dict(X=_dev(X), y=torch.as_tensor(y, device=DEV), D=s["D"], n=s["n"],
                       theta0=th0, key=(0, i), dkey=(1, 7 * i + 3),
                       temperature=s.get("T", 1.0)))
    fams.append((X, y, th0))
  eng = fused.FusedGLM(chains, kind, mode, eps, steps=steps, friction=friction, sigma=sigma,
                       prior_scale=pscale, save_slots=np.arange(iters), alpha=alpha, lmbd=lmbd)
  eng.run(iters, graph_iterations=graph)
  torch.cuda.synchronize()
  for i, s in enumerate(specs):
    X, y, th0 = fams[i]
    fam = om.Family(kind, X, y, sigma=sigma)
    pot = lambda th, idx, fam=fam: om.stochastic_potential(fam, th, idx, pscale)
    key = np.array([0, i], np.uint32)
    dkey = np.array([1, 7 * i + 3], np.uint32)
    if mode == "sghmc":
      r = ost.sghmc(pot, th0, iters, eps, steps, friction, s["N"], s["n"], [s["D"]], key=key,
                    dkey=dkey)
      _check(eng.momentum(i).cpu().numpy(), r["p"][-1], f"chain {i} momentum")
    else:
      r = ost.sgld(pot, th0, iters, eps, s.get("T", 1.0), s["N"], s["n"], [s["D"]], key=key,
                   dkey=dkey, rms=(alpha, lmbd) if mode == "sgld_rms" else None)
    tr = eng.sample_trace(i).cpu().numpy()
    for t in range(iters):
      _check(tr[t], r["theta"][t], f"chain {i} iteration {t}")
    np.testing.assert_array_equal(R.key_words(eng.keys[i]), r["key"])
    np.testing.assert_array_equal(R.key_words(eng.dkeys[i]), r["dkey"])
    Uref = r["U"][-1]
    assert abs(eng.U[i].item() - Uref) <= 1e-5 * abs(Uref) + 1e-3, (eng.U[i].item(), Uref)
  assert eng.iteration_count() == iters
  return eng


def test_fused_sgld_linear_regression():
  print("fused sgld linear", flush=True)
  eps = ost.polynomial_step_size_first_last(2e-4, 5e-5, 0.33, 30)
  _run_case("linear_gaussian", "sgld", [dict(N=2000, D=32, n=128)] * 2, 30, eps,
            pscale=np.sqrt(10.0))


def test_fused_sgld_linear_regression_full_size():
  """The SGLD linear-regression workload at its stated size: N = 10,000, D = 32,
  n = 128, one chain, the polynomial schedule from 1e-4 to 1e-5 over 10,000 iterations
  (its first 100 entries), data from jax-compatible draws of PRNGKey(0). The first
  step sizes exceed 2 / lambda_max (eps lambda_max = 4.5), so the chain first grows
  geometrically, as the reference's would; every iteration stays within 1e-5 of the
  float64 oracle relative to the iterate's scale."""
  print("fused sgld linear N=10000 D=32 n=128", flush=True)
  from oracle import jaxrandom as jr
  N, D, n, iters = 10_000, 32, 128, 100
  kx, kw, ke = jr.split(jr.PRNGKey(0), 3)
  X = jr.normal(kx, (N * D,)).reshape(N, D)
  w = jr.normal(kw, (D,))
  y = (X @ w + np.float32(0.5) * jr.normal(ke, (N,))).astype(np.float32)
  eps = ost.polynomial_step_size_first_last(1e-4, 1e-5, 0.33, 10_000)[:iters]
  ch = [dict(X=_dev(X), y=torch.as_tensor(y, device=DEV), D=D, n=n,
             theta0=np.zeros(D, np.float32), key=(0, 0), dkey=(0, 0))]
  eng = fused.FusedGLM(ch, "linear_gaussian", "sgld", eps, sigma=0.5,
                       prior_scale=float(np.sqrt(10.0)), save_slots=np.arange(iters))
  assert eng.loop_capable()
  eng.run(iters)
  torch.cuda.synchronize()
  fam = om.Family("linear_gaussian", X, y, sigma=0.5)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, np.sqrt(10.0))
  r = ost.sgld(pot, np.zeros(D), iters, eps, 1.0, N, n, [D], key=np.zeros(2, np.uint32),
               dkey=np.zeros(2, np.uint32))
  tr = eng.sample_trace(0).cpu().numpy()
  for t in range(iters):
    _check(tr[t], r["theta"][t], f"iteration {t}")
  np.testing.assert_array_equal(R.key_words(eng.keys[0]), r["key"])
  np.testing.assert_array_equal(R.key_words(eng.dkeys[0]), r["dkey"])


def test_fused_sgld_rms_logistic():
  print("fused sgld rms logistic", flush=True)
  eps = np.full(20, 1e-3, np.float32)
  _run_case("logistic", "sgld_rms", [dict(N=3000, D=64, n=100)], 20, eps)


def test_fused_sghmc_logistic_multiblock():
  print("fused sghmc logistic", flush=True)
  eps = np.full(6, 2e-3, np.float32)
  _run_case("logistic", "sghmc", [dict(N=20000, D=512, n=1024)] * 3, 6, eps, steps=5)


def test_fused_heterogeneous_table():
  print("fused heterogeneous", flush=True)
  specs = [dict(N=2000, D=16, n=64, T=1.0), dict(N=5000, D=64, n=200, T=2.5),
           dict(N=3000, D=256, n=96, T=1.0), dict(N=4000, D=1000, n=300, T=7.0)]
  eps = np.full(8, 1e-4, np.float32)
  _run_case("logistic", "sgld", specs, 8, eps)


def test_fused_graph_replay_equals_eager():
  print("fused graph", flush=True)
  eps = np.full(12, 2e-3, np.float32)
  spec = [dict(N=5000, D=128, n=256)] * 2
  e1 = _run_case("logistic", "sghmc", spec, 12, eps, steps=3, graph=4)
  e2 = _run_case("logistic", "sghmc", spec, 12, eps, steps=3, graph=0)
  assert torch.equal(e1.theta_flat, e2.theta_flat)
  assert torch.equal(e1.samples, e2.samples)


def test_fused_autotune_preserves_state():
  eps = np.full(4, 1e-3, np.float32)
  X, y = _data(4000, 512, "logistic", 0)
  ch = [dict(X=_dev(X), y=torch.as_tensor(y, device=DEV), D=512, n=1024, theta0=np.zeros(512),
             key=(0, c), dkey=(1, c)) for c in range(4)]
  a = fused.FusedGLM(ch, "logistic", "sghmc", eps, steps=2)
  b = fused.FusedGLM(ch, "logistic", "sghmc", eps, steps=2)
  a.autotune(reps=1)
  a.run(4)
  b.run(4)
  torch.cuda.synchronize()
  # block counts differ, so sums are reordered: equal to float32 rounding only
  np.testing.assert_allclose(a.theta_flat.cpu().numpy(), b.theta_flat.cpu().numpy(),
                             rtol=1e-4, atol=1e-5)


def test_fused_sghmc_logistic_c2_full_size():
  """The headline configuration at its stated size: SGHMC on logistic regression with
  N = 1e6, D = 512, n = 4096, 5 leapfrog steps, friction 0.5, 64 chains as two chain
  tables on two streams (as the bench runs it). Two iterations (10 gradient evaluations
  of 4096 rows per chain); chains 0 and 63 (one per table) against the float64 oracle:
  positions, momenta and keys, within 1e-5 relative (floor 1e-5 max|x|)."""
  print("fused SGHMC, C2 size", flush=True)
  N, D, n, C, its, steps, fric = 1_000_000, 512, 4096, 64, 2, 5, 0.5
  rng = np.random.default_rng(2024)
  X = (rng.standard_normal((N, D), dtype=np.float32) / np.float32(np.sqrt(D)))
  w = rng.standard_normal(D).astype(np.float32)
  y = (rng.random(N) < 1 / (1 + np.exp(-(X @ w)))).astype(np.float32)
  Xd, yd = torch.as_tensor(X, device=DEV), torch.as_tensor(y, device=DEV)
  eps = np.array([2e-3, 1.5e-3], np.float32)
  chains = [dict(X=Xd, y=yd, D=D, n=n, theta0=np.zeros(D, np.float32), key=(0, c),
                 dkey=(2, c)) for c in range(C)]
  grp = fused.TableGroup.split(chains, 2, "logistic", "sghmc", eps, steps=steps,
                               friction=fric, prior_scale=1.0)
  grp.run(its)
  torch.cuda.synchronize()
  fam = om.Family("logistic", X, y)
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  for c in (0, C - 1):
    r = ost.sghmc(pot, np.zeros(D), its, eps, steps, fric, N, n, [D],
                  key=np.array([0, c], np.uint32), dkey=np.array([2, c], np.uint32))
    _check(grp.theta(c).cpu().numpy(), r["theta"][-1], f"chain {c} theta")
    _check(grp.momentum(c).cpu().numpy(), r["p"][-1], f"chain {c} momentum")
    e, j = grp._locate(c)
    np.testing.assert_array_equal(R.key_words(e.keys[j]), r["key"])
    np.testing.assert_array_equal(R.key_words(e.dkeys[j]), r["dkey"])


def _ladder(R_=4, D=24, N=700, swap_every=3, iters=12, graph=0, saving=True, target_bytes=None):
  """One logistic problem, R_ rungs on a geometric temperature ladder, adjacent pairs
  tested in alternating even / odd rounds (rung 0 saved when ``saving``)."""
  X, y = _data(N, D, "logistic", seed=42)
  Xd, yd = _dev(X), torch.as_tensor(y, device=DEV)
  temps = [20.0 ** (j / (R_ - 1)) for j in range(R_)]
  chains = [dict(X=Xd, y=yd, D=D, n=64, theta0=np.full(D, 0.05 * j, np.float32), key=(2, j),
                 dkey=(3, j), temperature=temps[j], save=False) for j in range(R_)]
  eng = fused.FusedGLM(chains, "logistic", "sgld", np.full(iters, 2e-3, np.float32),
                       save_slots=np.arange(iters), n_slots=iters, target_bytes=target_bytes)
  pairs = [(j, j + 1) for j in range(R_ - 1)]
  ex = fused.ReplicaExchange(eng, pairs, [(temps[j], temps[j + 1]) for j in range(R_ - 1)],
                             keys=[(5, j) for j in range(R_ - 1)], swap_every=swap_every,
                             parity=[j % 2 for j in range(R_ - 1)],
                             save=[saving and j == 0 for j in range(R_ - 1)])
  return eng, ex


def test_exchange_ladder_rounds_and_permutation():
  """Each exchange launch tests exactly the pairs of the round's parity, every tested
  pair either keeps or exchanges its two states (theta, keys, U), and rung 0 is saved
  after the exchange."""
  print("exchange ladder", flush=True)
  eng, ex = _ladder()
  for it in range(12):
    eng._launch_steps()
    th = torch.stack(eng.theta()).clone()
    keys = eng.keys.clone()
    U = eng.U.clone()
    step0 = ex.step.clone()
    swaps0 = ex.swaps.clone()
    ex.launch()
    torch.cuda.synchronize()
    th1 = torch.stack(eng.theta())
    done = it + 1
    tested = set()
    if done % ex.swap_every == 0:
      par = (done // ex.swap_every - 1) % 2
      tested = {p for p in range(ex.P) if p % 2 == par}
    for p, (lo, hi) in enumerate(ex.pairs):
      dstep = int(ex.step[p] - step0[p])
      assert dstep == (1 if p in tested else 0)
      sw = int(ex.swaps[p] - swaps0[p])
      if sw:
        assert torch.equal(th1[lo], th[hi]) and torch.equal(th1[hi], th[lo])
        assert torch.equal(eng.keys[lo], keys[hi]) and torch.equal(eng.U[lo], U[hi])
    moved = {c for p in range(ex.P) if int(ex.swaps[p] - swaps0[p]) for c in ex.pairs[p]}
    for c in range(eng.C):
      if c not in moved:
        assert torch.equal(th1[c], th[c])
    np.testing.assert_array_equal(ex.sample_trace(0)[it].cpu().numpy(), th1[0].cpu().numpy())
  # rounds after iterations 3, 6, 9, 12 alternate parity 0, 1, 0, 1: pairs 0 and 2 are
  # tested twice, pair 1 twice
  assert ex.step.tolist() == [2, 2, 2]


def test_exchange_ladder_graph_equals_eager():
  print("exchange ladder graph", flush=True)
  a, exa = _ladder()
  b, exb = _ladder()
  a.run(12, graph_iterations=6)
  b.run(12)
  torch.cuda.synchronize()
  assert torch.equal(a.theta_flat, b.theta_flat)
  assert torch.equal(exa.swaps, exb.swaps) and torch.equal(exa.ssq, exb.ssq)
  assert torch.equal(exa.samples, exb.samples)


def _same_ladder(a, exa, b, exb):
  assert a.iteration_count() == b.iteration_count()
  assert torch.equal(a.theta_flat, b.theta_flat) and torch.equal(a.keys, b.keys)
  assert torch.equal(a.dkeys, b.dkeys) and torch.equal(a.U, b.U)
  assert torch.equal(exa.swaps, exb.swaps) and torch.equal(exa.ssq, exb.ssq)
  assert torch.equal(exa.step, exb.step) and torch.equal(exa.keys, exb.keys)
  assert torch.equal(a.samples, b.samples)


@pytest.mark.parametrize("how,iters,swap_every,target_bytes", [
    ("eager", 12, 3, None), ("graph", 24, 3, None), ("graph", 14, 4, None),
    ("loop", 12, 3, 1 << 30), ("loop-graph", 24, 4, 1 << 30), ("eager", 20, 3, None),
    ("loop", 20, 4, 1 << 30)])
def test_sparse_exchange_equals_launch_after_every_iteration(how, iters, swap_every,
                                                             target_bytes):
  """Without saving pairs the swap kernel is launched only after its test iterations
  (eagerly, from graphs, between loop launches); the chains, exchange statistics and
  keys equal those of a launch after every iteration. The schedule has 12 entries, so
  runs of 20 iterations cover the capped iteration count (12 % 3 == 0: a test after
  every iteration past the schedule; 12 % 4 == 0 too; 20 / 4 with table 12)."""
  print("sparse exchange", how, iters, swap_every, flush=True)
  a, exa = _ladder(saving=False, swap_every=swap_every, target_bytes=target_bytes)
  b, exb = _ladder(saving=False, swap_every=swap_every, target_bytes=target_bytes)
  exb.sparse = False  # the reference: launch after every iteration
  assert exa.sparse and a.table_len == 12
  if how.startswith("loop"):
    assert a.loop_capable() and not b.loop_capable()
  if how == "eager":
    a.run(iters, loop=False)
  elif how == "graph":
    a.run(iters, graph_iterations=6, loop=False)
  elif how == "loop":
    a.run(iters, loop=True)
  else:
    a.run(iters, graph_iterations=8, loop=True)
  b.run(iters)
  torch.cuda.synchronize()
  assert int(exa.step.sum()) > 0
  _same_ladder(a, exa, b, exb)


def test_table_group_autotune_restores_state():
  """TableGroup.autotune on one table of a group (the other, loop-capable, timed as it
  runs): afterwards the group runs exactly as a fresh group built with the chosen
  targets (chains, exchange statistics, keys)."""
  print("table group autotune", flush=True)
  a = [_ladder(saving=False, R_=4), _ladder(saving=False, R_=5, D=40, N=900)]
  b = [_ladder(saving=False, R_=4), _ladder(saving=False, R_=5, D=40, N=900)]
  ga = fused.TableGroup([e for e, _ in a])
  tuned = ga.autotune(candidates=[4 << 10, 64 << 10], reps=3, tables=[0])
  assert tuned["tables"] == [0] and len(tuned["target_bytes"]) == 1
  e_b = b[0][0]
  e_b.target_bytes = a[0][0].target_bytes
  e_b._build_table()
  assert list(e_b.nblocks_chain) == list(a[0][0].nblocks_chain)
  gb = fused.TableGroup([e for e, _ in b])
  ga.run(12, graph_iterations=6)
  gb.run(12, graph_iterations=6)
  torch.cuda.synchronize()
  for (ea, xa), (eb, xb) in zip(a, b):
    _same_ladder(ea, xa, eb, xb)


def test_sparse_exchange_graph_recaptured_after_eager_iterations():
  """A graph recorded at one position relative to the test iterations is not replayed
  at another: EngineGraph.replay raises, run() captures again."""
  print("sparse exchange graph phase", flush=True)
  a, exa = _ladder(saving=False, swap_every=3)
  b, exb = _ladder(saving=False, swap_every=3)
  exb.sparse = False
  g = a.capture(3, loop=False)
  a.run(1, loop=False)
  with pytest.raises(RuntimeError, match="another position"):
    g.replay()
  a.run(9, graph_iterations=3, loop=False)
  b.run(10)
  torch.cuda.synchronize()
  _same_ladder(a, exa, b, exb)


_LOOP_SHAPES = {
    # mixed D classes: the generic loop body
    "mixed": [(500, 8, 32), (800, 33, 64), (300, 100, 17)],
    # D <= 16, D <= 32 and D <= 64 tables: SGLD chains run resident (minibatches up to
    # 256 rows; the 300-row chains take the generic body in the same launch)
    "d16": [(500, 8, 32), (700, 16, 256), (300, 3, 17)],
    "d32": [(600, 17, 128), (700, 32, 256), (900, 24, 300), (400, 29, 40)],
    "d64": [(800, 33, 64), (600, 64, 200), (900, 40, 300)],
}


@pytest.mark.parametrize("mode,shapes,family", [
    ("sgld", "mixed", "logistic"), ("sgld_rms", "mixed", "logistic"),
    ("sghmc", "mixed", "logistic"), ("sgld", "d16", "logistic"),
    ("sgld_rms", "d16", "linear_gaussian"), ("sgld", "d32", "linear_gaussian"),
    ("sgld_rms", "d32", "logistic"), ("sghmc", "d32", "linear_gaussian"),
    ("sgld", "d64", "linear_gaussian"),
    ("sgld_rms", "d64", "logistic"), ("sghmc", "d64", "logistic")])
def test_loop_launch_equals_per_evaluation_launches(mode, shapes, family):
  """Single-block chains: many iterations in one loop launch give bit for bit the
  chains of one launch per evaluation (positions, rms_prop state, saved samples and
  potentials, U, var(lik), keys, iteration counter)."""
  print("loop launch", mode, shapes, family, flush=True)
  rng = np.random.default_rng(12)
  tables = []
  for _ in range(2):
    chains = []
    for c, (N, D, n) in enumerate(_LOOP_SHAPES[shapes]):
      X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32) if not tables else \
          tables[0][c][0]
      if family == "logistic":
        y = (rng.random(N) < 0.5).astype(np.float32) if not tables else tables[0][c][1]
      else:
        y = rng.standard_normal(N).astype(np.float32) if not tables else tables[0][c][1]
      chains.append((X, y, D, n))
    tables.append(chains)
  its = 23
  eps = np.linspace(0.01, 0.002, its).astype(np.float32)
  engines = []
  for _ in range(2):
    spec = []
    for c, (X, y, D, n) in enumerate(tables[0]):
      Xp = np.zeros((X.shape[0], (D + 3) // 4 * 4), np.float32)
      Xp[:, :D] = X
      spec.append(dict(X=torch.as_tensor(Xp, device=DEV), y=torch.as_tensor(y, device=DEV), D=D,
                       n=n,
                       theta0=np.zeros(D, np.float32), key=(1, c), dkey=(2, c)))
    kw = dict(sigma=0.7, prior_scale=2.0) if family == "linear_gaussian" else {}
    eng = fused.FusedGLM(spec, family, mode, eps, steps=3, friction=0.4,
                      save_slots=np.arange(its, dtype=np.int32), n_slots=its,
                      target_bytes=1 << 30, **kw)
    assert eng.loop_capable()
    engines.append(eng)
  engines[0].run(its, loop=False)
  engines[1].run(its, loop=True)
  torch.cuda.synchronize()
  a, b = engines
  assert a.iteration_count() == b.iteration_count() == its
  assert torch.isfinite(a.theta_flat).all()
  assert torch.equal(a.theta_flat, b.theta_flat)
  assert torch.equal(a.U, b.U) and torch.equal(a.var, b.var)
  if a.v_flat is not None:
    assert torch.equal(a.v_flat, b.v_flat)
  for c in range(3):
    np.testing.assert_array_equal(a.sample_trace(c).cpu().numpy(), b.sample_trace(c).cpu().numpy())
    np.testing.assert_array_equal(a.sample_potentials(c).cpu().numpy(),
                                  b.sample_potentials(c).cpu().numpy())
  np.testing.assert_array_equal(a.keys.cpu().numpy(), b.keys.cpu().numpy())
  np.testing.assert_array_equal(a.dkeys.cpu().numpy(), b.dkeys.cpu().numpy())


@pytest.mark.parametrize("shapes", ["d16", "d32", "d64"])
def test_short_loop_launches_equal_per_evaluation_launches(shapes):
  """Loop launches of 1, 2, 3 and 4 iterations (the key pipeline's start and end
  cases) give the chains, keys and saved samples of one launch per evaluation."""
  print("short loop launches", shapes, flush=True)
  rng = np.random.default_rng(5)
  its = 10
  eps = np.linspace(0.01, 0.004, its).astype(np.float32)
  data = []
  for c, (N, D, n) in enumerate(_LOOP_SHAPES[shapes]):
    X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
    y = (rng.random(N) < 0.5).astype(np.float32)
    data.append((X, y, D, n))
  engines = []
  for _ in range(2):
    spec = []
    for c, (X, y, D, n) in enumerate(data):
      Xp = np.zeros((X.shape[0], (D + 3) // 4 * 4), np.float32)
      Xp[:, :D] = X
      spec.append(dict(X=torch.as_tensor(Xp, device=DEV), y=torch.as_tensor(y, device=DEV), D=D,
                       n=n, theta0=np.zeros(D, np.float32), key=(3, c), dkey=(4, c)))
    eng = fused.FusedGLM(spec, "logistic", "sgld", eps, save_slots=np.arange(its, dtype=np.int32),
                         n_slots=its, target_bytes=1 << 30)
    assert eng.loop_capable()
    engines.append(eng)
  a, b = engines
  a.run(its, loop=False)
  for k in (1, 2, 3, 4):
    b.run(k, loop=True)
  torch.cuda.synchronize()
  assert a.iteration_count() == b.iteration_count() == its
  assert torch.equal(a.theta_flat, b.theta_flat) and torch.equal(a.U, b.U)
  for c in range(len(data)):
    np.testing.assert_array_equal(a.sample_trace(c).cpu().numpy(), b.sample_trace(c).cpu().numpy())
  np.testing.assert_array_equal(a.keys.cpu().numpy(), b.keys.cpu().numpy())
  np.testing.assert_array_equal(a.dkeys.cpu().numpy(), b.dkeys.cpu().numpy())


@pytest.mark.parametrize("tables,graph", [(2, 0), (3, 4)])
def test_table_group_equals_one_table(tables, graph):
  """Chains split over tables on separate streams give bit for bit the single table's
  positions, momenta and saved samples (same bytes-per-block target, so every chain
  keeps its block count and reduction order); graph replays included."""
  print("table group", tables, "graph", graph, flush=True)
  its = 8
  eps = np.full(its, 2e-3, np.float32)
  sizes = [(3000, 40, 256), (5000, 128, 512), (2500, 12, 64), (4000, 256, 300),
           (3000, 40, 128), (6000, 64, 700)]
  chains = []
  for c, (N, D, n) in enumerate(sizes):
    X, y = _data(N, D, "logistic", 20 + c)
    chains.append(dict(X=_dev(X), y=torch.as_tensor(y, device=DEV), D=D, n=n,
                       theta0=np.full(D, 0.01 * c, np.float32), key=(3, c), dkey=(c, 5)))
  kw = dict(steps=3, friction=0.7, target_bytes=64 * 1024,
            save_slots=np.arange(its, dtype=np.int32), n_slots=its)
  one = fused.FusedGLM(chains, "logistic", "sghmc", eps, **kw)
  grp = fused.TableGroup.split(chains, tables, "logistic", "sghmc", eps, **kw)
  assert len(grp.engines) == tables and grp.C == len(chains)
  assert any(nb > 1 for nb in one.nblocks_chain)
  one.run(its)
  grp.run(its, graph_iterations=graph)
  torch.cuda.synchronize()
  assert grp.iteration_count() == one.iteration_count() == its
  for c in range(len(chains)):
    assert torch.equal(grp.theta(c), one.theta(c)), c
    assert torch.equal(grp.momentum(c), one.momentum(c)), c
    assert torch.equal(grp.sample_trace(c), one.sample_trace(c)), c
    assert torch.equal(grp.sample_potentials(c), one.sample_potentials(c)), c


def test_table_group_with_exchanges_equals_tables_alone():
  """Two ladder tables, each with its own replica exchange, advanced as a table group
  on two streams (graph replays) equal the same tables run one after the other."""
  print("table group with exchanges", flush=True)
  grouped = [_ladder(R_=4), _ladder(R_=5, D=40, N=900)]
  alone = [_ladder(R_=4), _ladder(R_=5, D=40, N=900)]
  grp = fused.TableGroup([e for e, _ in grouped])
  grp.run(12, graph_iterations=6)
  for e, _ in alone:
    e.run(12)
  torch.cuda.synchronize()
  for (ga, exa), (gb, exb) in zip(grouped, alone):
    assert torch.equal(ga.theta_flat, gb.theta_flat)
    assert torch.equal(exa.swaps, exb.swaps) and torch.equal(exa.ssq, exb.ssq)
    assert torch.equal(exa.samples, exb.samples)
    assert torch.equal(exa.step, exb.step) and int(exa.step.sum()) > 0
