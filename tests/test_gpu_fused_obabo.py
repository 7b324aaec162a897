"""OBABO and SGGMC for GLM families on the fused multi-chain engine (integrator.py:138-346,
solver.py:434-567, alias.py:330-449, 542-619): trajectories, kinetic energies and
accept/reject decisions against the oracle, fused chains against the same chains run
one by one through the generic kernels, and the loop launch and graph replays against
eager launches."""
import numpy as np
import pytest
import torch

from oracle import model as om
from oracle import steps as ost

from jax_sgmc_amd import alias
from jax_sgmc_amd import data as DT
from jax_sgmc_amd import fused as FU
from jax_sgmc_amd import likelihoods as LK
from jax_sgmc_amd import potential as PT
from jax_sgmc_amd import random as R
from jax_sgmc_amd import scheduler as SC

pytestmark = pytest.mark.gpu
DEV = "cuda"


def logreg(N=3000, D=24, seed=1):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  return X, y


def linreg(N=2000, D=10, seed=2, sigma=0.5):
  rng = np.random.default_rng(seed)
  X = rng.standard_normal((N, D)).astype(np.float32)
  w = rng.standard_normal(D)
  y = (X @ w + sigma * rng.standard_normal(N)).astype(np.float32)
  return X, y


def _within(got, ref, what, rtol=1e-5):
  got = np.asarray(got, np.float64)
  ref = np.asarray(ref, np.float64)
  tol = rtol * np.abs(ref) + rtol * np.max(np.abs(ref))
  err = np.abs(got - ref)
  assert np.all(err <= tol), f"{what}: max err {err.max():.3g}, max tol {tol.max():.3g}"


@pytest.mark.parametrize("family,mass", [("logistic", False), ("linear_gaussian", False),
                                         ("logistic", True)])
def test_fused_obabo_state_matches_oracle(family, mass):
  """Positions, momenta, potentials and both kinetic-energy accumulators of two chains
  with their own keys after every iteration, against the oracle's obabo (with a shared
  diagonal mass: M^{1/2} in both O steps, M^-1 in the A step and the energies)."""
  print("fused obabo vs oracle", family, "mass", mass, flush=True)
  if family == "logistic":
    X, y = logreg()
    fam = om.Family("logistic", X, y)
  else:
    X, y = linreg(D=11)  # odd D: the last noise pair holds one element
    fam = om.Family("linear_gaussian", X, y, sigma=0.5)
  N, D = X.shape
  its, steps, fr, n = 10, 3, 1.5, 128
  eps = SC.polynomial_table(4e-3, 1e-3, its)
  Xd = torch.as_tensor(np.pad(X, ((0, 0), (0, (-D) % 4))), device=DEV)
  yd = torch.as_tensor(y, device=DEV)
  keys = [(0, 5), (7, 1)]
  dkeys = [(3, 0), (0, 9)]
  chains = [dict(X=Xd, y=yd, D=D, n=n, theta0=np.full(D, 0.1 * c, np.float32), key=keys[c],
                 dkey=dkeys[c]) for c in range(2)]
  inv_m = sqrt_m = None
  mk = {}
  if mass:
    inv = np.linspace(0.5, 2.0, D).astype(np.float32)
    sq = (1.0 / np.sqrt(inv)).astype(np.float32)
    inv_m = torch.as_tensor(inv, device=DEV)
    sqrt_m = torch.as_tensor(sq, device=DEV)
    mk = dict(inv_mass=inv.astype(np.float64), sqrt_mass=sq.astype(np.float64))
  eng = FU.FusedGLM(chains, family, "obabo", eps, steps=steps, friction=fr, sigma=0.5,
                    inv_mass=inv_m, sqrt_mass=sqrt_m)
  refs = [ost.obabo(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0),
                    np.full(D, 0.1 * c), its, eps, steps, fr, N, n, [D],
                    key=np.asarray(keys[c], np.uint32), dkey=np.asarray(dkeys[c], np.uint32),
                    **mk)
          for c in range(2)]
  for t in range(its):
    eng.iteration()
    torch.cuda.synchronize()
    for c in range(2):
      r = refs[c]
      _within(eng.theta(c).cpu().numpy(), r["theta"][t], f"theta c{c} t{t}")
      _within(eng.momentum(c).cpu().numpy(), r["p"][t], f"p c{c} t{t}", rtol=2e-5)
      _within([eng.U[c].item()], [r["U"][t]], f"U c{c} t{t}")
      ks, ke = eng.kinetic_energies(c).cpu().numpy()
      _within([ks, ke], [r["ke_start"][t], r["ke_end"][t]], f"KE c{c} t{t}")
  for c in range(2):
    assert tuple(eng.keys[c].cpu().numpy().view(np.uint32)) == tuple(refs[c]["key"])
    assert tuple(eng.dkeys[c].cpu().numpy().view(np.uint32)) == tuple(refs[c]["dkey"])


def test_fused_obabo_chains_match_generic_path():
  """alias.obabo with four chains and their own keys: the fused engine against the
  same chains run one by one through the generic OBABO kernels."""
  print("fused obabo vs generic, 4 chains", flush=True)
  X, y = logreg(N=2500, D=40, seed=5)
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  run = alias.obabo(pot, loader, cache_size=1, batch_size=100, integration_steps=4,
                    friction=0.8, first_step_size=5e-3, last_step_size=1e-3,
                    accepted_samples=10)
  keys = [R.PRNGKey(s, DEV) for s in (1, 2, 3, 4)]
  dkeys = [R.PRNGKey(10 + s, DEV) for s in range(4)]
  inits = [0.1 * np.random.default_rng(s).standard_normal(40).astype(np.float32)
           for s in range(4)]
  a = run(*[torch.as_tensor(v) for v in inits], iterations=16, keys=keys, data_keys=dkeys,
          fused=True)
  b = run(*[torch.as_tensor(v) for v in inits], iterations=16, keys=keys, data_keys=dkeys,
          fused=False)
  for c in range(4):
    sa, sb = a[c]["samples"], b[c]["samples"]
    assert a[c]["sample_count"] == b[c]["sample_count"] == 10
    _within(sa["variables"], sb["variables"], f"variables c{c}")
    _within(sa["energy"], sb["energy"], f"energy c{c}")


def test_fused_obabo_loop_and_graph_replay_match_eager():
  """Single-block chains: one loop launch of all iterations and graph replays of two
  iterations give the eager launches' samples bit for bit."""
  print("fused obabo loop / graph", flush=True)
  X, y = linreg(N=1500, D=12, seed=4)
  Xd = torch.as_tensor(np.pad(X, ((0, 0), (0, 4))), device=DEV)
  yd = torch.as_tensor(y, device=DEV)
  its = 9
  eps = SC.polynomial_table(3e-3, 1e-3, its)

  def make():
    chains = [dict(X=Xd, y=yd, D=12, n=64, theta0=np.zeros(12, np.float32), key=(1, c),
                   dkey=(c, 2), temperature=1.0 + c) for c in range(3)]
    return FU.FusedGLM(chains, "linear_gaussian", "obabo", eps, steps=2, friction=0.9,
                       sigma=0.5, save_slots=np.arange(its, dtype=np.int32), n_slots=its)

  e0, e1, e2 = make(), make(), make()
  assert e0.loop_capable()
  e0.run(its, loop=False)
  e1.run(its, loop=True)
  e2.run(its, graph_iterations=2, loop=False)
  torch.cuda.synchronize()
  assert e0.iteration_count() == e1.iteration_count() == e2.iteration_count() == its
  for c in range(3):
    ref = e0.sample_trace(c).cpu().numpy()
    np.testing.assert_array_equal(e1.sample_trace(c).cpu().numpy(), ref)
    np.testing.assert_array_equal(e2.sample_trace(c).cpu().numpy(), ref)
    np.testing.assert_array_equal(e1.sample_potentials(c).cpu().numpy(),
                                  e0.sample_potentials(c).cpu().numpy())
    np.testing.assert_array_equal(e1.kinetic_energies(c).cpu().numpy(),
                                  e0.kinetic_energies(c).cpu().numpy())


def _sggmc_alias(X, y, first=6e-3, last=4e-3, steps=3, fr=1.0, n=128):
  # distinct first and last step sizes: with first == last the polynomial schedule's
  # coefficients divide by zero (scheduler.py:512-519) and every step size is NaN
  loader = DT.DeviceNumpyDataLoader(x=X, y=y)
  pot = PT.minibatch_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  full = PT.full_potential(LK.GaussianPrior(1.0), LK.Logistic(), strategy="vmap")
  return alias.sggmc(pot, full, loader, cache_size=1, batch_size=n, integration_steps=steps,
                     friction_coefficient=fr, first_step_size=first, last_step_size=last)


@pytest.mark.parametrize("fused", [True, False])
def test_sggmc_matches_oracle(fused):
  """alias.sggmc (fused GLM engine or generic kernels) against the oracle's sggmc:
  accept decisions, acceptance ratios, positions and the kept full potential."""
  print("sggmc vs oracle fused", fused, flush=True)
  X, y = logreg()
  its = 20
  res = _sggmc_alias(X, y)(torch.zeros(24), iterations=its, fused=fused)[0]["samples"]
  fam = om.Family("logistic", X, y)
  eps = SC.polynomial_table(6e-3, 4e-3, its)
  r = ost.sggmc(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0),
                lambda th: om.full_potential(fam, th, 1.0, 3000), np.zeros(24), its, eps, 3,
                1.0, 3000, 128, [24])
  acc = r["accept"]
  assert acc.any() and (~acc).any(), "the case should exercise both branches"
  ratio = np.asarray(res["acceptance_ratio"], np.float64).reshape(-1)
  np.testing.assert_allclose(ratio, np.exp(r["log_alpha"]), rtol=1e-3, atol=1e-4)
  _within(res["variables"], r["theta"], "theta")
  np.testing.assert_allclose(np.asarray(res["potential"], np.float64).reshape(-1),
                             r["U_full"], rtol=1e-6)


def test_fused_sggmc_chains_match_generic_path():
  """Four SGGMC chains with their own keys in one fused table equal the same chains run
  one by one through the generic solver (ratios, positions, energies, kinetic
  energies, kept full potentials)."""
  print("fused sggmc vs generic, 4 chains", flush=True)
  X, y = logreg(N=2000, D=32, seed=9)
  its = 14
  run = _sggmc_alias(X, y, first=5e-3, last=3e-3, n=100)
  keys = [R.PRNGKey(s, DEV) for s in (3, 4, 5, 6)]
  inits = [0.05 * np.random.default_rng(s).standard_normal(32).astype(np.float32)
           for s in range(4)]
  a = run(*[torch.as_tensor(v) for v in inits], iterations=its, keys=keys, fused=True)
  b = run(*[torch.as_tensor(v) for v in inits], iterations=its, keys=keys, fused=False)
  for c in range(4):
    sa, sb = a[c]["samples"], b[c]["samples"]
    ra = np.asarray(sa["acceptance_ratio"], np.float64).reshape(-1)
    rb = np.asarray(sb["acceptance_ratio"], np.float64).reshape(-1)
    np.testing.assert_allclose(ra, rb, rtol=1e-3, atol=1e-4)
    _within(sa["variables"], sb["variables"], f"variables c{c}")
    _within(sa["energy"], np.asarray(sb["energy"]).reshape(-1), f"energy c{c}", rtol=1e-4)
    np.testing.assert_allclose(np.asarray(sa["kinetic_energy"], np.float64),
                               np.asarray(sb["kinetic_energy"], np.float64).reshape(-1),
                               rtol=1e-3, atol=1e-3)
    np.testing.assert_allclose(np.asarray(sa["potential"], np.float64),
                               np.asarray(sb["potential"], np.float64).reshape(-1), rtol=1e-6)


def test_fused_sggmc_graph_replay_matches_eager():
  print("fused sggmc graph replay", flush=True)
  X, y = logreg(N=3000, D=64, seed=8)
  Xd = torch.as_tensor(X, device=DEV)
  yd = torch.as_tensor(y, device=DEV)
  its = 8
  eps = np.full(its, 4e-3, np.float32)

  def make():
    chains = [dict(X=Xd, y=yd, D=64, n=256, theta0=np.zeros(64, np.float32), dkey=(0, c))
              for c in range(3)]
    eng = FU.FusedGLM(chains, "logistic", "sggmc", eps, steps=3, friction=1.0,
                      save_slots=np.arange(its, dtype=np.int32), n_slots=its)
    FU.SggmcMH(eng, keys=[(0, 10 + c) for c in range(3)])
    return eng

  e1, e2 = make(), make()
  e1.run(its)
  e2.run(its, graph_iterations=4)
  torch.cuda.synchronize()
  assert e1.iteration_count() == e2.iteration_count() == its
  for c in range(3):
    np.testing.assert_array_equal(e1.sample_trace(c).cpu().numpy(),
                                  e2.sample_trace(c).cpu().numpy())
    np.testing.assert_array_equal(e1.mh.acceptance_ratios(c).cpu().numpy(),
                                  e2.mh.acceptance_ratios(c).cpu().numpy())
    np.testing.assert_array_equal(e1.mh.full_potentials(c).cpu().numpy(),
                                  e2.mh.full_potentials(c).cpu().numpy())
