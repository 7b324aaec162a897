"""The reference's potential unit tests (tests/test_potential.py) restated on this
package's potential module, on its generic torch path (CPU tensors, no kernels), plus
the reference's doctest values (docs/usage/potential.rst) through the host data
loader."""
import itertools
import math

import numpy as np
import pytest
import torch

from oracle import jaxrandom as jr

from jax_sgmc_amd import data as DT
from jax_sgmc_amd import likelihoods as LK
from jax_sgmc_amd import potential as PT


def info(obs, bs, dim):
  return DT.MiniBatchInformation(observation_count=obs, batch_size=bs, mask=torch.ones(dim))


def exp_data(obs, dim, seed=0):
  k1, k2 = jr.split(jr.PRNGKey(seed))
  # exponential draws from the oracle's uniforms (inputs only)
  u1 = jr.uniform(k1, (obs,))
  u2 = jr.uniform(k2, (obs, dim))
  return {"scale": torch.as_tensor(-np.log1p(-u1)), "power": torch.as_tensor(-np.log1p(-u2))}


def power_lik(sample, obs):
  return sample["scale"] * obs["scale"] * torch.sum(torch.pow(sample["base"], obs["power"]))


@pytest.mark.parametrize("obs,dim", itertools.product([7, 11], [3, 5]))
def test_linear_potential(obs, dim):
  lik = lambda s, o: torch.sum(s * o)
  prior = lambda s: torch.sum(s)
  data = torch.tile(torch.arange(4.0), (dim, 1))
  for strat in ("map", "vmap"):
    pot = PT.minibatch_potential(prior, lik, strategy=strat)
    U, _ = pot(torch.ones(4), (data, info(obs, dim, dim)))
    assert math.isclose(U.item(), -np.sum(np.arange(4)) * obs - 4, rel_tol=1e-5)


@pytest.mark.parametrize("obs,dim", itertools.product([7, 11], [3, 5]))
def test_stochastic_potential_zero_and_zero_gradient(obs, dim):
  data = exp_data(obs, dim)
  sample = {"scale": torch.tensor(0.5), "base": torch.zeros(dim)}
  for strat in ("map", "vmap"):
    pot = PT.minibatch_potential(lambda _: 0.0, power_lik, strategy=strat)
    U, _ = pot(sample, (data, info(obs, obs, dim)))
    assert U.item() == 0.0
  data0 = dict(data, scale=torch.zeros(obs))
  pot = PT.minibatch_potential(lambda _: 0.0, power_lik, strategy="vmap")
  sample = {"scale": torch.tensor(0.5), "base": torch.rand(dim)}
  (_, _), g = pot.value_and_grad(sample, (data0, info(obs, obs, dim)))
  assert torch.all(g == 0)


@pytest.mark.parametrize("obs,dim", itertools.product([7, 11], [3, 5]))
def test_map_equals_vmap_value_and_gradient(obs, dim):
  data = exp_data(obs, dim, seed=1)
  prior = lambda s: torch.exp(-s["scale"])
  sample = {"scale": torch.tensor(0.5), "base": torch.rand(dim, generator=torch.Generator().manual_seed(0))}
  a = PT.minibatch_potential(prior, power_lik, strategy="map")
  b = PT.minibatch_potential(prior, power_lik, strategy="vmap")
  (Ua, _), ga = a.value_and_grad(sample, (data, info(obs, obs, dim)))
  (Ub, _), gb = b.value_and_grad(sample, (data, info(obs, obs, dim)))
  torch.testing.assert_close(Ua, Ub, rtol=1e-5, atol=0)
  torch.testing.assert_close(ga, gb, rtol=1e-5, atol=1e-6)
  # gradient against the analytic derivative of the same potential
  base = sample["base"]
  N = obs
  dscale = -N / obs * torch.sum(data["scale"] * torch.sum(torch.pow(base, data["power"]), 1)) \
      + torch.exp(-sample["scale"])
  torch.testing.assert_close(ga[dim], dscale, rtol=1e-5, atol=1e-6)  # sorted keys: base, scale


@pytest.mark.parametrize("obs,dim", itertools.product([7, 11], [3, 5]))
def test_stateful_potential_returns_first_row_state(obs, dim):
  data = exp_data(obs, dim)

  def lik(state, sample, obs_):
    scale = sample["scale"] * state["scale"]
    base = sample["base"] + state["base"]
    return scale * obs_["scale"] * torch.sum(torch.pow(base, obs_["power"])), sample

  sample = {"scale": torch.tensor([0.5]), "base": torch.ones(dim)}
  st = {"scale": torch.tensor([0.0]), "base": torch.zeros(dim)}
  for strat in ("map", "vmap"):
    pot = PT.minibatch_potential(lambda _: 0.0, lik, strategy=strat, has_state=True)
    _, new = pot(sample, (data, info(obs, obs, dim)), state=st)
    torch.testing.assert_close(new["scale"], sample["scale"])
    torch.testing.assert_close(new["base"], sample["base"])


@pytest.mark.parametrize("obs,dim,mb", itertools.product([7, 11], [3, 5], [2, 3]))
def test_full_potential_equals_minibatch_over_all_rows(obs, dim, mb):
  data = exp_data(obs, dim)
  prior = lambda s: torch.exp(-s["scale"][0])
  sample = {"scale": torch.tensor([0.5]), "base": torch.ones(dim)}
  ref, _ = PT.minibatch_potential(prior, power_lik, strategy="map")(
      sample, (data, info(obs, obs, mb)))
  loader = DT.NumpyDataLoader(device="cpu", **{k: v.numpy() for k, v in data.items()})
  init, fmap, _ = DT.full_reference_data(loader, cached_batches_count=2, mb_size=mb)
  for strat in ("map", "vmap"):
    full = PT.full_potential(prior, power_lik, strategy=strat)
    U, _ = full(sample, init(), fmap)
    torch.testing.assert_close(U, ref, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("obs,dim", itertools.product([7, 11], [3, 5]))
def test_variance_of_likelihoods(obs, dim):
  data = exp_data(obs, dim)
  sample = {"scale": torch.tensor([0.5]), "base": torch.ones(dim)}
  pot = PT.minibatch_potential(lambda _: 0.0, power_lik, strategy="map")
  per_row = [pot(sample, ({k: v[i:i + 1] for k, v in data.items()}, info(1, 1, 1)))[0]
             for i in range(obs)]
  true_var = torch.var(torch.stack(per_row), unbiased=False)
  _, (lks, _) = pot(sample, (data, info(obs, obs, obs)), likelihoods=True)
  torch.testing.assert_close(torch.var(lks, unbiased=False), true_var, rtol=1e-5, atol=1e-6)


def test_unknown_strategy_raises():
  with pytest.raises(NotImplementedError):
    PT.minibatch_potential(lambda _: 0.0, power_lik, strategy="scan")


def _norm_logpdf(x, loc, scale):
  return (torch.log(2 * math.pi * scale ** 2) + (x - loc) ** 2 / scale ** 2) / -2


def test_reference_doctest_values_through_host_loader():
  """docs/usage/potential.rst:27-118, 131-179: 838, var 7, 707."""
  mean = jr.normal(jr.PRNGKey(0), (100, 5))
  loader = DT.NumpyDataLoader(device="cpu", mean=mean)
  sample = {"mean": torch.zeros(5), "std": torch.ones(1)}
  b_init, b_get, _ = DT.random_reference_data(loader, cached_batches_count=50, mb_size=5)
  f_init, f_map, _ = DT.full_reference_data(loader, cached_batches_count=50, mb_size=3)
  rstate = b_init()
  fstate = f_init()

  def lik(s, obs):
    return torch.sum(_norm_logpdf(obs["mean"], s["mean"], s["std"]))

  prior = lambda _: 0.0
  pot = PT.minibatch_potential(prior, lik, strategy="map")
  _, batch = b_get(rstate, information=True)
  U, _ = pot(sample, batch)
  assert round(U.item()) == 838
  _, (lks, _) = pot(sample, batch, likelihoods=True)
  assert round(torch.var(lks, unbiased=False).item()) == 7
  U, _ = PT.full_potential(prior, lik, strategy="vmap")(sample, fstate, f_map)
  assert round(U.item()) == 707


def test_random_reference_data_argument_checks():
  loader = DT.NumpyDataLoader(device="cpu", x=np.zeros(5))
  with pytest.raises(ValueError):
    DT.random_reference_data(loader, 1, 6)
  with pytest.raises(ValueError):
    DT.random_reference_data(loader, 0, 2)


def test_adaptive_step_size_scheduler_matches_oracle():
  """scheduler.adaptive_step_size driven through init_scheduler with the solver's
  acceptance ratios reproduces the oracle's dual averaging (scheduler.py:376-444)."""
  from oracle import steps as ost
  from jax_sgmc_amd import scheduler as SC
  rng = np.random.default_rng(9)
  acc = rng.random(60).astype(np.float32)
  init, nxt, get = SC.init_scheduler(
      step_size=SC.adaptive_step_size(burn_in=40, initial_step_size=0.03,
                                      stabilization_constant=10, decay_constant=0.75,
                                      speed_constant=0.05, target_acceptance_rate=0.25),
      burn_in=SC.initial_burn_in(40))
  st, _ = init(60)
  got = []
  for a in acc:
    got.append(get(st).step_size)
    st = nxt(st, acceptance_ratio=a)
  ref = ost.adaptive_step_size(acc, 40, 0.03, 10, 0.75, 0.05, 0.25)
  np.testing.assert_array_equal(np.asarray(got, np.float32), ref)
  import pytest
  with pytest.raises(NotImplementedError):
    init.tables(st)
