"""Solvers: the iteration driver and the solver wrappers (solver.py of the reference).

- mcmc(solver, scheduler, strategy='map', saving=None) -> run(*states, schedulers,
  iterations) (solver.py:65-185): per chain, per iteration: schedule = get; (state,
  stats) = update(state, schedule); keep = burn_in and accept; save; next.
  Strategy 'map' runs the chains one after another, as the reference does;
  ``alias`` routes chains of a recognized GLM family through the fused multi-chain
  kernel instead.
- sgmc(integrator) (solver.py:188-217);
- parallel_tempering(integrator, sa_schedule) (solver.py:220-299): normal and hot
  chain with the reference's swap rule (the pair is kept when log u < log s and
  swapped otherwise);
- sggmc(integrator, full_potential_fn, full_data_map) (solver.py:434-577): OBABO
  proposal and a Metropolis-Hastings test on the full-data potential, accumulated in
  FP64;
- amagold(integrator, full_potential_fn, full_data_map) (solver.py:301-432):
  reversible-leapfrog proposal, the amortized test log a = min(0, U - U' + energy)
  on the FP64 full-data potential, momentum flipped on rejection.
Both MH solvers take ``mass_adaption`` (e.g. adaption.mass_matrix): the integrator
runs with the adapted mass and the strategy is updated with every post-test position.
"""
import math
from typing import Any, NamedTuple

import torch

from . import io as sgio
from . import random as R
from . import tree as T
from .tree import tree_map


def _on_device(sample):
  """The sample's leaves on its GPU (the device the integrators run on)."""
  dev = next((l.device for l in T.flatten(sample)[0]
              if isinstance(l, torch.Tensor) and l.is_cuda), torch.device("cuda"))
  return tree_map(lambda l: T.as_tensor(l, dev), sample)


class AMAGOLDState(NamedTuple):
  full_data_state: Any
  potential: Any
  integrator_state: Any
  key: Any
  acceptance_ratio: Any
  mass_state: Any


class SGGMCState(NamedTuple):
  full_data_state: Any
  potential: Any
  integrator_state: Any
  key: Any
  acceptance_ratio: Any
  mass_state: Any


def mcmc(solver, scheduler, strategy="map", saving=None, loading=None):
  if strategy not in ("map", "vmap", "pmap"):
    raise NotImplementedError(f"Strategy {strategy} is unknown. ")
  if loading is not None:
    raise NotImplementedError("Loading of checkpoints is currently not supported.")
  init_saving, save, postprocess = saving if saving is not None else sgio.no_save()
  sched_init, sched_next, sched_get = scheduler
  _, solver_update, solver_get = solver

  def run_chain(state, iterations, schedulers):
    main, static = sched_init(iterations)
    helpers = [sched_init(iterations, **kw)[0] for kw in (schedulers or [])]
    sstates = [main] + helpers
    saving_state = init_saving(solver_get(state), (state, sstates), static)
    for _ in range(int(iterations)):
      schedules = [sched_get(s) for s in sstates]
      state, stats = solver_update(state, *schedules)
      keep = bool(schedules[0].burn_in) and bool(schedules[0].accept)
      saving_state = save(saving_state, keep, solver_get(state))
      sstates = [sched_next(s, **(stats or {})) for s in sstates]
    return postprocess(saving_state)

  def run(*states, schedulers=None, iterations=100000):
    # 'vmap' and 'pmap' evaluate the same independent chains; on one GPU they run in
    # sequence like 'map'.
    return [run_chain(s, iterations, schedulers) for s in states]

  return run


def sgmc(integrator):
  init_i, update_i, get_i = integrator

  def init(*args, **kwargs):
    return init_i(*args, **kwargs)

  def update(state, schedule):
    return update_i(state, schedule), None

  def get(state):
    return get_i(state)

  return init, update, get


def parallel_tempering(integrator, sa_schedule=lambda n: 1 / n):
  init_i, update_i, get_i = integrator

  def init(normal_sample, tempered_sample, ssq_init=0.0, key=None, F=1.0, **kwargs):
    key = R.PRNGKey(0) if key is None else key
    k = R.split(key, 3)
    normal = init_i(normal_sample, key=k[1].clone(), **kwargs)
    hot = init_i(tempered_sample, key=k[2].clone(), **kwargs)
    return (normal, hot, torch.tensor(float(ssq_init), device=key.device), float(F), 0,
            k[0].clone())

  def update(state, normal_schedule, hot_schedule):
    normal, hot, ssq, F, step, key = state
    step += 1
    normal = update_i(normal, normal_schedule)
    hot = update_i(hot, hot_schedule)
    lr = torch.tensor(sa_schedule(step), dtype=torch.float32)
    ssq = (1 - lr.item()) * ssq + lr.item() * normal.variance
    temps = torch.tensor(1.0, dtype=torch.float32) / float(normal_schedule.temperature) - \
        torch.tensor(1.0, dtype=torch.float32) / float(hot_schedule.temperature)
    temps = temps.to(ssq.device)
    correction = temps * ssq / F
    log_s = temps * (normal.potential - hot.potential - correction)
    key, sub = R.split(key)
    log_u = torch.log(R.uniform(sub))
    if not bool((log_u < log_s).item()):
      normal, hot = hot, normal
    return (normal, hot, ssq, F, step, key.clone()), None

  def get(state):
    return get_i(state[0])

  return init, update, get


def sggmc(integrator_fn, full_potential_fn, full_data_map, mass_adaption=None):
  init_i, update_i, get_i = integrator_fn
  init_full, full_map, _ = full_data_map
  if mass_adaption:
    init_m, update_m, get_m = mass_adaption[:3]

  def init(init_sample, key=None, initial_mass=None, full_data_kwargs=None, **kwargs):
    fds = init_full(**(full_data_kwargs or {}))
    ims = kwargs.get("init_model_state")
    U, (fds, new_ms) = full_potential_fn(init_sample, fds, full_map, state=ims)
    kwargs["init_model_state"] = new_ms
    key = R.PRNGKey(0) if key is None else key
    k = R.split(key)
    ist = init_i(init_sample, key=k[0].clone(), **kwargs)
    mass_state = init_m(_on_device(init_sample), initial_mass) if mass_adaption else None
    z = torch.zeros((), device=k.device)
    return SGGMCState(full_data_state=fds, potential=U.to(torch.float64), integrator_state=ist,
                      key=k[1].clone(), acceptance_ratio=(z, z, z), mass_state=mass_state)

  def update(state, schedule):
    k = R.split(state.key)
    key, sub = k[0].clone(), k[1].clone()
    mass = get_m(state.mass_state) if mass_adaption else None
    prop = update_i(state.integrator_state, schedule, mass=mass)
    U_new, (fds, new_ms) = full_potential_fn(prop.positions, state.full_data_state, full_map,
                                             state=prop.model_state)
    U_new = U_new.to(torch.float64)
    d = (U_new - state.potential + prop.kinetic_energy_end.double()
         - prop.kinetic_energy_start.double())
    log_alpha = -1.0 / float(schedule.temperature) * d
    log_alpha = torch.where(log_alpha <= 0, log_alpha, torch.zeros_like(log_alpha))
    u = R.uniform(sub)
    accept = bool((torch.log(u.double()) < log_alpha).item())
    if accept:
      src, U_keep, ms = prop, U_new, new_ms
    else:
      src, U_keep, ms = state.integrator_state, state.potential, \
          state.integrator_state.model_state
    z = torch.zeros((), device=u.device)
    new_int = type(prop)(positions=src.positions, momentum=src.momentum, model_state=ms,
                         potential=src.potential, data_state=prop.data_state, key=prop.key,
                         kinetic_energy_start=z, kinetic_energy_end=z.clone())
    ratio = torch.exp(log_alpha)
    mass_state = update_m(state.mass_state, src.positions) if mass_adaption else None
    new = SGGMCState(full_data_state=fds, potential=U_keep, integrator_state=new_int, key=key,
                     acceptance_ratio=(ratio, torch.tensor(float(schedule.step_size)),
                                       prop.kinetic_energy_end - prop.kinetic_energy_start),
                     mass_state=mass_state)
    return new, {"acceptance_ratio": ratio}

  def get(state):
    d = get_i(state.integrator_state)
    d["acceptance_ratio"] = state.acceptance_ratio[0]
    d["step_size"] = state.acceptance_ratio[1]
    d["kinetic_energy"] = state.acceptance_ratio[2]
    d["potential"] = state.potential
    return d

  return init, update, get


def amagold(integrator_fn, full_potential_fn, full_data_map, mass_adaption=None):
  """solver.py:301-432. init splits the key into the integrator key and the MH key and
  evaluates the full potential at the initial sample; every update splits the MH key,
  integrates, evaluates the full potential at the proposal and accepts iff
  log u < min(0, U - U' + energy). A rejected proposal keeps the previous positions,
  potential and model state and flips the momentum; the integrator key and the data
  state come from the proposal either way."""
  init_i, update_i, get_i = integrator_fn
  init_full, full_map, _ = full_data_map
  if mass_adaption:
    init_m, update_m, get_m = mass_adaption[:3]

  def init(init_sample, key=None, initial_mass=None, full_data_kwargs=None, **kwargs):
    mass_state = init_m(_on_device(init_sample), initial_mass) if mass_adaption else None
    mass = get_m(mass_state) if mass_adaption else None
    fds = init_full(**(full_data_kwargs or {}))
    ims = kwargs.get("init_model_state")
    U, (fds, new_ms) = full_potential_fn(init_sample, fds, full_map, state=ims)
    kwargs["init_model_state"] = new_ms
    key = R.PRNGKey(0) if key is None else key
    k = R.split(key)
    ist = init_i(init_sample, key=k[0].clone(), mass=mass, **kwargs)
    z = torch.zeros((), device=k.device)
    return AMAGOLDState(full_data_state=fds, potential=U.to(torch.float64),
                        integrator_state=ist, key=k[1].clone(), acceptance_ratio=(z, z),
                        mass_state=mass_state)

  def update(state, schedule):
    k = R.split(state.key)
    key, sub = k[0].clone(), k[1].clone()
    mass = get_m(state.mass_state) if mass_adaption else None
    prop = update_i(state.integrator_state, schedule, mass=mass)
    U_new, (fds, new_ms) = full_potential_fn(prop.positions, state.full_data_state, full_map,
                                             state=prop.model_state)
    U_new = U_new.to(torch.float64)
    log_alpha = state.potential - U_new + prop.potential.double()
    log_alpha = torch.where(log_alpha > 0.0, torch.zeros_like(log_alpha), log_alpha)
    u = R.uniform(sub)
    accept = bool((torch.log(u.double()) < log_alpha).item())
    if accept:
      src, U_keep, ms, mom = prop, U_new, new_ms, prop.momentum
    else:
      old = state.integrator_state
      src, U_keep, ms = old, state.potential, old.model_state
      mom = tree_map(lambda x: -1.0 * x, old.momentum)
    new_int = type(prop)(positions=src.positions, momentum=mom, potential=src.potential,
                         model_state=ms, data_state=prop.data_state, key=prop.key,
                         extra_fields=prop.extra_fields)
    ratio = torch.exp(log_alpha)
    mass_state = update_m(state.mass_state, src.positions) if mass_adaption else None
    new = AMAGOLDState(full_data_state=fds, potential=U_keep, integrator_state=new_int,
                       key=key, acceptance_ratio=(ratio, torch.tensor(float(schedule.step_size))),
                       mass_state=mass_state)
    return new, {"acceptance_ratio": ratio}

  def get(state):
    d = get_i(state.integrator_state)
    d["acceptance_ratio"] = state.acceptance_ratio[0]
    d["step_size"] = state.acceptance_ratio[1]
    return d

  return init, update, get
