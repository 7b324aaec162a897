"""SG-MCMC integrators on the HIP step kernels (integrator.py of the reference).

Each constructor returns the reference's triplet:
- langevin_diffusion(potential_fn, batch_fn, adaption=None) -> (init_fn, update_fn,
  get_fn)  (integrator.py:767-924);
- friction_leapfrog(potential_fn, batch_fn, steps, friction, const_mass,
  noise_model) -> (init_fn, integrate, get_fn)  (integrator.py:563-765);
- obabo(potential_fn, batch_fn, steps, friction, const_mass) -> (init_fn, integrate,
  get_fn)  (integrator.py:138-346);
- reversible_leapfrog(potential_fn, batch_fn, steps, friction, const_mass) ->
  (init_fn, integrate, get_fn)  (integrator.py:349-560), AMAGOLD's proposal.
States are NamedTuples with the reference's fields; parameters are kept as one flat
float32 device vector (ravel order) plus the pytree structure, and every update
returns a new state (the input state is not modified).
"""
from typing import Any, NamedTuple

import numpy as np
import torch

from . import _lib
from . import random as R
from . import tree as T
from .adaption import MassMatrix, Tensor


class LeapfrogState(NamedTuple):
  positions: Any
  momentum: Any
  potential: Any
  model_state: Any
  data_state: Any
  key: Any
  extra_fields: Any = None


class ObaboState(NamedTuple):
  positions: Any
  momentum: Any
  potential: Any
  model_state: Any
  data_state: Any
  key: Any
  kinetic_energy_start: Any
  kinetic_energy_end: Any


class LangevinState(NamedTuple):
  latent_variables: Any
  model_state: Any
  key: Any
  adapt_state: Any
  data_state: Any
  potential: Any
  variance: Any


def _device_of(sample):
  leaves, _ = T.flatten(sample)
  for l in leaves:
    if isinstance(l, torch.Tensor) and l.is_cuda:
      return l.device
  return torch.device("cuda")


def _to_device_tree(sample, device):
  return T.tree_map(lambda l: T.as_tensor(l, device), sample)


def init_mass(mass):
  """integrator.py:99-116: diagonal mass -> (inverse, square root)."""
  inv = T.tree_map(lambda x: torch.pow(x, -1.0), mass)
  sq = T.tree_map(torch.sqrt, mass)
  return MassMatrix(inv=Tensor(1, inv), sqrt=Tensor(1, sq))


def random_tree(key, a):
  """integrator.py:119-135."""
  lay = T.Layout(a)
  return lay.unravel(R.random_tree(key, lay))


def _flat_mass(mass, lay, device):
  if mass is None:
    return None, None
  if getattr(mass.inv, "ndim", 1) == 2 or getattr(mass.sqrt, "ndim", 1) == 2:
    raise NotImplementedError("the step kernels take a diagonal mass")
  inv = lay.ravel(mass.inv.tensor, device)
  sq = lay.ravel(mass.sqrt.tensor, device)
  return inv.contiguous(), sq.contiguous()


def _new_key(device):
  return torch.empty(2, dtype=torch.int32, device=device)


def langevin_diffusion(potential_fn, batch_fn, adaption=None):
  """Stochastic-gradient Langevin step, optionally preconditioned (rms_prop)."""
  batch_init, batch_get, _ = batch_fn
  if adaption is not None:
    adapt_init, adapt_update, adapt_get = adaption[:3]
  fused_rms = adaption is not None and getattr(adaption, "kind", "") == "rms_prop"

  def init_fn(init_sample, key=None, adaption_kwargs=None, batch_kwargs=None,
              init_model_state=None):
    dev = _device_of(init_sample)
    sample = _to_device_tree(init_sample, dev)
    key = R.PRNGKey(0, dev) if key is None else key
    adaption_kwargs = adaption_kwargs or {}
    batch_kwargs = batch_kwargs or {}
    adapt_state = adapt_init(sample, **adaption_kwargs) if adaption is not None else None
    return LangevinState(latent_variables=sample, model_state=init_model_state, key=key,
                         adapt_state=adapt_state, data_state=batch_init(**batch_kwargs),
                         potential=torch.zeros((), device=dev),
                         variance=torch.ones((), device=dev))

  def get_fn(state):
    return {"variables": state.latent_variables, "likelihood": -state.potential,
            "model_state": state.model_state}

  def update_fn(state, parameters):
    dev = state.key.device
    lay = T.Layout(state.latent_variables)
    theta = lay.ravel(state.latent_variables, dev)
    data_state, mini_batch = batch_get(state.data_state, information=True)
    (U, (lik, new_ms)), g = potential_fn.value_and_grad(
        state.latent_variables, mini_batch, state=state.model_state, likelihoods=True)
    variance = torch.var(lik, unbiased=False)
    g = g.contiguous()
    table, nb = lay.leaf_table(dev, _lib.PAIRS_PER_BLOCK)
    key_out = _new_key(dev)
    eps = float(parameters.step_size)
    temp = float(parameters.temperature)
    adapt_state = state.adapt_state
    if adaption is None or fused_rms:
      v = None
      if fused_rms:
        v0, alpha, lmbd = adapt_state.state
        v = v0.clone()
      else:
        alpha, lmbd = 0.9, 1e-5
      _lib.call("sgmc_sgld_step", _lib.ptr(theta), _lib.ptr(g), _lib.ptr(v), lay.size,
                _lib.ptr(table), nb, _lib.ptr(state.key), _lib.ptr(key_out), eps, temp, alpha,
                lmbd, _lib.stream())
      if fused_rms:
        adapt_state = adapt_state._replace(state=(v, alpha, lmbd))
    else:
      key_out, split = R.split(state.key)
      noise = R.random_tree(split, lay)
      adapt_state = adapt_update(adapt_state, theta, g, mini_batch=mini_batch)
      manifold = adapt_get(adapt_state, theta, g, mini_batch=mini_batch)
      G = lay.ravel(manifold.g_inv.tensor, dev)
      sG = lay.ravel(manifold.sqrt_g_inv.tensor, dev)
      gam = lay.ravel(manifold.gamma.tensor, dev)
      sg = (-eps) * g
      sn = (2.0 * temp * eps) ** 0.5 * noise
      theta = theta + ((eps * gam + G * sg) + sG * sn)
    return LangevinState(latent_variables=lay.unravel(theta), model_state=new_ms, key=key_out,
                         adapt_state=adapt_state, data_state=data_state,
                         potential=U.to(state.potential.dtype), variance=variance)

  return init_fn, update_fn, get_fn


def friction_leapfrog(potential_fn, batch_fn, steps=10, friction=0.25, const_mass=None,
                      noise_model=None):
  """Original SGHMC leapfrog (integrator.py:563-765). With a noise model (e.g.
  adaption.fisher_information) the momentum noise is sqrt(2 eps) (cb_diff_sqrt xi)
  with cb_diff_sqrt from the model at the new position and gradient
  (integrator.py:632-650); without one it is C (sqrt(2 eps) xi)."""
  init_data, get_data, _ = batch_fn
  if noise_model:
    init_nm, update_nm, get_nm = noise_model[:3]
  cmass = init_mass(T.tree_map(lambda x: T.as_tensor(x, "cuda"), const_mass)) \
      if const_mass else None

  def init_fn(init_sample, key=None, batch_kwargs=None, init_model_state=None):
    dev = _device_of(init_sample)
    sample = _to_device_tree(init_sample, dev)
    key = R.PRNGKey(0, dev) if key is None else key
    return LeapfrogState(positions=sample, momentum=sample,
                         potential=torch.zeros((), device=dev), model_state=init_model_state,
                         data_state=init_data(**(batch_kwargs or {})), key=key,
                         extra_fields=init_nm(sample) if noise_model else None)

  def integrate(state, parameters, mass=None):
    dev = state.key.device
    lay = T.Layout(state.positions)
    m = mass if mass else cmass
    inv_m, sqrt_m = _flat_mass(m, lay, dev)
    leaves, _ = T.flatten(friction)
    if len(leaves) == 1 and not isinstance(leaves[0], torch.Tensor) and \
        not isinstance(friction, (dict, list, tuple)):
      fr, fr_arr = float(friction), None
    else:
      fr, fr_arr = 0.0, lay.ravel(friction, dev).contiguous()
    table, nb = lay.leaf_table(dev, _lib.PAIRS_PER_BLOCK)
    theta = lay.ravel(state.positions, dev)
    p = torch.empty_like(theta)
    k_a, k_b = _new_key(dev), _new_key(dev)
    _lib.call("sgmc_sghmc_resample", _lib.ptr(p), _lib.ptr(sqrt_m), _lib.ptr(table), nb,
              _lib.ptr(state.key), _lib.ptr(k_a), _lib.stream())
    eps = float(parameters.step_size)
    data_state, ms, U = state.data_state, state.model_state, torch.zeros((), device=dev)
    extra = state.extra_fields
    fr_flat = None
    if noise_model:
      fr_flat = fr_arr if fr_arr is not None else torch.full((lay.size,), fr, device=dev)
    for _ in range(int(steps)):
      _lib.call("sgmc_sghmc_position", _lib.ptr(theta), _lib.ptr(p), _lib.ptr(inv_m),
                _lib.ptr(fr_arr), fr, lay.size, eps, _lib.stream())
      data_state, mini_batch = get_data(data_state, information=True)
      ms_before = ms
      (U, ms), g = potential_fn.value_and_grad(lay.unravel(theta), mini_batch, state=ms)
      g = g.contiguous()
      if noise_model:
        extra = update_nm(extra, theta, g, fr_flat, mini_batch=mini_batch, step_size=eps)
        nm = get_nm(extra, theta, g, fr_flat, mini_batch=mini_batch, step_size=eps,
                    model_state=ms_before)
        cb = lay.ravel(nm.cb_diff_sqrt.tensor, dev).contiguous()
        _lib.call("sgmc_sghmc_momentum_noise_model", _lib.ptr(p), _lib.ptr(g), _lib.ptr(cb),
                  _lib.ptr(table), nb, _lib.ptr(k_a), _lib.ptr(k_b), eps, _lib.stream())
      else:
        _lib.call("sgmc_sghmc_momentum", _lib.ptr(p), _lib.ptr(g), _lib.ptr(fr_arr), fr,
                  _lib.ptr(table), nb, _lib.ptr(k_a), _lib.ptr(k_b), eps, _lib.stream())
      k_a, k_b = k_b, k_a
    return LeapfrogState(positions=lay.unravel(theta), momentum=lay.unravel(p), potential=U,
                         model_state=ms, data_state=data_state, key=k_a, extra_fields=extra)

  def get_fn(state):
    return {"variables": state.positions, "energy": state.potential,
            "model_state": state.model_state}

  return init_fn, integrate, get_fn


def obabo(potential_fn, batch_fn, steps=10, friction=1.0, const_mass=None):
  """OBABO with two fresh minibatches per step (integrator.py:138-346)."""
  init_data, get_data, _ = batch_fn
  cmass = init_mass(T.tree_map(lambda x: T.as_tensor(x, "cuda"), const_mass)) \
      if const_mass else None
  ws_cache = {}

  def init_fn(init_sample, key=None, batch_kwargs=None, init_model_state=None):
    dev = _device_of(init_sample)
    sample = _to_device_tree(init_sample, dev)
    key = R.PRNGKey(0, dev) if key is None else key
    z = torch.zeros((), device=dev)
    return ObaboState(positions=sample, momentum=T.tree_map(torch.zeros_like, sample),
                      potential=z, model_state=init_model_state,
                      data_state=init_data(**(batch_kwargs or {})), key=key,
                      kinetic_energy_start=z.clone(), kinetic_energy_end=z.clone())

  def integrate(state, parameters, mass=None):
    dev = state.key.device
    lay = T.Layout(state.positions)
    m = mass if mass else cmass
    inv_m, sqrt_m = _flat_mass(m, lay, dev)
    table, nb = lay.leaf_table(dev, _lib.PAIRS_PER_BLOCK)
    wkey = (str(dev), nb)
    if wkey not in ws_cache:
      ws_cache[wkey] = torch.zeros(int(_lib.query("sgmc_obabo_workspace_floats", nb)),
                                   device=dev)
    ws = ws_cache[wkey]
    theta = lay.ravel(state.positions, dev)
    p = lay.ravel(state.momentum, dev)
    eps = float(parameters.step_size)
    temp = float(parameters.temperature)
    fr = float(friction)
    key = state.key
    ks = state.kinetic_energy_start.reshape(1).to(torch.float32).clone()
    ke = state.kinetic_energy_end.reshape(1).to(torch.float32).clone()
    data_state, ms = state.data_state, state.model_state
    U = torch.zeros((), device=dev)
    for _ in range(int(steps)):
      keys = R.split(key, 3)
      key = keys[0].clone()
      ks_new = torch.empty_like(ks)
      _lib.call("sgmc_obabo_o", _lib.ptr(p), _lib.ptr(sqrt_m), _lib.ptr(inv_m), _lib.ptr(table),
                nb, _lib.ptr(keys[1]), fr, eps, temp, 0, _lib.ptr(ks), _lib.ptr(ks_new),
                _lib.ptr(ws), _lib.stream())
      ks = ks_new
      data_state, mb = get_data(data_state, information=True)
      (U0, ms), g0 = potential_fn.value_and_grad(lay.unravel(theta), mb, state=ms)
      _lib.call("sgmc_obabo_ba", _lib.ptr(theta), _lib.ptr(p), _lib.ptr(g0.contiguous()),
                _lib.ptr(inv_m), lay.size, eps, _lib.stream())
      data_state, mb = get_data(data_state, information=True)
      (U1, ms), g1 = potential_fn.value_and_grad(lay.unravel(theta), mb, state=ms)
      _lib.call("sgmc_obabo_b", _lib.ptr(p), _lib.ptr(g1.contiguous()), lay.size, eps,
                _lib.stream())
      ke_new = torch.empty_like(ke)
      _lib.call("sgmc_obabo_o", _lib.ptr(p), _lib.ptr(sqrt_m), _lib.ptr(inv_m), _lib.ptr(table),
                nb, _lib.ptr(keys[2]), fr, eps, temp, 1, _lib.ptr(ke), _lib.ptr(ke_new),
                _lib.ptr(ws), _lib.stream())
      ke = ke_new
      U = 0.5 * (U0 + U1)
    return ObaboState(positions=lay.unravel(theta), momentum=lay.unravel(p), potential=U,
                      model_state=ms, data_state=data_state, key=key,
                      kinetic_energy_start=ks.reshape(()), kinetic_energy_end=ke.reshape(()))

  def get_fn(state):
    return {"variables": state.positions, "energy": state.potential,
            "model_state": state.model_state}

  return init_fn, integrate, get_fn


def reversible_leapfrog(potential_fn, batch_fn, steps=10, friction=0.25, const_mass=None):
  """AMAGOLD's reversible leapfrog (integrator.py:349-560): half position step,
  ``steps`` x {position step (skipped at step 0), gradient, momentum
  p' = ((1 - eps f) p - eps g + sqrt(4 f eps) M^{1/2} xi) / (1 + eps f), energy
  += 0.5 eps (p + p') . M^-1 g}, final half position step. The momentum is drawn once
  in init_fn and carried across iterations; the state's ``potential`` is the
  accumulated energy. Each momentum launch also applies the position step that
  follows it."""
  init_data, get_data, _ = batch_fn
  cmass = init_mass(T.tree_map(lambda x: T.as_tensor(x, "cuda"), const_mass)) \
      if const_mass else None
  leaves, _ = T.flatten(friction)
  if len(leaves) != 1 or (isinstance(leaves[0], torch.Tensor) and leaves[0].numel() != 1):
    raise NotImplementedError("reversible_leapfrog takes a scalar friction")
  fr = float(leaves[0])
  ws_cache = {}

  def init_fn(init_sample, key=None, batch_kwargs=None, init_model_state=None, mass=None):
    dev = _device_of(init_sample)
    sample = _to_device_tree(init_sample, dev)
    lay = T.Layout(sample)
    _, sqrt_m = _flat_mass(mass if mass else cmass, lay, dev)
    key = R.PRNGKey(0, dev) if key is None else key
    table, nb = lay.leaf_table(dev, _lib.PAIRS_PER_BLOCK)
    p = torch.zeros(lay.size, device=dev)
    key_out = _new_key(dev)
    _lib.call("sgmc_sghmc_resample", _lib.ptr(p), _lib.ptr(sqrt_m), _lib.ptr(table), nb,
              _lib.ptr(key), _lib.ptr(key_out), _lib.stream())
    return LeapfrogState(positions=sample, momentum=lay.unravel(p),
                         potential=torch.zeros((), device=dev), model_state=init_model_state,
                         data_state=init_data(**(batch_kwargs or {})), key=key_out,
                         extra_fields=None)

  def integrate(state, parameters, mass=None):
    dev = state.key.device
    lay = T.Layout(state.positions)
    inv_m, sqrt_m = _flat_mass(mass if mass else cmass, lay, dev)
    table, nb = lay.leaf_table(dev, _lib.PAIRS_PER_BLOCK)
    wkey = (str(dev), nb)
    if wkey not in ws_cache:
      ws_cache[wkey] = torch.zeros(int(_lib.query("sgmc_obabo_workspace_floats", nb)),
                                   device=dev)
    ws = ws_cache[wkey]
    theta = lay.ravel(state.positions, dev)
    p = lay.ravel(state.momentum, dev)
    eps = float(parameters.step_size)
    half = float(np.float32(0.5) * np.float32(eps))
    _lib.call("sgmc_amagold_position", _lib.ptr(theta), _lib.ptr(p), _lib.ptr(inv_m), lay.size,
              half, _lib.stream())
    energy = torch.zeros(1, device=dev)
    data_state, ms = state.data_state, state.model_state
    keys = (_new_key(dev), _new_key(dev))
    key = state.key
    for step in range(int(steps)):
      data_state, mb = get_data(data_state, information=True)
      (_, ms), g = potential_fn.value_and_grad(lay.unravel(theta), mb, state=ms)
      e_out = torch.empty(1, device=dev)
      k_out = keys[step % 2]
      nxt = eps if step < int(steps) - 1 else half
      _lib.call("sgmc_amagold_momentum", _lib.ptr(p), _lib.ptr(g.contiguous()), _lib.ptr(inv_m),
                _lib.ptr(sqrt_m), _lib.ptr(table), nb, _lib.ptr(key), _lib.ptr(k_out), eps, fr,
                _lib.ptr(energy), _lib.ptr(e_out), _lib.ptr(theta), nxt, _lib.ptr(ws),
                _lib.stream())
      key, energy = k_out, e_out
    if int(steps) == 0:
      _lib.call("sgmc_amagold_position", _lib.ptr(theta), _lib.ptr(p), _lib.ptr(inv_m),
                lay.size, half, _lib.stream())
      key = state.key.clone()
    return LeapfrogState(positions=lay.unravel(theta), momentum=lay.unravel(p),
                         potential=energy.reshape(()), model_state=ms, data_state=data_state,
                         key=key, extra_fields=state.extra_fields)

  def get_fn(state):
    return {"variables": state.positions, "energy": state.potential,
            "model_state": state.model_state}

  return init_fn, integrate, get_fn
