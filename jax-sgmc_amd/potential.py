"""Stochastic and full potentials (potential.py:94-293 of the reference).

``minibatch_potential(prior, likelihood, strategy, has_state, is_batched,
temperature)`` returns a callable with the reference's signature
``fn(sample, reference_data, state=None, mask=None, likelihoods=False)`` returning
``(U, state)`` or ``(U, (likelihoods, state))`` with
    U = (-N * mean(lik) - prior(sample)) / T            (mask is None)
    U = (-N / n * dot(lik, mask) - prior(sample)) / T    (masked)
and a ``value_and_grad`` method standing in for jax.value_and_grad (integrator.py:166,
593, 792).

Dispatch: a recognized family (likelihoods.LinearGaussian / Logistic with a
GaussianPrior or ZeroPrior, stateless, unmasked, single-leaf sample) runs the
``sgmc_glm_potential`` kernel: gradient and potential in one launch. Any other torch
callable runs through torch on the GPU (vmap over observations for 'map'/'vmap',
autograd for the gradient) - the reference's arbitrary-likelihood path.
``full_potential`` runs the FP64 full-data kernel for a recognized family on a
device-resident dataset, and the masked batch scan otherwise.
"""
from functools import partial

import torch

from . import _lib
from . import likelihoods as LK
from . import tree as T


def _check_strategy(strategy):
  if strategy not in ("map", "vmap", "pmap"):
    raise NotImplementedError(f"Strategy {strategy} is unknown")
  if strategy == "pmap":
    raise NotImplementedError("Strategy pmap needs several devices; the library runs on one "
                              "GPU (use 'map' or 'vmap')")


def _prior_spec(prior):
  if isinstance(prior, LK.GaussianPrior):
    return 1, prior.scale
  if isinstance(prior, LK.ZeroPrior):
    return 0, 1.0
  return None


class _Workspace:
  """Per-(n, D) device workspace; zero-filled once, counters self-resetting."""

  def __init__(self):
    self.cache = {}

  def get(self, device, n, D):
    key = (str(device), n, D)
    if key not in self.cache:
      nb = _lib.query("sgmc_glm_blocks", n, D)
      ws = torch.zeros(int(_lib.query("sgmc_glm_workspace_floats", n, D, nb)), device=device)
      self.cache[key] = (ws, nb)
    return self.cache[key]


class StochasticPotential:

  def __init__(self, prior, likelihood, strategy="map", has_state=False, is_batched=False,
               temperature=1.0):
    _check_strategy(strategy)
    self.prior = prior
    self.likelihood = likelihood
    self.strategy = strategy
    self.has_state = has_state
    self.is_batched = is_batched
    self.temperature = float(temperature)
    self._ws = _Workspace()
    self.fast = (isinstance(likelihood, LK.GLMLikelihood) and _prior_spec(prior) is not None
                 and not has_state)

  # --- fast path -------------------------------------------------------------
  def _fast_ok(self, sample, reference_data, state, mask):
    if not self.fast or mask is not None or state is not None:
      return False
    leaves, _ = T.flatten(sample)
    if len(leaves) != 1 or not isinstance(leaves[0], torch.Tensor) or leaves[0].dim() != 1:
      return False
    batch, _ = reference_data
    x = batch.get(self.likelihood.x) if isinstance(batch, dict) else None
    return (x is not None and x.dim() == 2 and x.dtype == torch.float32 and x.is_cuda
            and x.stride(1) == 1 and x.stride(0) % 4 == 0 and x.data_ptr() % 16 == 0
            and x.shape[1] <= 1024)

  def _fast(self, sample, reference_data, want_lik):
    batch, info = reference_data
    w = T.flatten(sample)[0][0].contiguous().to(torch.float32)
    X = batch[self.likelihood.x]
    y = batch[self.likelihood.y].to(torch.float32).contiguous()
    n, D = X.shape
    N = int(info.observation_count)
    has_prior, pscale = _prior_spec(self.prior)
    dev = X.device
    ws, nb = self._ws.get(dev, n, D)
    U = torch.empty((), device=dev)
    g = torch.empty(D, device=dev)
    lik = torch.empty(n, device=dev) if want_lik else None
    var = torch.empty((), device=dev) if want_lik else None
    _lib.call("sgmc_glm_potential", self.likelihood.family, _lib.ptr(X), X.stride(0),
              _lib.ptr(y), None, n, N, D, _lib.ptr(w), self.likelihood.sigma, has_prior, pscale,
              self.temperature, _lib.ptr(U), _lib.ptr(g), _lib.ptr(lik), _lib.ptr(var),
              _lib.ptr(ws), ws.numel(), nb, _lib.stream())
    return U, g, lik, var

  # --- generic torch path ----------------------------------------------------
  def _generic(self, sample, reference_data, state, mask, want_grad):
    batch, info = reference_data
    leaves, td = T.flatten(sample)
    params = [torch.as_tensor(l, dtype=torch.float32).detach().clone().requires_grad_(want_grad)
              for l in leaves]
    s = T.unflatten(td, params)
    with torch.enable_grad() if want_grad else torch.no_grad():
      if self.is_batched:
        if self.has_state:
          lik, new_state = self.likelihood(state, s, batch)
        else:
          lik, new_state = self.likelihood(s, batch), None
      else:
        if self.has_state:
          fn = lambda obs: self.likelihood(state, s, obs)
          lik, states = torch.func.vmap(fn)(batch)
          new_state = T.tree_map(lambda a, o: torch.as_tensor(a)[0].reshape(
              torch.as_tensor(o).shape), states, state) if state is not None else None
        else:
          lik = torch.func.vmap(lambda obs: torch.as_tensor(self.likelihood(s, obs)))(batch)
          new_state = None
      lik = torch.as_tensor(lik).reshape(-1)
      N = info.observation_count
      n = info.batch_size
      if mask is None:
        Ulik = -N * torch.mean(lik, dim=0)
      else:
        Ulik = -N / n * torch.dot(lik, torch.as_tensor(mask, dtype=lik.dtype))
      U = (Ulik - torch.as_tensor(self.prior(s), device=lik.device)) / self.temperature
    grads = None
    if want_grad:
      gs = torch.autograd.grad(U, params, allow_unused=True)
      grads = [torch.zeros_like(p) if g is None else g for g, p in zip(gs, params)]
    return U.detach(), grads, lik.detach(), new_state

  # --- reference signature ---------------------------------------------------
  def __call__(self, sample, reference_data, state=None, mask=None, likelihoods=False):
    if self._fast_ok(sample, reference_data, state, mask):
      U, _, lik, _ = self._fast(sample, reference_data, likelihoods)
      return (U, (lik, None)) if likelihoods else (U, None)
    U, _, lik, new_state = self._generic(sample, reference_data, state, mask, False)
    return (U, (lik, new_state)) if likelihoods else (U, new_state)

  def value_and_grad(self, sample, reference_data, state=None, likelihoods=False):
    """((U, aux), grad) with grad as a flat float32 vector in ravel order; aux is the
    new state, or (likelihoods, state)."""
    if self._fast_ok(sample, reference_data, state, None):
      U, g, lik, _ = self._fast(sample, reference_data, likelihoods)
      aux = (lik, None) if likelihoods else None
      return (U, aux), g
    U, grads, lik, new_state = self._generic(sample, reference_data, state, None, True)
    g = torch.cat([x.reshape(-1) for x in grads]) if grads else torch.zeros(0)
    aux = (lik, new_state) if likelihoods else new_state
    return (U, aux), g


  def differentiable(self, w, reference_data):
    """U(w) of a recognized GLM family on one minibatch as a differentiable torch value:
    the forward launch of sgmc_glm_potential also writes the gradient, and the backward
    returns grad_output times that gradient (the hand-written backward in glm.hip), so
    ``U.backward()`` launches nothing more. ``w`` is the weight vector (one leaf)."""
    if not self._fast_ok(w, reference_data, None, None):
      raise ValueError("differentiable() needs a GLM family potential (GaussianPrior or "
                       "ZeroPrior), a float32 device weight vector and a device batch with "
                       "16-byte aligned rows, D <= 1024")
    return GLMPotentialFn.apply(w, self, reference_data)


class GLMPotentialFn(torch.autograd.Function):
  """Minibatch potential of a GLM family (potential.py:159-214) whose backward is the
  kernel's gradient (jax.value_and_grad at integrator.py:166, 593, 792)."""

  @staticmethod
  def forward(ctx, w, potential, reference_data):
    U, g, _, _ = potential._fast(w.detach(), reference_data, False)
    ctx.save_for_backward(g)
    return U

  @staticmethod
  def backward(ctx, grad_U):
    (g,) = ctx.saved_tensors
    return grad_U * g, None, None


def minibatch_potential(prior, likelihood, strategy="map", has_state=False, is_batched=False,
                        temperature=1.0):
  """potential.py:94-216."""
  return StochasticPotential(prior, likelihood, strategy, has_state, is_batched, temperature)


class FullPotential:

  def __init__(self, prior, likelihood, strategy="map", has_state=False, is_batched=False,
               temperature=1.0):
    assert strategy != "pmap", "Pmap is currently not supported"
    _check_strategy(strategy)
    self.prior = prior
    self.likelihood = likelihood
    self.temperature = float(temperature)
    self.batch_potential = minibatch_potential(LK.ZeroPrior(), likelihood, strategy,
                                               has_state, is_batched)
    self.fast = (isinstance(likelihood, LK.GLMLikelihood) and _prior_spec(prior) is not None
                 and not has_state)
    self._ws = {}

  def _fast_full(self, sample, loader):
    data = loader.get_full_data()
    X = data[self.likelihood.x]
    y = data[self.likelihood.y].to(torch.float32).contiguous()
    w = T.as_tensor(T.flatten(sample)[0][0], X.device).to(torch.float32).contiguous()
    N, D = X.shape
    rpb = 1024
    key = (N, rpb)
    if key not in self._ws:
      self._ws[key] = torch.zeros(int(_lib.query("sgmc_glm_full_workspace_doubles", N, rpb)),
                                  dtype=torch.float64, device=X.device)
    ws = self._ws[key]
    out = torch.empty((), dtype=torch.float64, device=X.device)
    has_prior, pscale = _prior_spec(self.prior)
    _lib.call("sgmc_glm_full_potential", self.likelihood.family, _lib.ptr(X), X.stride(0),
              _lib.ptr(y), N, D, _lib.ptr(w), self.likelihood.sigma, has_prior, pscale,
              self.temperature, _lib.ptr(out), _lib.ptr(ws), ws.numel(), rpb, _lib.stream())
    return out

  def __call__(self, sample, data_state, full_data_map_fn, state=None):
    loader = getattr(full_data_map_fn, "loader", None)
    from .data import DeviceDataLoader
    leaves, _ = T.flatten(sample)
    if (self.fast and state is None and isinstance(loader, DeviceDataLoader) and len(leaves) == 1
        and isinstance(leaves[0], torch.Tensor) and leaves[0].dim() == 1):
      X = loader.get_full_data()[self.likelihood.x]
      if X.stride(0) % 4 == 0 and X.shape[1] <= 1024:
        U = self._fast_full(sample, loader)
        # the ordered scan visits every row once; its end position is data-independent
        N = loader.static_information["observation_count"]
        bs = full_data_map_fn.batch_size
        nb = -(-N // bs)
        line = (data_state.current_line + nb * bs) % N if data_state is not None else 0
        new_ds = type(data_state)(current_line=line) if data_state is not None else None
        return U, (new_ds, None)

    def body(batch_and_info, mask, carry):
      pot, st = self.batch_potential(sample, batch_and_info, carry, mask)
      _, info = batch_and_info
      return pot * info.batch_size / info.observation_count, st

    if data_state is None:
      results, new_state = full_data_map_fn(body, state, masking=True, information=True)
    else:
      data_state, (results, new_state) = full_data_map_fn(body, data_state, state,
                                                          masking=True, information=True)
    U = (torch.sum(results) - torch.as_tensor(self.prior(sample), device=results.device)) \
        / self.temperature
    return U, (data_state, new_state)


def full_potential(prior, likelihood, strategy="map", has_state=False, is_batched=False,
                   temperature=1.0):
  """potential.py:219-293."""
  return FullPotential(prior, likelihood, strategy, has_state, is_batched, temperature)
