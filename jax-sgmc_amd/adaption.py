"""Manifold adaption on ravelled parameters (adaption.py:53-293 of the reference).

``rms_prop()`` returns the (init, update, get) triplet of adaption.py:225-293, wrapped
like the reference's ``adaption`` decorator (adaption.py:107-222): positional
arguments are ravelled to one flat vector (tree_flatten order), ``get`` returns a
``Manifold`` of diagonal ``Tensor``s unravelled to the sample's structure.

langevin_diffusion recognizes rms_prop and fuses update, get and the step into the
``sgmc_sgld_step`` kernel (one pass over the parameters); other strategies built with
``adaption`` run their torch update/get functions.
"""
from typing import Any, NamedTuple

import torch

from . import tree as T


class Tensor(NamedTuple):
  ndim: int
  tensor: Any


class Manifold(NamedTuple):
  g_inv: Tensor
  sqrt_g_inv: Tensor
  gamma: Tensor


class MassMatrix(NamedTuple):
  inv: Tensor
  sqrt: Tensor


class NoiseModel(NamedTuple):
  cb_diff_sqrt: Tensor
  b_sqrt: Tensor


class AdaptionState(NamedTuple):
  state: Any
  layout: Any
  flat_potential: Any = None


class FlatPotential:
  """The minibatch potential seen from a flat parameter vector (the reference's
  ``flat_potential``, adaption.py:146-158): ``observation_grads`` gives the gradient of
  every observation's likelihood, the rows of d likelihoods / d theta that the
  reference takes one observation at a time (adaption.py:418-425)."""

  def __init__(self, potential, layout):
    self.potential = potential
    self.layout = layout

  def observation_grads(self, flat, mini_batch, model_state=None):
    pot = self.potential
    lik = pot.likelihood
    lay = self.layout
    batch, _ = mini_batch
    flat = flat.detach().to(torch.float32)
    if pot.is_batched:
      if pot.has_state:
        fn = lambda th: torch.as_tensor(lik(model_state, lay.unravel(th), batch)[0]).reshape(-1)
      else:
        fn = lambda th: torch.as_tensor(lik(lay.unravel(th), batch)).reshape(-1)
      return torch.func.jacrev(fn)(flat)
    if pot.has_state:
      one = lambda th, obs: torch.as_tensor(lik(model_state, lay.unravel(th), obs)[0]).reshape(())
    else:
      one = lambda th, obs: torch.as_tensor(lik(lay.unravel(th), obs)).reshape(())
    return torch.func.vmap(torch.func.grad(one), in_dims=(None, 0))(flat, batch)


class AdaptionStrategy(NamedTuple):
  init: Any
  update: Any
  get: Any
  kind: str = "custom"


def adaption(quantity=tuple, kind="custom"):
  """Decorator: makes an (init, update, get) strategy on 1-D vectors work on pytrees."""

  def wrap(fn):
    def strategy(*args, **kwargs):
      init, update, get = fn(*args, **kwargs)
      potential = kwargs.get("minibatch_potential")

      def new_init(x0, *init_args, **init_kwargs):
        lay = T.Layout(x0)
        dev = T.flatten(x0)[0][0].device if T.flatten(x0)[0] else None
        flat = lay.ravel(x0, dev)
        st = init(flat, *[None if a is None else lay.ravel(a, dev) for a in init_args],
                  **init_kwargs)
        fp = FlatPotential(potential, lay) if potential is not None else None
        return AdaptionState(state=st, layout=lay, flat_potential=fp)

      def _flat(state, args):
        return [a if (isinstance(a, torch.Tensor) and a.dim() == 1 and
                      a.numel() == state.layout.size) else state.layout.ravel(a)
                for a in args]

      def new_update(state, *update_args, mini_batch=None, **update_kwargs):
        flat = _flat(state, update_args)
        if state.flat_potential is None:
          st = update(state.state, *flat, **update_kwargs)
        else:
          assert mini_batch, "Adaption strategy requires mini-batch"
          st = update(state.state, *flat, mini_batch, state.flat_potential, **update_kwargs)
        return state._replace(state=st)

      def new_get(state, *get_args, mini_batch=None, **get_kwargs):
        flat = _flat(state, get_args)
        if state.flat_potential is None:
          q = get(state.state, *flat, **get_kwargs)
        else:
          q = get(state.state, *flat, mini_batch=mini_batch,
                  flat_potential=state.flat_potential, **get_kwargs)
        out = []
        for v in q:
          if v.dim() == 1:
            out.append(Tensor(ndim=1, tensor=state.layout.unravel(v)))
          else:
            out.append(Tensor(ndim=2, tensor=v))
        return quantity(*out)

      return AdaptionStrategy(new_init, new_update, new_get, kind)

    return strategy

  return wrap


@adaption(quantity=Manifold, kind="rms_prop")
def rms_prop():
  """v0 = 1; v <- alpha v + (1 - alpha) g^2; G = (lmbd + sqrt(v))^-1; sqrt(G); Gamma = 0."""

  def init(sample, alpha=0.9, lmbd=1e-5):
    return (torch.ones_like(sample), float(alpha), float(lmbd))

  def update(state, sample, sample_grad, *args, **kwargs):
    v, alpha, lmbd = state
    return (alpha * v + (1 - alpha) * sample_grad * sample_grad, alpha, lmbd)

  def get(state, sample, sample_grad, *args, **kwargs):
    v, _, lmbd = state
    g = 1.0 / (lmbd + torch.sqrt(v))
    return g, torch.sqrt(g), torch.zeros_like(g)

  return init, update, get


@adaption(quantity=MassMatrix, kind="mass_matrix")
def mass_matrix(diagonal=True, burn_in=1000):
  """adaption.py:296-369: running mean and sum of squared deviations of the accepted
  samples (Welford); at iteration == burn_in the mass becomes M^-1 = ssq / iteration,
  M^{1/2} = sqrt(iteration / ssq) (diagonal) or via the eigendecomposition (full).
  Before that, M^-1 = init_cov and M^{1/2} = 1 / sqrt(init_cov) (init_cov = ones by
  default)."""

  def _matrix(iteration, ssq):
    inv = ssq / iteration
    if diagonal:
      return inv, torch.sqrt(iteration / ssq)
    w, v = torch.linalg.eigh(ssq / iteration)
    return inv, v.T @ (torch.diag(torch.sqrt(w)) @ v)

  def init(sample, init_cov=None):
    mean = torch.zeros_like(sample)
    ssq = torch.zeros_like(sample) if diagonal else torch.zeros(
        (sample.numel(), sample.numel()), dtype=sample.dtype, device=sample.device)
    cov = torch.ones_like(sample) if init_cov is None else init_cov
    if diagonal:
      return (0, mean, ssq, cov, 1 / torch.sqrt(cov))
    return (0, mean, ssq, torch.diag(cov), torch.diag(1 / torch.sqrt(cov)))

  def update(state, sample, *args, **kwargs):
    iteration, mean, ssq, m_inv, m_sqrt = state
    iteration += 1
    new_mean = (iteration - 1) / iteration * mean + 1 / iteration * sample
    if diagonal:
      ssq = ssq + (sample - mean) * (sample - new_mean)
    else:
      ssq = ssq + torch.outer(sample - mean, sample - new_mean)
    if iteration == burn_in:
      m_inv, m_sqrt = _matrix(iteration, ssq)
    return iteration, new_mean, ssq, m_inv, m_sqrt

  def get(state, *args, **kwargs):
    return state[3], state[4]

  return init, update, get


@adaption(quantity=NoiseModel, kind="fisher_information")
def fisher_information(minibatch_potential=None, diagonal=True):
  """adaption.py:371-457: SGHMC noise model from the empirical Fisher information of
  the minibatch. With g the potential's gradient and G_i the gradient of observation
  i's likelihood: ssq = sum_i (G_i - g / N)^2 (or outer products), v = ssq / (n - 1),
  b = eps v / 2. Diagonal: c = friction - b, entries c <= 0 replaced by the smallest
  positive entry; returns (sqrt(c), sqrt(friction - c)). Full: singular vectors times
  the square roots of the singular values of diag(friction) - b and of b."""
  assert minibatch_potential, "Fisher information requires potential function."

  def init(*args):
    return None

  def update(*args, **kwargs):
    return None

  def get(state, sample, sample_grad, friction, *args, mini_batch, flat_potential,
          step_size=1.0, model_state=None, **kwargs):
    _, info = mini_batch
    N = info.observation_count
    n = info.batch_size
    sg = sample_grad / N
    G = flat_potential.observation_grads(sample, mini_batch, model_state)
    diff = G - sg
    eps = torch.as_tensor(step_size, dtype=torch.float32, device=sample.device)
    if diagonal:
      ssq = torch.sum(diff * diff, dim=0)
      v = (1 / (n - 1)) * ssq
      b = (0.5 * eps) * v
      corr = friction - b
      inf = torch.full_like(corr, float("inf"))
      smallest = torch.min(torch.where(corr <= 0, inf, corr))
      pos = torch.where(corr <= 0, smallest, corr)
      return torch.sqrt(pos), torch.sqrt(friction - pos)
    ssq = diff.T @ diff
    v = (1 / (n - 1)) * ssq
    b = (0.5 * eps) * v
    u_cb, s_cb, _ = torch.linalg.svd(torch.diag(friction) - b)
    u, s_, _ = torch.linalg.svd(b)
    return u_cb @ torch.sqrt(s_cb), u @ torch.sqrt(s_)

  return init, update, get
