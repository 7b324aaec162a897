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


class AdaptionState(NamedTuple):
  state: Any
  layout: Any


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

      def new_init(x0, *init_args, **init_kwargs):
        lay = T.Layout(x0)
        dev = T.flatten(x0)[0][0].device if T.flatten(x0)[0] else None
        flat = lay.ravel(x0, dev)
        st = init(flat, *[lay.ravel(a, dev) for a in init_args], **init_kwargs)
        return AdaptionState(state=st, layout=lay)

      def new_update(state, *update_args, mini_batch=None, **update_kwargs):
        flat = [a if (isinstance(a, torch.Tensor) and a.dim() == 1 and
                      a.numel() == state.layout.size) else state.layout.ravel(a)
                for a in update_args]
        return AdaptionState(state=update(state.state, *flat, **update_kwargs),
                             layout=state.layout)

      def new_get(state, *get_args, mini_batch=None, **get_kwargs):
        flat = [a if (isinstance(a, torch.Tensor) and a.dim() == 1 and
                      a.numel() == state.layout.size) else state.layout.ravel(a)
                for a in get_args]
        q = get(state.state, *flat, **get_kwargs)
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


def mass_matrix(*args, **kwargs):
  """adaption.py:296-369 (mass adaption for HMC) is not part of this build's path."""
  raise NotImplementedError("mass_matrix adaption is not provided by this build")


def fisher_information(*args, **kwargs):
  """adaption.py:371-457 (SGHMC noise model) is not part of this build's path."""
  raise NotImplementedError("fisher_information noise model is not provided by this build")
