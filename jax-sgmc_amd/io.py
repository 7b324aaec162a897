"""Sample collection (the part of io.py the samplers' run functions use).

The reference saves each kept sample through a host callback into a collector
(io.py:626-771) or carries them through the scan (no_save, io.py:775-879); both hand
back ``{"sample_count": k, "samples": <tree of stacked samples>}``. Here samples stay
on the device while the chain runs and are stacked once at the end, converted to
numpy when ``to_numpy`` is set (save_to_numpy=True in the aliases).
"""
import numpy as np
import torch

from . import tree as T


def _stack(samples, to_numpy):
  if not samples:
    return {}

  def st(*xs):
    if xs[0] is None:
      return None
    t = torch.stack([torch.as_tensor(x) for x in xs])
    return t.cpu().numpy() if to_numpy else t

  leaves0, td = T.flatten(samples[0])
  cols = list(zip(*[T.flatten(s)[0] for s in samples]))
  return T.unflatten(td, [st(*c) for c in cols])


def save(data_collector=None, to_numpy=True):
  """Keeps every sample with keep=True; returns (init, save, postprocess)."""

  def init(sample, *args, **kwargs):
    return []

  def save_fn(state, keep, sample, **kwargs):
    if keep:
      state.append(T.tree_map(lambda x: x.clone() if isinstance(x, torch.Tensor) else x,
                              sample))
    return state

  def post(state):
    return {"sample_count": len(state), "samples": _stack(state, to_numpy)}

  return init, save_fn, post


def no_save():
  """io.py:775-879: samples returned through the loop (device tensors)."""
  return save(to_numpy=False)


class MemoryCollector:
  """Placeholder for the reference's in-memory collector: selects save(to_numpy)."""
