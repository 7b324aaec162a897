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


def _leaf_names(tree, prefix=()):
  """'/'-joined dict keys (positions for sequences and named tuples) of every leaf, in
  flatten order."""
  if tree is None:
    return []
  if isinstance(tree, dict):
    return [n for k in sorted(tree) for n in _leaf_names(tree[k], prefix + (str(k),))]
  if isinstance(tree, (list, tuple)):
    return [n for i, v in enumerate(tree) for n in _leaf_names(v, prefix + (str(i),))]
  return ["/".join(prefix) if prefix else "leaf"]


def _numpy_dtype(leaf):
  if isinstance(leaf, torch.Tensor):
    return torch.empty(0, dtype=leaf.dtype).numpy().dtype
  return np.asarray(leaf).dtype


class MemoryCollector:
  """Stores the kept samples of each chain in host RAM as numpy arrays (io.py:502-616).

  register_chain allocates, per leaf of the initial sample, an array of
  static_information.samples_collected entries; save writes one sample; finalize
  optionally writes ``save_dir/chain_<id>.npz`` (one array per leaf, named by its
  '/'-joined dict keys); finished returns the samples in the sample's tree structure."""

  def __init__(self, save_dir=None):
    self._samples = []
    self._counts = []
    self._treedefs = []
    self._names = []
    self._dir = save_dir

  def register_data_loader(self, data_loader):
    """Registers the data loader whose state a checkpoint would hold (no checkpoints)."""

  def register_chain(self, init_sample=None, static_information=None, **unused_kwargs):
    assert init_sample is not None, "Need a sample-pytree to allocate memory."
    leaves, treedef = T.flatten(init_sample)
    total = int(static_information.samples_collected)
    cache = [np.zeros((total,) + tuple(np.shape(l)), dtype=_numpy_dtype(l)) for l in leaves]
    self._samples.append(cache)
    self._counts.append(0)
    self._treedefs.append(treedef)
    self._names.append(_leaf_names(init_sample) if self._dir else
                       [f"leaf~{i}" for i in range(len(leaves))])
    return len(self._samples) - 1

  def save(self, chain_id, values):
    """Writes the leaves of one sample (flatten order) as the next entry."""
    k = self._counts[chain_id]
    for arr, v in zip(self._samples[chain_id], values):
      arr[k] = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
    self._counts[chain_id] = k + 1

  def finalize(self, chain_id):
    if self._dir:
      import os
      os.makedirs(self._dir, exist_ok=True)
      np.savez(os.path.join(self._dir, f"chain_{chain_id}.npz"),
               **dict(zip(self._names[chain_id], self._samples[chain_id])))

  def finished(self, chain_id):
    return T.unflatten(self._treedefs[chain_id], self._samples[chain_id])

  def checkpoint(self, chain_id, state):
    """Called every nth step (checkpoints are not supported, as in the reference)."""

  def resume(self):
    """Restores a checkpoint (not supported, as in the reference)."""


def save(data_collector=None, checkpoint_every=0, to_numpy=True):
  """io.py:626-771: keeps every sample with keep=True; returns (init, save,
  postprocess). With a data collector, each kept sample goes to the host collector as
  it is produced and postprocess returns what the collector holds; without one, the
  samples stay on the device and are stacked once at the end."""
  if checkpoint_every != 0:
    raise NotImplementedError("Checkpointing is not supported yet.")
  if data_collector is not None:

    def init_c(sample, init_checkpoint=None, static_information=None, **kwargs):
      cid = data_collector.register_chain(init_sample=sample,
                                          static_information=static_information)
      return {"chain_id": cid, "count": 0}

    def save_c(state, keep, sample, **kwargs):
      if keep:
        data_collector.save(state["chain_id"], T.flatten(sample)[0])
        state = dict(state, count=state["count"] + 1)
      return state

    def post_c(state, *unused):
      data_collector.finalize(state["chain_id"])
      return {"sample_count": state["count"],
              "samples": data_collector.finished(state["chain_id"])}

    return init_c, save_c, post_c

  def init(sample, *args, **kwargs):
    return []

  def save_fn(state, keep, sample, **kwargs):
    if keep:
      state.append(T.tree_map(lambda x: x.clone() if isinstance(x, torch.Tensor) else x,
                              sample))
    return state

  def post(state, *unused):
    return {"sample_count": len(state), "samples": _stack(state, to_numpy)}

  return init, save_fn, post


def no_save():
  """io.py:775-879: samples returned through the loop (device tensors)."""
  return save(to_numpy=False)
