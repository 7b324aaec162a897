"""Reference data access: device-resident datasets, random minibatches, ordered
full-data batching.

Mirrors jax_sgmc.data:
- MiniBatchInformation (data/core.py:49-61), CacheState (data/core.py:240-259);
- DeviceNumpyDataLoader (data/numpy_loader.py:89-144): the whole dataset lives in
  HBM; a random batch is ``key, split = split(state); idx = randint(split, (n,), 0,
  N); tree_index(data, idx)``, run as the ``sgmc_draw_batch`` and ``sgmc_gather_rows``
  kernels;
- NumpyDataLoader (data/numpy_loader.py:147-460): host arrays, numpy PCG64 index draws
  seeded with ``SeedSequence(seed).spawn(1)[0]`` (seed defaults to the chain id), a
  cache of ``cache_size`` batches refreshed when exhausted (data/core.py:664-835);
- random_reference_data / full_reference_data (data/core.py:477-629) with the same
  argument checks and exceptions.

2-D float datasets are stored with their rows padded to a multiple of 4 floats (zero
padding), the layout the GLM kernels read; arrays returned to the user are views
without the padding.
"""
import math
from typing import Any, NamedTuple

import numpy as np
import torch

from . import _lib
from . import random as R
from . import tree as T


class MiniBatchInformation(NamedTuple):
  observation_count: Any
  mask: Any
  batch_size: Any


class CacheState(NamedTuple):
  cached_batches: Any = None
  cached_batches_count: Any = None
  current_line: Any = None
  chain_id: Any = None
  state: Any = None
  valid: Any = None
  callback_uuid: Any = None
  token: Any = None


class DataLoader:
  pass


class DeviceDataLoader(DataLoader):
  pass


class HostDataLoader(DataLoader):
  pass


def _to_device_array(a, device):
  if isinstance(a, torch.Tensor):
    t = a.to(device)
    if t.dtype == torch.float64:
      t = t.to(torch.float32)
    if t.dtype == torch.int64:
      t = t.to(torch.int32)
    t = t.contiguous()
  else:
    a = np.asarray(a)
    if a.dtype == np.float64:
      a = a.astype(np.float32)
    if a.dtype == np.int64:
      a = a.astype(np.int32)
    t = torch.as_tensor(np.ascontiguousarray(a), device=device)
  if t.dim() == 2 and t.dtype == torch.float32 and t.shape[1] % 4:
    ld = (t.shape[1] + 3) // 4 * 4
    store = torch.zeros((t.shape[0], ld), dtype=t.dtype, device=device)
    store[:, :t.shape[1]] = t
    return store[:, :t.shape[1]]
  return t


def gather(leaf, idx):
  """tree_index for one leaf (data/core.py:654-659) with the row-gather kernel. A
  2-D leaf keeps its padded row stride, so a gathered batch has the layout of the
  dataset."""
  n = int(idx.numel())
  N = int(leaf.shape[0])
  es = leaf.element_size()
  if leaf.dim() == 2 and leaf.stride(1) == 1 and leaf.stride(0) >= leaf.shape[1]:
    ld = leaf.stride(0)
    store = torch.empty((n, ld), dtype=leaf.dtype, device=leaf.device)
    _lib.call("sgmc_gather_rows", _lib.ptr(leaf), N, _lib.ptr(idx), _lib.ptr(store), n,
              ld * es, ld * es, _lib.stream())
    return store[:, :leaf.shape[1]]
  if not leaf.is_contiguous():
    leaf = leaf.contiguous()
  row_elems = int(np.prod(leaf.shape[1:])) if leaf.dim() > 1 else 1
  out = torch.empty((n,) + tuple(leaf.shape[1:]), dtype=leaf.dtype, device=leaf.device)
  _lib.call("sgmc_gather_rows", _lib.ptr(leaf), N, _lib.ptr(idx), _lib.ptr(out), n,
            row_elems * es, row_elems * es, _lib.stream())
  return out


def tree_index(pytree, index):
  return T.tree_map(lambda leaf: gather(leaf, index), pytree)


class DeviceNumpyDataLoader(DeviceDataLoader):
  """data/numpy_loader.py:89-144."""

  def __init__(self, copy=True, device="cuda", **reference_data):
    del copy
    self.device = torch.device(device)
    counts = []
    self._reference_data = {}
    for name, array in reference_data.items():
      counts.append(len(array))
      self._reference_data[name] = _to_device_array(array, self.device)
    if counts and any(c != counts[0] for c in counts):
      raise ValueError("All reference_data arrays must have the same length in the first "
                       "dimension.")
    self._observation_count = counts[0] if counts else 0

  @property
  def reference_data(self):
    return self._reference_data

  @property
  def static_information(self):
    return {"observation_count": self._observation_count}

  def init_random_data(self, *args, **kwargs):
    key = kwargs.get("key")
    if key is None:
      key = R.PRNGKey(0, self.device)
    return key

  def draw_indices(self, state, batch_size):
    """(new_key, idx) with idx = randint(split(state)[1], (n,), 0, N)."""
    new_key = torch.empty(2, dtype=torch.int32, device=self.device)
    idx = torch.empty(batch_size, dtype=torch.int32, device=self.device)
    _lib.call("sgmc_draw_batch", _lib.ptr(state), _lib.ptr(new_key), _lib.ptr(idx), batch_size,
              self._observation_count, _lib.stream())
    return new_key, idx

  def get_random_data(self, state, batch_size):
    new_key, idx = self.draw_indices(state, batch_size)
    batch = tree_index(self._reference_data, idx)
    info = MiniBatchInformation(observation_count=self._observation_count,
                                batch_size=batch_size,
                                mask=torch.ones(batch_size, dtype=torch.bool,
                                                device=self.device))
    return new_key, (batch, info)

  def get_full_data(self):
    return self._reference_data


class NumpyDataLoader(HostDataLoader):
  """Host-resident arrays with numpy PCG64 index draws (data/numpy_loader.py:147-460).

  Random pipelines draw with replacement (default), shuffle (successive permutations
  of all observations, consumed as one stream) or shuffle in epochs (each epoch a
  fresh permutation; the last batch of an epoch is padded and its surplus entries
  masked). Ordered pipelines walk the rows with wrap-around. Chain state can be saved
  and restored (save_state / load_state). Batches are copied to the device."""

  def __init__(self, copy=True, device="cuda", **reference_data):
    self.device = torch.device(device)
    counts = []
    self._reference_data = {}
    for name, array in reference_data.items():
      counts.append(len(array))
      self._reference_data[name] = np.array(array, copy=copy)
    if counts and any(c != counts[0] for c in counts):
      raise ValueError("All reference_data arrays must have the same length in the first "
                       "dimension.")
    self._observation_count = counts[0] if counts else 0
    self._chains = []

  @property
  def reference_data(self):
    return self._reference_data

  @property
  def static_information(self):
    return {"observation_count": self._observation_count}

  def initializer_batch(self, mb_size=None):
    """Zero batch of mb_size observations with the dtypes a batch has on the device
    (numpy_loader.py:60-86)."""
    out = {}
    for k, v in self._reference_data.items():
      t = _to_device_array(v[:1], "cpu")
      shape = (mb_size,) + tuple(v.shape[1:]) if mb_size is not None else tuple(v.shape[1:])
      out[k] = torch.zeros(shape, dtype=t.dtype, device=self.device)
    return out

  # ---------------------------------------------------------------- chains ----
  def register_random_pipeline(self, cache_size=1, mb_size=None, in_epochs=False,
                               shuffle=False, **kwargs):
    if mb_size > self._observation_count:
      raise ValueError(f"The batch size cannot be bigger than the observation count. "
                       f"Provided {mb_size} and {self._observation_count}")
    if in_epochs and not shuffle:
      raise ValueError("in_epochs = True can only be used for shuffle = True.")
    chain_id = len(self._chains)
    seed = kwargs.get("seed", chain_id)
    rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])
    N = self._observation_count
    self._chains.append({"type": "random", "rng": rng, "mb_size": mb_size,
                         "cache_size": cache_size, "shuffle": bool(shuffle),
                         "in_epochs": bool(in_epochs), "buffer": None, "draws": -(-N // mb_size),
                         "remaining": 0})
    return chain_id

  def register_ordered_pipeline(self, cache_size=1, mb_size=None, **kwargs):
    assert mb_size <= self._observation_count, (
        f"The batch size cannot be bigger than the observation count. Provided {mb_size} "
        f"and {self._observation_count}")
    chain_id = len(self._chains)
    self._chains.append({"type": "ordered", "idx_offset": 0, "mb_size": mb_size,
                         "cache_size": cache_size})
    return chain_id

  def save_state(self, chain_id):
    """Everything needed to continue chain_id from here (numpy_loader.py:185-205)."""
    c = self._chains[chain_id]
    if c["type"] == "ordered":
      return {"ordered": c["idx_offset"]}
    fields = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.items()
              if k != "rng"}
    return {"random": (c["rng"].bit_generator.state, fields)}

  def load_state(self, chain_id, data):
    """Restores a state from save_state into chain_id (numpy_loader.py:207-224)."""
    kind, value = dict(data).popitem()
    c = self._chains[chain_id]
    if kind == "ordered":
      c["idx_offset"] = value
    elif kind == "random":
      rng_state, fields = value
      c["rng"].bit_generator.state = rng_state
      for k, v in fields.items():
        c[k] = v.copy() if isinstance(v, np.ndarray) else v
    else:
      raise ValueError(f"Chain type {kind} is unknown.")

  # --------------------------------------------------------------- indices ----
  def _permutation(self, chain):
    N = self._observation_count
    return chain["rng"].choice(np.arange(0, N), size=N, replace=False)

  def _shuffled(self, chain):
    """Shuffle without epochs (numpy_loader.py:393-433): the batches read a ring buffer
    of (floor(N / n) + 2) n slots that is refilled with a fresh permutation, written
    right after the unread entries, whenever fewer than n entries remain; the stream
    of batches is the concatenation of successive permutations."""
    N, n = self._observation_count, chain["mb_size"]
    slots = (N // n + 2) * n
    if chain["remaining"] < n:
      perm = self._permutation(chain)
      if chain["buffer"] is None:
        chain["draws"] = 0
        chain["buffer"] = np.zeros(slots, dtype=np.int64)
      pos = np.mod(np.arange(N) + chain["draws"] * n + chain["remaining"], slots)
      chain["buffer"][pos] = perm
      chain["remaining"] += N
    take = np.mod(np.arange(n) + chain["draws"] * n, slots)
    sel = chain["buffer"][take].copy()
    chain["draws"] = (chain["draws"] + 1) % (N // n + 2)
    chain["remaining"] -= n
    return sel, np.ones(n, dtype=bool)

  def _epochs(self, chain):
    """Shuffle in epochs (numpy_loader.py:435-460): ceil(N / n) batches per epoch from
    one permutation; entries past N in the last batch are masked."""
    N, n = self._observation_count, chain["mb_size"]
    per_epoch = -(-N // n)
    if chain["draws"] == per_epoch:
      perm = self._permutation(chain)
      if chain["buffer"] is None:
        chain["buffer"] = np.zeros(per_epoch * n, dtype=np.int64)
      chain["buffer"][:N] = perm
      chain["draws"] = 0
    lo = n * chain["draws"]
    sel = chain["buffer"][lo:lo + n].copy()
    chain["draws"] += 1
    return sel, np.arange(lo, lo + n) < N

  def _indices(self, chain):
    N = self._observation_count
    if chain["type"] == "ordered":
      idx = np.arange(chain["mb_size"]) + chain["idx_offset"]
      mask = idx < N
      if chain["idx_offset"] + chain["mb_size"] >= N:
        chain["idx_offset"] = 0
      else:
        chain["idx_offset"] += chain["mb_size"]
      return np.mod(idx, N), mask
    if chain["in_epochs"]:
      return self._epochs(chain)
    if chain["shuffle"]:
      return self._shuffled(chain)
    sel = chain["rng"].choice(np.arange(0, N), size=chain["mb_size"], replace=True)
    return sel, np.ones(chain["mb_size"], dtype=bool)

  def get_batches(self, chain_id):
    """cache_size batches of mb_size observations: ({name: [cache, mb, ...]}, mask)."""
    chain = self._chains[chain_id]
    pairs = [self._indices(chain) for _ in range(chain["cache_size"])]
    idx = np.array([p[0] for p in pairs])
    mask = np.array([p[1] for p in pairs], dtype=bool)
    out = {k: _to_device_array(v[idx], self.device) for k, v in self._reference_data.items()}
    return out, torch.as_tensor(mask, device=self.device)


def random_reference_data(data_loader, cached_batches_count, mb_size, verify_calls=False):
  """data/core.py:477-520."""
  del verify_calls
  N = data_loader.static_information["observation_count"]
  if N < mb_size:
    raise ValueError(f"Batch size cannot be bigger than the number of total observations. "
                     f"Got {N} and {mb_size}.")
  if cached_batches_count <= 0 or mb_size <= 0:
    raise ValueError(f"Cache size and batch size must be positive, got"
                     f"{cached_batches_count} and {mb_size}.")
  if isinstance(data_loader, HostDataLoader):
    return _random_host(data_loader, cached_batches_count, mb_size)
  if isinstance(data_loader, DeviceDataLoader):
    if not cached_batches_count == 1:
      raise ValueError("No caching on device.")
    return _random_device(data_loader, mb_size)
  raise TypeError("The DataLoader must inherit from HostDataLoader or DeviceDataLoader")


def _random_device(loader, mb_size):
  """data/core.py:838-863."""

  def init_fn(**kwargs):
    return CacheState(state=loader.init_random_data(**kwargs))

  def batch_fn(state, information=False):
    key, (batch, info) = loader.get_random_data(state.state, mb_size)
    new = CacheState(state=key)
    return (new, (batch, info)) if information else (new, batch)

  batch_fn.loader = loader
  batch_fn.mb_size = mb_size
  return init_fn, batch_fn, lambda: None


def _random_host(loader, count, mb_size):
  """data/core.py:664-835 without the host callback: the cache is refilled from the
  host when all cached batches have been used."""
  N = loader.static_information["observation_count"]

  def init_fn(**kwargs):
    cid = loader.register_random_pipeline(count, mb_size=mb_size, **kwargs)
    data, mask = loader.get_batches(cid)
    return CacheState(cached_batches=data, cached_batches_count=count, current_line=0,
                      chain_id=cid, valid=mask)

  def batch_fn(state, information=False):
    if state.current_line == state.cached_batches_count:
      data, mask = loader.get_batches(state.chain_id)
      state = state._replace(cached_batches=data, valid=mask, current_line=0)
    line = state.current_line % state.cached_batches_count
    batch = {k: v[line] for k, v in state.cached_batches.items()}
    info = MiniBatchInformation(observation_count=N, batch_size=mb_size,
                                mask=state.valid[line])
    new = state._replace(current_line=line + 1)
    return (new, (batch, info)) if information else (new, batch)

  return init_fn, batch_fn, lambda: None


def full_reference_data(data_loader, cached_batches_count=100, mb_size=None):
  """data/core.py:522-629. For a DeviceDataLoader the ordered batch size is
  ``cached_batches_count`` (data/core.py:552-553), as in the reference."""
  N = data_loader.static_information["observation_count"]
  if isinstance(data_loader, DeviceDataLoader):
    bs = cached_batches_count
  elif isinstance(data_loader, HostDataLoader):
    bs = mb_size
  else:
    raise TypeError("The DataLoader must inherit from HostDataLoader or DeviceDataLoader")
  if mb_size is not None and N < mb_size:
    raise ValueError(f"Batch size cannot be bigger than the number of total observations. "
                     f"Got {N} and {mb_size}.")
  num_iterations = int(math.ceil(N / bs))

  def init_fn(offset=0, **kwargs):
    if offset >= N:
      raise ValueError(f"The offset cannot be greater than the total observation count. "
                       f"Given {offset} and {N}.")
    return CacheState(current_line=int(offset))

  def batch_scan(fun, data_state, carry, masking=False, information=False, device_count=1):
    del device_count
    line = data_state.current_line
    results = []
    dev = getattr(data_loader, "device", torch.device("cuda"))
    for it in range(num_iterations):
      idx = np.mod(np.arange(bs) + line, N)
      line = int((idx[-1] + 1) % N)
      mask = torch.as_tensor(it * bs + np.arange(bs) < N, device=dev)
      if isinstance(data_loader, DeviceDataLoader):
        batch = tree_index(data_loader.get_full_data(),
                           torch.as_tensor(idx.astype(np.int32), device=dev))
      else:
        batch = {k: _to_device_array(v[idx], dev) for k, v in
                 data_loader.reference_data.items()}
      info = MiniBatchInformation(observation_count=N, batch_size=bs, mask=mask)
      b = (batch, info) if information else batch
      if masking:
        r, carry = fun(b, mask, carry)
      else:
        r, carry = fun(b, carry)
      results.append(r)
    if masking:
      out = torch.stack([torch.as_tensor(r) for r in results])
    else:
      out = T.tree_map(lambda *xs: torch.cat(list(xs))[:N], *results)
    return CacheState(current_line=line), (out, carry)

  batch_scan.loader = data_loader
  batch_scan.batch_size = bs
  return init_fn, batch_scan, lambda: None
