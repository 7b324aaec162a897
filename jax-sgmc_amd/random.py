"""jax.random's PRNG API on the GPU (legacy Threefry-2x32 derivation).

Keys are int32 tensors of shape (2,) holding the uint32 words of a jax key; every
draw runs a HIP kernel of libsgmc (include/sgmc.h, "RNG"). These are the random
numbers of the reference's hot path (integrator.py:119-135, 208, 630, 736, 871;
data/numpy_loader.py:124, 132-134; solver.py:254, 283-284, 483, 508, 528).
"""
import numpy as np
import torch

from . import _lib


def _device(device=None):
  return torch.device(device) if device is not None else torch.device("cuda")


def PRNGKey(seed, device=None):
  """jax.random.PRNGKey(seed) = (0, seed) for 0 <= seed < 2**31."""
  seed = int(seed)
  if not 0 <= seed < 2 ** 31:
    raise ValueError("seed must lie in [0, 2**31)")
  return torch.tensor([0, seed], dtype=torch.int32, device=_device(device))


def key_from_words(words, device=None):
  w = np.asarray(words, dtype=np.uint32).view(np.int32)
  return torch.as_tensor(w.copy(), device=_device(device))


def key_words(key):
  """uint32 numpy view of a key (or of a stack of keys)."""
  return key.detach().cpu().numpy().view(np.uint32)


def split(key, num=2):
  out = torch.empty((num, 2), dtype=torch.int32, device=key.device)
  _lib.call("sgmc_split", _lib.ptr(key), _lib.ptr(out), num, _lib.stream())
  return out


def bits(key, shape):
  shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
  n = int(np.prod(shape)) if shape else 1
  out = torch.empty(n, dtype=torch.int32, device=key.device)
  _lib.call("sgmc_random_bits", _lib.ptr(key), _lib.ptr(out), n, _lib.stream())
  return out.reshape(shape)


def uniform(key, shape=(), minval=0.0, maxval=1.0):
  shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
  n = int(np.prod(shape)) if shape else 1
  out = torch.empty(n, dtype=torch.float32, device=key.device)
  _lib.call("sgmc_uniform", _lib.ptr(key), _lib.ptr(out), n, float(minval), float(maxval),
            _lib.stream())
  return out.reshape(shape)


def normal(key, shape=()):
  shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
  n = int(np.prod(shape)) if shape else 1
  out = torch.empty(n, dtype=torch.float32, device=key.device)
  _lib.call("sgmc_normal", _lib.ptr(key), _lib.ptr(out), n, _lib.stream())
  return out.reshape(shape)


def randint(key, shape, minval, maxval):
  shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
  n = int(np.prod(shape)) if shape else 1
  out = torch.empty(n, dtype=torch.int32, device=key.device)
  _lib.call("sgmc_randint", _lib.ptr(key), _lib.ptr(out), n, int(minval), int(maxval),
            _lib.stream())
  return out.reshape(shape)


def random_tree(key, layout):
  """random_tree(key, tree) (integrator.py:119-135) as one flat float32 vector in
  ravel order; ``layout`` is a tree.Layout."""
  table, nb = layout.leaf_table(key.device, _lib.PAIRS_PER_BLOCK)
  out = torch.empty(layout.size, dtype=torch.float32, device=key.device)
  _lib.call("sgmc_tree_normal", _lib.ptr(key), _lib.ptr(table), nb, _lib.ptr(out),
            _lib.stream())
  return out
