"""Index streams of the host data loader, restated (oracle only).

TEST INFRASTRUCTURE. NumpyDataLoader (data/numpy_loader.py:231-460) draws its indices
from numpy's PCG64 seeded with SeedSequence(seed).spawn(1)[0] (seed = the chain's
registration order unless given). Restated here as whole streams rather than per-call
state machines:
- with replacement: each batch is rng.choice(arange(N), n, replace=True)
  (numpy_loader.py:385-391);
- shuffle: the batches are consecutive n-chunks of perm_1 ++ perm_2 ++ ..., perm_k =
  rng.choice(arange(N), N, replace=False) drawn when the previous ones run out
  (numpy_loader.py:393-433, whose ring buffer of (floor(N/n) + 2) n slots is refilled
  right after the unread entries);
- shuffle in epochs: each epoch is one permutation padded to ceil(N/n) n entries; the
  padding holds index 0 and is masked (numpy_loader.py:435-460);
- ordered: arange(n) + offset, wrapped mod N, masked past N, offset reset to 0 after
  the batch that reaches N (numpy_loader.py:358-371).
"""
import numpy as np


def _rng(seed):
  return np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])


def index_stream(N, n, batches, mode="replace", seed=0):
  """(idx [batches, n], mask [batches, n]) for mode in replace, shuffle, epochs,
  ordered."""
  idx = np.zeros((batches, n), np.int64)
  mask = np.ones((batches, n), bool)
  if mode == "replace":
    rng = _rng(seed)
    for b in range(batches):
      idx[b] = rng.choice(np.arange(N), size=n, replace=True)
  elif mode == "shuffle":
    rng = _rng(seed)
    stream = np.zeros(0, np.int64)
    for b in range(batches):
      while stream.size < (b + 1) * n:
        stream = np.concatenate([stream, rng.choice(np.arange(N), size=N, replace=False)])
      idx[b] = stream[b * n:(b + 1) * n]
  elif mode == "epochs":
    rng = _rng(seed)
    per = -(-N // n)
    epoch = None
    for b in range(batches):
      k = b % per
      if k == 0:
        epoch = np.zeros(per * n, np.int64)
        epoch[:N] = rng.choice(np.arange(N), size=N, replace=False)
      idx[b] = epoch[k * n:(k + 1) * n]
      mask[b] = np.arange(k * n, (k + 1) * n) < N
  elif mode == "ordered":
    off = 0
    for b in range(batches):
      raw = np.arange(n) + off
      idx[b] = np.mod(raw, N)
      mask[b] = raw < N
      off = 0 if off + n >= N else off + n
  else:
    raise ValueError(mode)
  return idx, mask
