"""Minimal pytree utilities with jax's flattening order.

Leaves of a dict are visited in sorted key order, lists/tuples in order, NamedTuples
field by field, None is an empty subtree (jax.tree_util semantics as the reference
relies on them: random_tree's leaf order, integrator.py:130-134, and ravel_pytree
in the adaption decorator, adaption.py:103, 132-136).
"""
import numpy as np
import torch


def _is_namedtuple(x):
  return isinstance(x, tuple) and hasattr(x, "_fields")


def flatten(tree):
  """Returns (leaves, treedef)."""
  if tree is None:
    return [], ("none",)
  if isinstance(tree, dict):
    keys = sorted(tree.keys())
    leaves, defs = [], []
    for k in keys:
      l, d = flatten(tree[k])
      leaves += l
      defs.append((k, d, len(l)))
    return leaves, ("dict", defs)
  if _is_namedtuple(tree):
    leaves, defs = [], []
    for v in tree:
      l, d = flatten(v)
      leaves += l
      defs.append((d, len(l)))
    return leaves, ("namedtuple", type(tree), defs)
  if isinstance(tree, (list, tuple)):
    leaves, defs = [], []
    for v in tree:
      l, d = flatten(v)
      leaves += l
      defs.append((d, len(l)))
    return leaves, ("list" if isinstance(tree, list) else "tuple", defs)
  return [tree], ("leaf",)


def unflatten(treedef, leaves):
  leaves = list(leaves)
  out, rest = _unflatten(treedef, leaves)
  assert not rest
  return out


def _unflatten(td, leaves):
  kind = td[0]
  if kind == "none":
    return None, leaves
  if kind == "leaf":
    return leaves[0], leaves[1:]
  if kind == "dict":
    out = {}
    for k, d, _ in td[1]:
      out[k], leaves = _unflatten(d, leaves)
    return out, leaves
  if kind == "namedtuple":
    vals = []
    for d, _ in td[2]:
      v, leaves = _unflatten(d, leaves)
      vals.append(v)
    return td[1](*vals), leaves
  vals = []
  for d, _ in td[1]:
    v, leaves = _unflatten(d, leaves)
    vals.append(v)
  return (vals if kind == "list" else tuple(vals)), leaves


def tree_map(fn, tree, *rest):
  leaves, td = flatten(tree)
  others = [flatten(r)[0] for r in rest]
  return unflatten(td, [fn(*xs) for xs in zip(leaves, *others)])


def as_tensor(x, device=None, dtype=torch.float32):
  if isinstance(x, torch.Tensor):
    return x.to(device=device, dtype=dtype) if device is not None else x.to(dtype)
  return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


class Layout:
  """Flat-vector layout of a pytree of float tensors (ravel_pytree order)."""

  def __init__(self, tree):
    leaves, self.treedef = flatten(tree)
    self.shapes = [tuple(torch.as_tensor(l).shape) if not isinstance(l, torch.Tensor)
                   else tuple(l.shape) for l in leaves]
    self.sizes = [int(np.prod(s)) if len(s) else 1 for s in self.shapes]
    self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
    self.size = int(self.offsets[-1])
    self._tables = {}

  @property
  def nleaves(self):
    return len(self.sizes)

  def ravel(self, tree, device=None):
    leaves, _ = flatten(tree)
    parts = [as_tensor(l, device).reshape(-1) for l in leaves]
    if not parts:
      return torch.zeros(0, device=device)
    return torch.cat(parts).contiguous()

  def unravel(self, flat):
    leaves = [flat[self.offsets[i]:self.offsets[i + 1]].reshape(self.shapes[i])
              for i in range(self.nleaves)]
    return unflatten(self.treedef, leaves)

  def leaf_table_host(self, pairs_per_block):
    """[L, B, (offset, size) x L, (leaf, first_pair) x B] as int64."""
    blocks = []
    for l, s in enumerate(self.sizes):
      m = (s + 1) // 2
      for p in range(0, m, pairs_per_block):
        blocks.append((l, p))
    tbl = [self.nleaves, len(blocks)]
    for o, s in zip(self.offsets[:-1], self.sizes):
      tbl += [int(o), int(s)]
    for l, p in blocks:
      tbl += [l, p]
    return np.asarray(tbl, dtype=np.int64), len(blocks)

  def leaf_table(self, device, pairs_per_block):
    key = (str(device), pairs_per_block)
    if key not in self._tables:
      host, nb = self.leaf_table_host(pairs_per_block)
      self._tables[key] = (torch.as_tensor(host, device=device), nb)
    return self._tables[key]
