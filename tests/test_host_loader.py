"""Host data pipeline of NumpyDataLoader (data/numpy_loader.py:147-460): the index
streams of the random (with replacement, shuffle, shuffle in epochs) and ordered
pipelines against the oracle's whole-stream restatement, the oracle pinned by the
permutation structure of each mode, and the reference's own loader tests restated
(tests/test_data.py:62-371): batch shapes and dtypes, duplicate counts of a
shuffle, every element once per epoch, cache-size invariance, checkpoints."""
import numpy as np
import pytest
import torch

from oracle import loader as ol

import __graft_entry__

__graft_entry__.load()
from jax_sgmc_amd import data as DT  # noqa: E402


def _loader(N=13):
  return DT.NumpyDataLoader(device="cpu", ordered_indices=np.arange(N),
                            x=np.arange(2 * N, dtype=np.float64).reshape(N, 2),
                            y=np.zeros((N, 4, 3), np.float32))


def _stream(dl, cid, batches_per_call, calls):
  idx, msk = [], []
  for _ in range(calls):
    b, m = dl.get_batches(cid)
    idx.append(b["ordered_indices"].numpy().reshape(batches_per_call, -1))
    msk.append(m.numpy().reshape(batches_per_call, -1))
  return np.concatenate(idx), np.concatenate(msk)


# ------------------------------------------------------------- oracle pins ----

@pytest.mark.parametrize("N,n", [(13, 3), (12, 3), (10, 10), (7, 1)])
def test_oracle_shuffle_is_a_sequence_of_permutations(N, n):
  idx, mask = ol.index_stream(N, n, 40, "shuffle", seed=5)
  flat = idx.reshape(-1)
  assert mask.all()
  for k in range(flat.size // N):
    assert sorted(flat[k * N:(k + 1) * N]) == list(range(N))


@pytest.mark.parametrize("N,n", [(13, 3), (12, 3), (10, 10), (7, 1)])
def test_oracle_epochs_each_index_once_per_epoch(N, n):
  per = -(-N // n)
  idx, mask = ol.index_stream(N, n, 3 * per, "epochs", seed=2)
  for e in range(3):
    sl = slice(e * per, (e + 1) * per)
    assert sorted(idx[sl][mask[sl]]) == list(range(N))
    assert mask[sl][:-1].all() and mask[sl][-1].sum() == N - (per - 1) * n
  first = np.random.default_rng(np.random.SeedSequence(2).spawn(1)[0]).choice(
      np.arange(N), size=N, replace=False)
  np.testing.assert_array_equal(idx[:per].reshape(-1)[:N], first)


def test_oracle_ordered_wraps_and_masks():
  idx, mask = ol.index_stream(7, 3, 4, "ordered")
  np.testing.assert_array_equal(idx, [[0, 1, 2], [3, 4, 5], [6, 0, 1], [0, 1, 2]])
  np.testing.assert_array_equal(mask, [[1, 1, 1], [1, 1, 1], [1, 0, 0], [1, 1, 1]])


# ----------------------------------------------------- loader against oracle ----

@pytest.mark.parametrize("mode", ["replace", "shuffle", "epochs"])
@pytest.mark.parametrize("N,n,cache", [(13, 3, 1), (13, 3, 5), (12, 4, 2), (9, 9, 3)])
def test_random_pipeline_streams_match_oracle(mode, N, n, cache):
  dl = _loader(N)
  cid = dl.register_random_pipeline(cache_size=cache, mb_size=n, seed=11,
                                    shuffle=mode != "replace", in_epochs=mode == "epochs")
  calls = 12
  idx, mask = _stream(dl, cid, cache, calls)
  ridx, rmask = ol.index_stream(N, n, cache * calls, mode, seed=11)
  np.testing.assert_array_equal(idx, ridx)
  np.testing.assert_array_equal(mask, rmask)


def test_ordered_pipeline_matches_oracle():
  dl = _loader(13)
  cid = dl.register_ordered_pipeline(cache_size=2, mb_size=5)
  idx, mask = _stream(dl, cid, 2, 6)
  ridx, rmask = ol.index_stream(13, 5, 12, "ordered")
  np.testing.assert_array_equal(idx, ridx)
  np.testing.assert_array_equal(mask, rmask)


def test_default_seed_is_registration_order():
  dl = _loader(13)
  dl.register_random_pipeline(cache_size=1, mb_size=3)
  cid = dl.register_random_pipeline(cache_size=1, mb_size=3, shuffle=True)
  idx, _ = _stream(dl, cid, 1, 5)
  np.testing.assert_array_equal(idx, ol.index_stream(13, 3, 5, "shuffle", seed=1)[0])


# ------------------------------------------ the reference's tests, restated ----

@pytest.mark.parametrize("shuffle,in_epochs", [(False, False), (True, False), (True, True)])
def test_batch_shapes_and_dtypes(shuffle, in_epochs):
  dl = _loader()
  cid = dl.register_random_pipeline(cache_size=5, mb_size=7, shuffle=shuffle,
                                    in_epochs=in_epochs)
  b, m = dl.get_batches(cid)
  assert b["x"].shape == (5, 7, 2) and b["x"].dtype == torch.float32
  assert b["y"].shape == (5, 7, 4, 3)
  assert b["ordered_indices"].dtype == torch.int32
  assert m.shape == (5, 7)
  init = dl.initializer_batch(7)
  assert init["x"].shape == (7, 2) and init["x"].dtype == torch.float32
  assert init["y"].shape == (7, 4, 3) and init["ordered_indices"].dtype == torch.int32


def test_shuffle_duplicates_like_the_reference():
  N, n = 13, 3
  per = -(-N // n)
  dup = per * n - N
  dl = _loader(N)
  a = dl.register_random_pipeline(cache_size=per, mb_size=n, shuffle=True)
  b = dl.register_random_pipeline(cache_size=1, mb_size=n, shuffle=True)
  sa = dl.get_batches(a)[0]["ordered_indices"].numpy()
  sb = np.array([dl.get_batches(b)[0]["ordered_indices"].numpy() for _ in range(per)])
  for s in (sa, sb):
    _, cnt = np.unique(s, return_counts=True)
    assert cnt.size == N and np.sum(cnt == 1) == N - dup and np.sum(cnt == 2) == dup


def test_in_epochs_every_element_once():
  N, n = 13, 3
  per = -(-N // n)
  dl = _loader(N)
  a = dl.register_random_pipeline(cache_size=per, mb_size=n, shuffle=True, in_epochs=True)
  sa, ma = dl.get_batches(a)
  _, cnt = np.unique(sa["ordered_indices"].numpy()[ma.numpy()], return_counts=True)
  assert np.sum(cnt == 1) == N


def test_in_epochs_needs_shuffle():
  with pytest.raises(ValueError):
    _loader().register_random_pipeline(cache_size=1, mb_size=2, in_epochs=True)


@pytest.mark.parametrize("shuffle,in_epochs", [(False, False), (True, False), (True, True)])
def test_cache_size_does_not_change_the_order(shuffle, in_epochs):
  dl = _loader()
  small = dl.register_random_pipeline(3, 2, seed=0, shuffle=shuffle, in_epochs=in_epochs)
  big = dl.register_random_pipeline(5, 2, seed=0, shuffle=shuffle, in_epochs=in_epochs)
  sb, _ = dl.get_batches(small)
  bb, _ = dl.get_batches(big)
  for k in sb:
    np.testing.assert_array_equal(sb[k].numpy(), bb[k].numpy()[:3])


@pytest.mark.parametrize("shuffle,in_epochs", [(False, False), (True, False), (True, True)])
def test_checkpoint_random_pipeline(shuffle, in_epochs):
  N = 13
  dl = _loader(N)
  a = dl.register_random_pipeline(shuffle=shuffle, in_epochs=in_epochs, cache_size=3,
                                  mb_size=2)
  for _ in range(N // 6):
    dl.get_batches(a)
  ck = dl.save_state(a)
  after = dl.get_batches(a)
  again = dl.get_batches(a)
  new = dl.register_random_pipeline(cache_size=3, mb_size=2, shuffle=shuffle,
                                    in_epochs=in_epochs)
  dl.load_state(new, ck)
  for ref in (after, again):
    got = dl.get_batches(new)
    for k in ref[0]:
      np.testing.assert_array_equal(got[0][k].numpy(), ref[0][k].numpy())
    np.testing.assert_array_equal(got[1].numpy(), ref[1].numpy())


def test_checkpoint_ordered_pipeline():
  dl = _loader(13)
  a = dl.register_ordered_pipeline(cache_size=3, mb_size=2)
  dl.get_batches(a)
  ck = dl.save_state(a)
  after = dl.get_batches(a)
  new = dl.register_ordered_pipeline(cache_size=3, mb_size=2)
  dl.load_state(new, ck)
  got = dl.get_batches(new)
  np.testing.assert_array_equal(got[0]["ordered_indices"].numpy(),
                                after[0]["ordered_indices"].numpy())
