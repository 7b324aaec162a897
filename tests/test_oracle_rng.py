"""Pins the oracle's jax.random restatement (CPU).

Known answers used:
- Random123 threefry2x32 (20 rounds) known-answer vectors (kat_vectors of the
  Random123 distribution).
- Values printed in JAX's public documentation for its default (legacy) PRNG:
  split(PRNGKey(0)), normal(PRNGKey(0), (1,)), normal(PRNGKey(0), (10,)),
  normal of the first subkey, uniform(PRNGKey(0)).
- The reference's own doctests, which are functions of normal(PRNGKey(0), (100, 5))
  and of numpy's PCG64 host indices: docs/usage/potential.rst:27-118, 159-164 and
  docs/usage/data.rst:85-89, 116-134.
"""
import numpy as np
import pytest
from scipy import special

from oracle import jaxrandom as jr


@pytest.mark.parametrize("key,ctr,expected", [
    ((0x00000000, 0x00000000), (0x00000000, 0x00000000), (0x6b200159, 0x99ba4efe)),
    ((0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff), (0x1cb996fc, 0xbb002be7)),
    ((0x13198a2e, 0x03707344), (0x243f6a88, 0x85a308d3), (0xc4923a9c, 0x483df7a0)),
])
def test_threefry_random123_kat(key, ctr, expected):
  y0, y1 = jr.threefry2x32_block(key, [ctr[0]], [ctr[1]])
  assert (int(y0[0]), int(y1[0])) == expected


def test_threefry_kat_fails_on_planted_rotation_error(monkeypatch):
  # A wrong rotation constant must break the KAT (the test above is not vacuous).
  monkeypatch.setattr(jr, "_ROTATIONS", ((13, 15, 26, 6), (17, 29, 16, 25)))
  y0, y1 = jr.threefry2x32_block((0, 0), [0], [0])
  assert (int(y0[0]), int(y1[0])) != (0x6b200159, 0x99ba4efe)


def test_jax_documented_split_and_draws():
  k = jr.PRNGKey(0)
  np.testing.assert_array_equal(k, [0, 0])
  s = jr.split(k)
  np.testing.assert_array_equal(s, [[4146024105, 967050713], [2718843009, 1272950319]])
  np.testing.assert_array_equal(jr.split(s[0]), [[2384771982, 3928867769],
                                                 [1278412471, 2182328957]])
  assert jr.normal(k, (1,))[0] == np.float32(-0.20584226)
  assert jr.normal(s[1], (1,))[0] == np.float32(-1.2515389)
  ref10 = np.array([-0.3721109, 0.26423115, -0.18252768, -0.7368197, -0.44030377,
                    -0.1521442, -0.67135346, -0.5908641, 0.73168886, 0.5673026], np.float32)
  np.testing.assert_array_equal(jr.normal(k, (10,)), ref10)
  assert jr.uniform(k) == np.float32(0.41845703)


def _host_indices(seed, N, mb, draws):
  rng = np.random.default_rng(np.random.SeedSequence(seed).spawn(1)[0])
  return [rng.choice(np.arange(N), size=mb, replace=True) for _ in range(draws)]


def test_reference_doctest_potential_values():
  """docs/usage/potential.rst: minibatch U rounds to 838, var(lik) to 7, full U to
  707, for mean = normal(PRNGKey(0), (100, 5)), sample mean 0 / std 1, batch of 5
  from the host loader with seed 0 (chain id 0)."""
  mean = jr.normal(jr.PRNGKey(0), (100, 5)).astype(np.float64)
  idx = _host_indices(0, 100, 5, 1)[0]

  def lik(x):
    return np.sum((np.log(2 * np.pi) + x ** 2) / -2, axis=-1)

  l = lik(mean[idx])
  assert round(-100 * np.mean(l)) == 838
  assert round(np.var(l)) == 7
  assert round(-np.sum(lik(mean))) == 707


def test_reference_doctest_host_indices():
  """docs/usage/data.rst:85-89 (seed 0, batch 2 of 10 -> [8, 9]) and :116-134
  (fourth draw of batch 3 -> [4, 6, 6])."""
  np.testing.assert_array_equal(_host_indices(0, 10, 2, 1)[0], [8, 9])
  np.testing.assert_array_equal(_host_indices(0, 10, 3, 4)[3], [4, 6, 6])


def test_split_layout_against_threefry_definition():
  k = np.array([123, 456], np.uint32)
  for num in (1, 2, 3, 4, 7):
    s = jr.split(k, num)
    ctr = np.arange(2 * num, dtype=np.uint32)
    y0, y1 = jr.threefry2x32_block(k, ctr[:num], ctr[num:])
    flat = np.concatenate([y0, y1])
    np.testing.assert_array_equal(s.ravel(), flat)


def test_random_bits_odd_padding():
  k = np.array([5, 6], np.uint32)
  for S in (1, 3, 5, 9):
    b = jr.random_bits(k, (S,))
    m = (S + 1) // 2
    x1 = np.concatenate([np.arange(m, S), [0]]).astype(np.uint32)[:m]
    y0, y1 = jr.threefry2x32_block(k, np.arange(m, dtype=np.uint32), x1)
    np.testing.assert_array_equal(b, np.concatenate([y0, y1])[:S])


def test_erfinv_against_scipy():
  x = np.concatenate([np.linspace(-0.999999, 0.999999, 20001),
                      1 - np.logspace(-7, -1, 200), -(1 - np.logspace(-7, -1, 200))])
  x = x.astype(np.float32)
  got = jr.erfinv_f32(x).astype(np.float64)
  ref = special.erfinv(x.astype(np.float64))
  rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
  inner = np.abs(x) <= 0.99
  # well-conditioned range: a few float32 ulp; near +-1 the float32 argument
  # 1 - x*x loses digits (as in the float32 formula being restated)
  assert np.max(rel[inner]) < 1e-6, np.max(rel[inner])
  assert np.max(rel) < 1e-5, np.max(rel)
  assert np.median(rel) < 2e-7


def test_erfinv_planted_coefficient_error_is_caught(monkeypatch):
  bad = list(jr._ERFINV_LT5)
  bad[7] = np.float32(0.2466)  # one coefficient, fourth digit
  monkeypatch.setattr(jr, "_ERFINV_LT5", bad)
  x = np.linspace(-0.9, 0.9, 101).astype(np.float32)
  rel = np.abs(jr.erfinv_f32(x) - special.erfinv(x.astype(np.float64)))
  assert np.max(rel) > 1e-5


def test_logf_log1p_against_numpy():
  x = np.exp(np.linspace(-80, 80, 20001)).astype(np.float32)
  got = jr.logf_f32(x).astype(np.float64)
  ref = np.log(x.astype(np.float64))
  ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
  assert np.max(np.abs(got - ref) / np.maximum(ulp, 1e-45)) <= 2.0
  t = np.linspace(-0.99, 3.0, 20001).astype(np.float32)
  got = jr.log1p_f32(t).astype(np.float64)
  ref = np.log1p(t.astype(np.float64))
  assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)) < 3e-7


def test_normal_moments_and_uniform_range():
  z = jr.normal(np.array([1, 2], np.uint32), (200000,)).astype(np.float64)
  assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01
  u = jr.uniform(np.array([3, 4], np.uint32), (100000,))
  assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.005


def _randint_bruteforce(key, n, lo, hi):
  k1, k2 = jr.split(key)
  h = [int(v) for v in jr.random_bits(k1, (n,))]
  l = [int(v) for v in jr.random_bits(k2, (n,))]
  span = (hi - lo) if hi > lo else 1
  mult = ((65536 % span) ** 2 % 2 ** 32) % span
  return [lo + (((a % span) * mult + (b % span)) % 2 ** 32) % span for a, b in zip(h, l)]


@pytest.mark.parametrize("lo,hi", [(0, 10), (0, 10000), (0, 65536), (0, 65537), (0, 10 ** 6),
                                   (-3, 4), (5, 5)])
def test_randint_definition(lo, hi):
  key = np.array([17, 29], np.uint32)
  np.testing.assert_array_equal(jr.randint(key, (257,), lo, hi),
                                _randint_bruteforce(key, 257, lo, hi))


def test_randint_uniformity():
  idx = jr.randint(np.array([1, 1], np.uint32), (100000,), 0, 10000)
  assert idx.min() >= 0 and idx.max() < 10000
  counts = np.bincount(idx // 100, minlength=100)
  chi2 = np.sum((counts - 1000.0) ** 2 / 1000.0)
  assert chi2 < 160  # 99 dof, p ~ 1e-4
