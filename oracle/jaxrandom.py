"""numpy restatement of jax.random's legacy Threefry-2x32 derivation (oracle only).

The reference draws every random number of the hot path through jax.random
(third-party; jax 0.4.x, whose default is the original, non-partitionable Threefry
derivation). Call sites: noise in ``random_tree`` (integrator.py:119-135), key
splits in every integrator (integrator.py:208, 630, 736, 871), minibatch indices in
``DeviceNumpyDataLoader.get_random_data`` (data/numpy_loader.py:132-134), the reSGLD
swap uniform (solver.py:283-284) and the SGGMC slice uniform (solver.py:508, 528).

Everything here is written op by op in float32/uint32 numpy arithmetic, which rounds
after every operation (no fused multiply-add), so the result of each function is a
fixed function of its inputs.

Pinning:
- ``threefry2x32_block``: Random123 known-answer vectors for threefry2x32-20
  (tests/test_oracle_rng.py).
- ``split``/``random_bits``/``uniform``/``normal``: the JAX documentation's printed
  key values for ``split(PRNGKey(0))`` and the reference's own doctest values, which
  depend on ``normal(PRNGKey(0), (100, 5))`` (docs/usage/potential.rst:27, 100-101,
  117-118, 163-164). ``erfinv_f32`` is checked against scipy's float64 erfinv, and
  ``logf_f32``/``log1p_f32`` against numpy's float64 log/log1p.
- ``randint``: exact modular construction; checked against a brute-force
  restatement of its definition and for uniformity. The exact jax bit-for-bit
  value of an index draw has no in-tree known answer: parity unpinned beyond the
  construction itself.
"""
import numpy as np

U32 = np.uint32
F32 = np.float32

_ROTATIONS = ((13, 15, 26, 6), (17, 29, 16, 24))
_PARITY = 0x1BD11BDA


def _u32(a):
  return np.asarray(a, dtype=np.uint32)


def threefry2x32_block(key, x0, x1):
  """Threefry-2x32 with 20 rounds on counter pairs (x0[i], x1[i]) under ``key``.

  The key schedule is (k0, k1, k0 ^ k1 ^ 0x1BD11BDA); a key injection follows every
  four rounds, the i-th adding ks[i % 3], ks[(i + 1) % 3] + i.
  """
  k0 = _u32(key[0])
  k1 = _u32(key[1])
  ks = (k0, k1, k0 ^ k1 ^ U32(_PARITY))
  x0 = _u32(x0).copy()
  x1 = _u32(x1).copy()
  with np.errstate(over="ignore"):
    x0 = x0 + ks[0]
    x1 = x1 + ks[1]
    for group in range(5):
      for r in _ROTATIONS[group % 2]:
        x0 = x0 + x1
        x1 = (x1 << U32(r)) | (x1 >> U32(32 - r))
        x1 = x1 ^ x0
      x0 = x0 + ks[(group + 1) % 3]
      x1 = x1 + ks[(group + 2) % 3] + U32(group + 1)
  return x0, x1


def threefry_2x32(key, count):
  """jax's ``threefry_2x32(keypair, count)``: the counts are flattened, padded with
  one zero if their number is odd, split into halves (x0, x1), hashed pairwise, and
  the two output halves are concatenated and trimmed to the input length."""
  count = _u32(count).ravel()
  n = count.size
  if n % 2:
    count = np.concatenate([count, _u32([0])])
  m = count.size // 2
  y0, y1 = threefry2x32_block(key, count[:m], count[m:])
  out = np.concatenate([y0, y1])
  return out[:n]


def PRNGKey(seed):
  """``jax.random.PRNGKey(seed)`` for 0 <= seed < 2**31: the pair (0, seed)."""
  seed = int(seed)
  if not 0 <= seed < 2 ** 31:
    raise ValueError("seed must lie in [0, 2**31)")
  return _u32([0, seed])


def split(key, num=2):
  """``jax.random.split(key, num)``: threefry_2x32(key, iota(2 num)) as (num, 2)."""
  counts = np.arange(2 * num, dtype=np.uint32)
  return threefry_2x32(key, counts).reshape(num, 2)


def random_bits(key, shape):
  """32-bit ``_random_bits(key, 32, shape)``: threefry_2x32(key, iota(size))."""
  shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
  size = int(np.prod(shape)) if shape else 1
  if size >= 2 ** 32 - 1:
    raise NotImplementedError("more than 2**32 - 2 words per key")
  return threefry_2x32(key, np.arange(size, dtype=np.uint32)).reshape(shape)


def bits_to_unit_float(bits):
  """Mantissa construction of ``uniform``: bitcast((bits >> 9) | 0x3F800000) - 1."""
  fb = (_u32(bits) >> U32(9)) | U32(0x3F800000)
  return fb.view(np.float32) - F32(1.0)


def uniform(key, shape=(), minval=0.0, maxval=1.0):
  """``jax.random.uniform`` in float32: max(lo, f * (hi - lo) + lo)."""
  lo = F32(minval)
  hi = F32(maxval)
  f = bits_to_unit_float(random_bits(key, shape if shape else (1,)))
  out = np.maximum(lo, f * (hi - lo) + lo).astype(np.float32)
  return out.reshape(shape) if shape else out.reshape(())


# --- float32 elementary functions, each a published method restated op by op ---

# Cephes single-precision logf polynomial (S. Moshier, Cephes Math Library), in the
# arrangement Eigen/XLA use for float32 log on the CPU.
_LOG_P = [F32(c) for c in (7.0376836292e-2, -1.1514610310e-1, 1.1676998740e-1,
                           -1.2420140846e-1, 1.4249322787e-1, -1.6668057665e-1,
                           2.0000714765e-1, -2.4999993993e-1, 3.3333331174e-1)]
_LOG_Q1 = F32(-2.12194440e-4)
_LOG_Q2 = F32(0.693359375)
_SQRTHF = F32(0.707106781186547524)


def logf_f32(x):
  """Natural log of positive normal float32 values (Cephes logf).

  x = 2**e * m with m in [0.5, 1); if m < sqrt(1/2) the mantissa is doubled and e
  decremented; log x = t - t^2/2 + t^3 P(t) + e*(q1 + q2) with t = m - 1.
  """
  x = np.asarray(x, dtype=np.float32)
  xb = x.view(np.uint32)
  e = ((xb >> U32(23)).astype(np.int32) - 126).astype(np.float32)
  m = ((xb & U32(0x807FFFFF)) | U32(0x3F000000)).view(np.float32)
  small = m < _SQRTHF
  t = m - F32(1.0)
  t = np.where(small, t + m, t).astype(np.float32)
  e = np.where(small, e - F32(1.0), e).astype(np.float32)
  x2 = t * t
  x3 = x2 * t
  p = _LOG_P
  y = p[0] * t + p[1]
  y1 = p[3] * t + p[4]
  y2 = p[6] * t + p[7]
  y = y * t + p[2]
  y1 = y1 * t + p[5]
  y2 = y2 * t + p[8]
  y = y * x3 + y1
  y = y * x3 + y2
  y = y * x3
  y = y + _LOG_Q1 * e
  r = t - F32(0.5) * x2
  r = r + y
  r = r + _LOG_Q2 * e
  return r.astype(np.float32)


# Cephes log1p rational approximation (unity.c, LP/LQ coefficients), used for
# |x| < sqrt(2) - 1; larger arguments take log(1 + x).
_L1P_NUM = [F32(c) for c in (4.5270000862445199635215e-5, 4.9854102823193375972212e-1,
                             6.5787325942061044846969e0, 2.9911919328553073277375e1,
                             6.0949667980987787057556e1, 5.7112963590585538103336e1,
                             2.0039553499201281259648e1)]
_L1P_DEN = [F32(c) for c in (1.0, 1.5062909083469192043167e1, 8.3047565967967209469434e1,
                             2.2176239823732856465394e2, 3.0909872225312059774938e2,
                             2.1642788614495947685003e2, 6.0118660497603843919306e1)]
_L1P_SMALL = F32(0.41421356237309504880)


def _horner(coeffs, x):
  acc = np.zeros_like(x)
  for c in coeffs:
    acc = acc * x + c
  return acc.astype(np.float32)


def log1p_f32(x):
  x = np.asarray(x, dtype=np.float32)
  large = logf_f32(x + F32(1.0))
  x2 = x * x
  ratio = _horner(_L1P_NUM, x) / _horner(_L1P_DEN, x)
  small = x + ((F32(-0.5) * x2) + ((x * x2) * ratio))
  return np.where(np.abs(x) < _L1P_SMALL, small, large).astype(np.float32)


# Giles (2010), "Approximating the erfinv function", single-precision branch
# coefficients, in highest-degree-first order.
_ERFINV_LT5 = [F32(c) for c in (2.81022636e-08, 3.43273939e-07, -3.5233877e-06,
                                -4.39150654e-06, 0.00021858087, -0.00125372503,
                                -0.00417768164, 0.246640727, 1.50140941)]
_ERFINV_GE5 = [F32(c) for c in (-0.000200214257, 0.000100950558, 0.00134934322,
                                -0.00367342844, 0.00573950773, -0.0076224613,
                                0.00943887047, 1.00167406, 2.83297682)]


def erfinv_f32(x):
  """Inverse error function, float32: w = -log1p(-x*x); polynomial in w - 2.5
  (w < 5) or sqrt(w) - 3 (w >= 5), times x; +-inf at x = +-1."""
  x = np.asarray(x, dtype=np.float32)
  w = -log1p_f32(x * (-x))
  lt = w < F32(5.0)
  with np.errstate(invalid="ignore"):
    w = np.where(lt, w - F32(2.5), np.sqrt(w) - F32(3.0)).astype(np.float32)
  p = np.where(lt, _ERFINV_LT5[0], _ERFINV_GE5[0]).astype(np.float32)
  for a, b in zip(_ERFINV_LT5[1:], _ERFINV_GE5[1:]):
    p = np.where(lt, a, b).astype(np.float32) + p * w
  res = (p * x).astype(np.float32)
  with np.errstate(invalid="ignore"):
    edge = (x * F32(np.inf)).astype(np.float32)
  return np.where(np.abs(x) == F32(1.0), edge, res).astype(np.float32)


_SQRT2 = F32(np.sqrt(2.0))
_NORMAL_LO = np.nextafter(F32(-1.0), F32(0.0), dtype=np.float32)


def bits_to_normal(bits):
  """Standard normal from raw 32-bit words: sqrt(2) * erfinv(u), u uniform on
  [nextafter(-1, 0), 1)."""
  f = bits_to_unit_float(bits)
  u = np.maximum(_NORMAL_LO, f * (F32(1.0) - _NORMAL_LO) + _NORMAL_LO).astype(np.float32)
  return (_SQRT2 * erfinv_f32(u)).astype(np.float32)


def normal(key, shape=()):
  """``jax.random.normal(key, shape)`` in float32."""
  sh = shape if shape else (1,)
  out = bits_to_normal(random_bits(key, sh))
  return out.reshape(shape) if shape else out.reshape(())


def randint(key, shape, minval, maxval):
  """``jax.random.randint(key, shape, minval, maxval)`` for int32 output.

  k1, k2 = split(key); hi, lo = bits(k1), bits(k2); span = maxval - minval (1 if
  empty); mult = ((2**16 % span)**2 mod 2**32) % span;
  offset = ((hi % span) * mult + lo % span) % span, all in wrapping uint32.
  """
  shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
  k1, k2 = split(key)
  hi = random_bits(k1, shape)
  lo = random_bits(k2, shape)
  minval = int(minval)
  maxval = int(maxval)
  span = U32(1) if maxval <= minval else U32((maxval - minval) & 0xFFFFFFFF)
  with np.errstate(over="ignore"):
    mult = _u32(U32(2 ** 16) % span)
    mult = _u32(mult * mult) % span
    off = ((hi % span) * mult + (lo % span)) % span
  return (np.int64(minval) + off.astype(np.int64)).astype(np.int32)


def random_tree_flat(key, leaf_sizes):
  """``random_tree(key, tree)`` (integrator.py:119-135) for a tree whose leaves, in
  tree_flatten order, have the given sizes: split(key, n_leaves), then one
  ``normal`` per leaf. Returns the concatenated flat noise."""
  keys = split(key, len(leaf_sizes))
  parts = [normal(keys[i], (int(s),)) for i, s in enumerate(leaf_sizes)]
  return np.concatenate(parts) if parts else np.zeros(0, np.float32)
