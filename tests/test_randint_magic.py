"""The index draws reduce Threefry words modulo the dataset size without a division
(csrc/sgmc_device.h, mod_magic): r = a - umulhi(a, floor(2^32 / d)) d, minus d once if
r >= d. This checks that reduction against Python's % on divisors of every magnitude,
including powers of two, 2^31 +- 1 and 2^32 - 1, and the extreme dividends."""
import random

import pytest


def _magic(d):
  return (1 << 32) // d if d >= 2 else 0


def _mod_magic(a, d, m):
  if d < 2:
    return 0
  r = (a - ((a * m) >> 32) * d) & 0xFFFFFFFF
  return r - d if r >= d else r


@pytest.mark.parametrize("seed", [0, 1])
def test_mod_magic_equals_remainder(seed):
  rng = random.Random(seed)
  divisors = [1, 2, 3, 5, 7, 10, 127, 128, 1000, 10_000, 50_000, 1_000_000, 3_000_017,
              2**31 - 1, 2**31, 2**31 + 1, 2**32 - 1] + [rng.randrange(1, 2**32) for _ in range(500)]
  for d in divisors:
    m = _magic(d)
    q = (2**32 - 1) // d
    for a in [0, 1, d - 1, d, d + 1, q * d, q * d - 1, 2**32 - 1, 2**32 - 2] + \
             [rng.randrange(0, 2**32) for _ in range(100)]:
      a %= 2**32
      assert _mod_magic(a, d, m) == a % d, (a, d)
