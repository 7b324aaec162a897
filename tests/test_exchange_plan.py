"""Launch plans of a fused chain table with a replica exchange (host logic, CPU).

The swap kernel (sgmc_resgld_swap) tests a round after an iteration when
min(iterations done, schedule length) is a multiple of swap_every, and otherwise does
nothing unless a pair saves samples. A sparse exchange (no saving pairs) is launched
only after those iterations: the plan of any number of iterations, from any start,
per-iteration or in loop launches, must put a swap launch after exactly the iterations
the kernel's rule selects (checked by brute force over starts and lengths), cut loop
launches at them, and launch after every iteration when a pair saves."""
import itertools

import pytest

from jax_sgmc_amd import fused


def engine(S, L, sparse=True, mh=False):
  eng = object.__new__(fused.FusedGLM)
  eng.table_len = L
  eng.host_iter = 0
  eng.mh = object() if mh else None
  ex = object.__new__(fused.ReplicaExchange)
  ex.engine, ex.swap_every, ex.sparse = eng, S, sparse
  eng.exchange = ex
  return eng, ex


def device_tests(start, n, S, L):
  """Iteration counts after which the swap kernel tests a round."""
  return [d for d in range(start + 1, start + n + 1) if min(d, L) % S == 0]


def swaps_of(ops, start):
  done, out = start, []
  for op in ops:
    if op[0] == "it":
      done += 1
    elif op[0] == "loop":
      done += op[1]
    elif op[0] == "swap":
      out.append(done)
  return done, out


@pytest.mark.parametrize("S,L", [(3, 12), (3, 10), (5, 5), (1, 4), (50, 5000), (4, 1)])
def test_sparse_plans_follow_the_kernel_rule(S, L):
  eng, _ = engine(S, L)
  for start, n in itertools.product(range(0, L + 8), (1, 2, 3, 7, 25)):
    eng.host_iter = start
    for loop in (False, True):
      ops = eng._plan(n, loop)
      done, swaps = swaps_of(ops, start)
      assert done == start + n
      assert swaps == device_tests(start, n, S, L), (start, n, loop, ops)
      if loop:
        # a loop launch never runs past an iteration that needs the swap kernel
        d = start
        for op in ops:
          if op[0] == "loop":
            inside = [t for t in device_tests(d, op[1], S, L) if t != d + op[1]]
            assert not inside, (start, n, ops)
            d += op[1]
      else:
        assert sum(op[0] == "it" for op in ops) == n


def test_dense_and_mh_plans():
  eng, ex = engine(3, 12, sparse=False)
  eng.host_iter = 2
  ops = eng._plan(4, False)
  assert ops == [("it",), ("swap",)] * 4
  assert ex.iterations_to_next_launch(5) == 1
  eng, _ = engine(3, 12, mh=True)
  ops = eng._plan(3, False)
  assert ops == [("it",), ("mh",), ("it",), ("mh",), ("it",), ("swap",), ("mh",)]


def test_loop_chunks_without_exchange():
  eng, _ = engine(3, 12)
  eng.exchange = None
  assert eng._plan(2500, True) == [("loop", 1000), ("loop", 1000), ("loop", 500)]
