"""Random thinning (scheduler.py:599-666): jax.random.choice(replace=False, p) draws by
the Gumbel top-k construction, which samples without replacement with probability
proportional to p (successive, Plackett-Luce sampling). Pinned on the CPU by the exact
inclusion probabilities of that sampling over 20,000 keys (uniforms from the oracle's
jax-compatible generator); a sign mistake in the construction fails the same check.
Zero-probability iterations (burn-in) are never selected."""
import numpy as np

from oracle import jaxrandom as jr

import __graft_entry__

__graft_entry__.load()
from jax_sgmc_amd import scheduler as SC  # noqa: E402

TINY = np.finfo(np.float32).tiny


def _inclusion(p, k):
  """Exact probability that item i is among k = 2 successive draws without replacement."""
  assert k == 2
  p = np.asarray(p, np.float64) / np.sum(p)
  return np.array([p[i] + sum(p[j] * p[i] / (1 - p[j]) for j in range(len(p)) if j != i)
                   for i in range(len(p))])


def _frequencies(select, p, k, keys):
  counts = np.zeros(len(p))
  for key in keys:
    counts[select(jr.uniform(key, (len(p),), TINY, 1.0), p, k)] += 1
  return counts / len(keys)


def test_gumbel_topk_inclusion_probabilities():
  p = np.array([0.05, 0.1, 0.15, 0.2, 0.2, 0.3], np.float32)
  keys = jr.split(jr.PRNGKey(0), 20000)
  f = _frequencies(SC.gumbel_topk, p, 2, keys)
  P = _inclusion(p, 2)
  tol = 4.5 * np.sqrt(P * (1 - P) / len(keys))
  assert np.all(np.abs(f - P) <= tol), (f, P)
  # a planted sign mistake (+log p) selects against the weights and fails the check
  wrong = lambda u, q, k: np.argsort(-(-np.log(-np.log(u))) + np.log(q))[:k]
  fw = _frequencies(wrong, p, 2, keys[:4000])
  assert np.any(np.abs(fw - P) > 4.5 * np.sqrt(P * (1 - P) / 4000))


def test_zero_probability_iterations_are_never_selected():
  p = np.concatenate([np.zeros(30), np.linspace(1e-3, 5e-4, 70)]).astype(np.float32)
  for s in range(50):
    u = jr.uniform(jr.PRNGKey(s), (100,), TINY, 1.0)
    sel = SC.gumbel_topk(u, p, 20)
    assert len(sel) == 20 and len(set(sel.tolist())) == 20 and sel.min() >= 30
    assert np.all(np.diff(sel) > 0)
