"""Likelihood families and potentials, restated in float64 numpy (oracle only).

Stochastic potential (potential.py:159-214), no mask on the stochastic path
(potential.py:182-183):
    U(theta) = (-N * mean_i lik_i - log_prior(theta)) / T
    grad U   = (-(N / n) * sum_i grad lik_i - grad log_prior) / T
Full potential (potential.py:264-291 with data/core.py:571-572, 944-949): ordered
batches of size mb starting at ``offset`` with wrap-around; each row counted once:
    U_full = (-sum_i lik_i - log_prior) / T.

Families (the likelihood a user writes in the reference as ``likelihood(sample,
obs)``):
- linear-Gaussian, jax.scipy.stats.norm.logpdf form
  lik_i = (log(2 pi s^2) + (y_i - x_i.w)^2 / s^2) / (-2);
- logistic, lik_i = y_i z_i - log(1 + exp(z_i)), z_i = x_i.w;
- MLP softmax cross-entropy: relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3,
  lik_i = log_softmax(logits_i)[label_i]; parameters flattened in sorted-key order
  (W1, W2, W3, b1, b2, b3), as jax's tree_flatten orders a dict.
Gaussian prior: log p(theta) = sum_j (log(2 pi s^2) + theta_j^2 / s^2) / (-2).

Pinning: linear-Gaussian and the potential scaling against the reference's doctest
values (docs/usage/potential.rst:100-118, 163-164) and closed forms
(tests/test_potential.py:84); every gradient against float64 central finite
differences (tests/test_oracle_model.py).
"""
import numpy as np

LOG_2PI = np.log(2.0 * np.pi)


def gaussian_logprior(theta, scale):
  theta = np.asarray(theta, np.float64)
  s2 = float(scale) ** 2
  val = np.sum((np.log(2.0 * np.pi * s2) + theta * theta / s2) / -2.0)
  return val, -theta / s2


def linear_gaussian(w, X, y, sigma):
  """Per-row lik (n,) and per-row gradients d lik_i / d w as (n, D)."""
  w = np.asarray(w, np.float64)
  X = np.asarray(X, np.float64)
  y = np.asarray(y, np.float64)
  s2 = float(sigma) ** 2
  r = y - X @ w
  lik = (np.log(2.0 * np.pi * s2) + r * r / s2) / -2.0
  dlik = X * (r / s2)[:, None]
  return lik, dlik


def _softplus(z):
  return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def _sigmoid(z):
  return 0.5 * (1.0 + np.tanh(0.5 * z))


def logistic(w, X, y):
  w = np.asarray(w, np.float64)
  X = np.asarray(X, np.float64)
  y = np.asarray(y, np.float64)
  z = X @ w
  lik = y * z - _softplus(z)
  dlik = X * (y - _sigmoid(z))[:, None]
  return lik, dlik


def mlp_shapes(d_in, h1, h2, d_out):
  """Leaf shapes in tree_flatten (sorted-key) order: W1, W2, W3, b1, b2, b3."""
  return [(d_in, h1), (h1, h2), (h2, d_out), (h1,), (h2,), (d_out,)]


def mlp_unflatten(theta, dims):
  shapes = mlp_shapes(*dims)
  out, off = [], 0
  for s in shapes:
    size = int(np.prod(s))
    out.append(np.asarray(theta[off:off + size], np.float64).reshape(s))
    off += size
  W1, W2, W3, b1, b2, b3 = out
  return W1, W2, W3, b1, b2, b3


def mlp_softmax(theta, X, labels, dims):
  """Returns lik (n,) and the summed-with-weights gradient helper.

  Because the per-row gradient of an MLP is large, this returns (lik, grad_fn)
  where grad_fn(c) = sum_i c_i d lik_i / d theta for a weight vector c.
  """
  W1, W2, W3, b1, b2, b3 = mlp_unflatten(theta, dims)
  X = np.asarray(X, np.float64)
  labels = np.asarray(labels).astype(np.int64)
  a1 = X @ W1 + b1
  h1 = np.maximum(a1, 0.0)
  a2 = h1 @ W2 + b2
  h2 = np.maximum(a2, 0.0)
  logits = h2 @ W3 + b3
  m = logits.max(axis=1, keepdims=True)
  lse = m[:, 0] + np.log(np.sum(np.exp(logits - m), axis=1))
  lik = logits[np.arange(len(labels)), labels] - lse

  def grad_fn(c):
    c = np.asarray(c, np.float64)
    p = np.exp(logits - lse[:, None])
    dlog = -p
    dlog[np.arange(len(labels)), labels] += 1.0
    dlog *= c[:, None]
    gW3 = h2.T @ dlog
    gb3 = dlog.sum(0)
    dh2 = dlog @ W3.T
    da2 = dh2 * (a2 > 0)
    gW2 = h1.T @ da2
    gb2 = da2.sum(0)
    dh1 = da2 @ W2.T
    da1 = dh1 * (a1 > 0)
    gW1 = X.T @ da1
    gb1 = da1.sum(0)
    return np.concatenate([g.ravel() for g in (gW1, gW2, gW3, gb1, gb2, gb3)])

  return lik, grad_fn


class Family:
  """Binds a likelihood family to a dataset: evaluate(theta, idx) gives the
  per-row lik and the gradient of sum_i c_i lik_i for weights c."""

  def __init__(self, kind, X, y, **kw):
    self.kind = kind
    self.X = np.asarray(X)
    self.y = np.asarray(y)
    self.kw = kw

  @property
  def N(self):
    return self.X.shape[0]

  def evaluate(self, theta, idx):
    Xb = self.X[idx]
    yb = self.y[idx]
    if self.kind == "linear_gaussian":
      lik, dlik = linear_gaussian(theta, Xb, yb, self.kw["sigma"])
      return lik, (lambda c: c @ dlik)
    if self.kind == "logistic":
      lik, dlik = logistic(theta, Xb, yb)
      return lik, (lambda c: c @ dlik)
    if self.kind == "mlp":
      return mlp_softmax(theta, Xb, yb, self.kw["dims"])
    raise ValueError(self.kind)


def stochastic_potential(family, theta, idx, prior_scale, temperature=1.0):
  """(U, grad U, lik) of the minibatch potential at rows ``idx``."""
  N = family.N
  n = len(idx)
  lik, gfun = family.evaluate(theta, idx)
  lp, glp = gaussian_logprior(theta, prior_scale)
  U = (-N * np.mean(lik) - lp) / temperature
  g = (gfun(np.full(n, -N / n)) - glp) / temperature
  return U, g, lik


def full_potential(family, theta, prior_scale, mb, offset=0, temperature=1.0):
  """Full-data potential, summed batch by batch in the order the reference's
  ordered device batching visits the rows (wrap-around at N, masked by batch
  position), accumulated in float64."""
  N = family.N
  nb = int(np.ceil(N / mb))
  total = 0.0
  line = offset
  for it in range(nb):
    idx = np.mod(np.arange(mb) + line, N)
    line = (idx[-1] + 1) % N
    mask = (it * mb + np.arange(mb)) < N
    lik, _ = family.evaluate(theta, idx)
    pot = -N / mb * np.dot(lik, mask)
    total += pot * mb / N
  lp, _ = gaussian_logprior(theta, prior_scale)
  return (total - lp) / temperature
