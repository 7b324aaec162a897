"""Float32 port of the Bayesian-MLP samplers, for the timed CPU baseline (oracle only).

TEST INFRASTRUCTURE: bench.py's cpu_baseline leg times it, tests/ check it against the
float64 oracle (oracle/model.py, oracle/steps.py). The likelihood is the MLP softmax
cross-entropy of oracle/model.py (potential.py:159-214 under value_and_grad,
integrator.py:166, 792): forward and hand-differentiated backward in float32 numpy,
whose matrix products go to numpy's BLAS (sgemm on all the host's BLAS threads); the
index draw, the jax-compatible noise tree and the parameter updates run in the C port
(oracle/cport.c, OpenMP over parameters):
- pSGLD: langevin_diffusion with rms_prop (integrator.py:859-922, adaption.py:254-291);
- OBABO steps with identity mass (integrator.py:203-273) and the full-data potential
  (potential.py:219-293) summed in float64, the pieces of sggmc (solver.py:500-567).
Keys and data keys follow the reference's splits exactly, as in oracle/steps.py.
"""
import numpy as np

from oracle import cport
from oracle import jaxrandom as jr

F32 = np.float32
NCLS = 10


class MLPPort:
  """The sampler state lives in numpy float32 arrays owned by the caller."""

  def __init__(self, X, labels, dims, n, prior_scale=1.0, temperature=1.0, threads=1,
               fast=True):
    self.X = np.ascontiguousarray(X, F32)
    self.labels = np.asarray(labels).astype(np.int64)
    self.N = self.X.shape[0]
    self.D, self.H1, self.H2 = dims
    self.n = int(n)
    self.sizes = [self.D * self.H1, self.H1 * self.H2, self.H2 * NCLS, self.H1, self.H2, NCLS]
    self.P = int(sum(self.sizes))
    self.off = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
    self.s2 = F32(prior_scale) ** 2
    self.log_norm = F32(np.log(F32(2 * np.pi) * self.s2))
    self.T = F32(temperature)
    self.threads = int(threads)
    self.fast = fast
    self.g = np.empty(self.P, F32)

  def _leaf(self, v, i):
    shp = [(self.D, self.H1), (self.H1, self.H2), (self.H2, NCLS), (self.H1,), (self.H2,),
           (NCLS,)][i]
    return v[self.off[i]:self.off[i + 1]].reshape(shp)

  def forward(self, theta, Xb):
    W1, W2, W3, b1, b2, b3 = (self._leaf(theta, i) for i in range(6))
    a1 = Xb @ W1 + b1
    h1 = np.maximum(a1, F32(0))
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, F32(0))
    logits = h2 @ W3 + b3
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    se = e.sum(axis=1, keepdims=True)
    lse = m + np.log(se)
    return a1, h1, a2, h2, logits, lse, e / se

  def likelihoods(self, theta, rows):
    Xb = self.X[rows]
    *_, logits, lse, _ = self.forward(theta, Xb)
    return logits[np.arange(len(rows)), self.labels[rows]] - lse[:, 0]

  def potential_grad(self, theta, idx):
    """U (float) and grad U (float32 [P], in self.g) on the rows idx."""
    Xb = self.X[idx]
    lab = self.labels[idx]
    n = len(idx)
    a1, h1, a2, h2, logits, lse, p = self.forward(theta, Xb)
    lik = logits[np.arange(n), lab] - lse[:, 0]
    lp = np.sum((self.log_norm + theta * theta / self.s2) / F32(-2), dtype=np.float64)
    U = (-self.N * float(np.mean(lik, dtype=np.float64)) - lp) / float(self.T)
    c = F32(-self.N / n) / self.T
    dL = -p
    dL[np.arange(n), lab] += F32(1)
    dL *= c
    W2, W3 = self._leaf(theta, 1), self._leaf(theta, 2)
    g = self.g
    np.matmul(h2.T, dL, out=self._leaf(g, 2))
    self._leaf(g, 5)[:] = dL.sum(axis=0)
    dA2 = (dL @ W3.T) * (a2 > 0)
    np.matmul(h1.T, dA2, out=self._leaf(g, 1))
    self._leaf(g, 4)[:] = dA2.sum(axis=0)
    dA1 = (dA2 @ W2.T) * (a1 > 0)
    np.matmul(Xb.T, dA1, out=self._leaf(g, 0))
    self._leaf(g, 3)[:] = dA1.sum(axis=0)
    g += theta / self.s2 / self.T
    return U, g

  def psgld(self, theta, v, key, dkey, eps, alpha=0.9, lmbd=1e-5):
    """len(eps) pSGLD iterations in place (theta, v float32 [P]; key, dkey uint32[2]).
    Returns the last potential."""
    U = 0.0
    for e in np.asarray(eps, F32):
      key[:], sub = jr.split(key)
      idx = cport.draw_indices(dkey, self.n, self.N, fast=self.fast)
      xi = cport.tree_normals(sub, self.sizes, self.threads, fast=self.fast)
      U, g = self.potential_grad(theta, idx)
      ns = float(np.sqrt(F32(F32(2) * self.T) * e))
      cport.sgld_update(theta, v, g, xi, float(e), ns, alpha, lmbd, True, self.threads,
                        fast=self.fast)
    return U

  def obabo_steps(self, theta, p, key, dkey, eps, friction, steps):
    """`steps` OBABO steps in place; returns (KE_start, KE_end) accumulated from zero."""
    e = F32(eps)
    perm = np.exp(F32(-F32(friction)) * e).astype(F32)
    decay = float(np.sqrt(perm))
    noise = float(np.sqrt((F32(1) - perm) * self.T))
    ks = ke = 0.0
    for _ in range(steps):
      k, s1, s2 = jr.split(key, 3)
      key[:] = k
      xi = cport.tree_normals(s1, self.sizes, self.threads, fast=self.fast)
      idx = cport.draw_indices(dkey, self.n, self.N, fast=self.fast)
      _, g = self.potential_grad(theta, idx)
      ks += 0.5 * cport.obabo_first(theta, p, g, xi, float(e), decay, noise, self.threads,
                                    fast=self.fast)
      idx = cport.draw_indices(dkey, self.n, self.N, fast=self.fast)
      _, g = self.potential_grad(theta, idx)
      xi = cport.tree_normals(s2, self.sizes, self.threads, fast=self.fast)
      ke += 0.5 * cport.obabo_second(p, g, xi, float(e), decay, noise, self.threads,
                                     fast=self.fast)
    return ks, ke

  def full_potential(self, theta, mb):
    """(-sum_i lik_i - log p(theta)) / T over all rows, batches of mb, float64 sums."""
    tot = 0.0
    for r0 in range(0, self.N, mb):
      rows = np.arange(r0, min(self.N, r0 + mb))
      tot += float(np.sum(self.likelihoods(theta, rows), dtype=np.float64))
    t = theta.astype(np.float64)
    lp = np.sum((np.log(2 * np.pi * float(self.s2)) + t * t / float(self.s2)) / -2.0)
    return (-tot - lp) / float(self.T)
