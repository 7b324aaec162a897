"""Likelihood and prior families the potential recognizes and runs on HIP kernels.

Each family is also an ordinary callable with the reference's signatures
(``likelihood(sample, observation)`` for one observation, ``prior(sample)``;
potential.py:37-40, 128-137), written in torch, so the same object can be evaluated
by the generic path. ``minibatch_potential``/``full_potential`` dispatch a family to
its kernels:
- LinearGaussian: jax.scipy.stats.norm.logpdf(y - x.w, scale=sigma)
  = (log(2 pi sigma^2) + (y - x.w)^2 / sigma^2) / (-2);
- Logistic: y z - log(1 + e^z), z = x.w (Bernoulli log-likelihood with logits z);
- MLPSoftmax: log-softmax cross-entropy of a ReLU MLP (see mlp.py);
- GaussianPrior: sum over all parameters of norm.logpdf(theta, 0, scale).
"""
import math

import torch

from . import _lib
from . import tree as T


def _single_leaf(sample):
  leaves, _ = T.flatten(sample)
  if len(leaves) != 1:
    raise ValueError("this likelihood family expects a sample with exactly one leaf "
                     "(the weight vector)")
  return leaves[0]


class GLMLikelihood:
  family = None

  def __init__(self, x="x", y="y"):
    self.x, self.y = x, y

  def z(self, sample, obs):
    w = _single_leaf(sample)
    return obs[self.x] @ w


class LinearGaussian(GLMLikelihood):
  family = _lib.LIK_LINEAR_GAUSSIAN

  def __init__(self, sigma, x="x", y="y"):
    super().__init__(x, y)
    self.sigma = float(sigma)

  def __call__(self, sample, obs):
    r = obs[self.y] - self.z(sample, obs)
    s2 = self.sigma ** 2
    return (math.log(2 * math.pi * s2) + r * r / s2) / -2.0


class Logistic(GLMLikelihood):
  family = _lib.LIK_LOGISTIC
  sigma = 1.0

  def __call__(self, sample, obs):
    z = self.z(sample, obs)
    return obs[self.y] * z - torch.nn.functional.softplus(z)


class GaussianPrior:
  """log p(theta) = sum_j norm.logpdf(theta_j, 0, scale) over every leaf."""

  def __init__(self, scale=1.0):
    self.scale = float(scale)

  def __call__(self, sample):
    leaves, _ = T.flatten(sample)
    s2 = self.scale ** 2
    tot = 0.0
    for l in leaves:
      l = torch.as_tensor(l)
      tot = tot + torch.sum((math.log(2 * math.pi * s2) + l * l / s2) / -2.0)
    return tot


class ZeroPrior:
  """The constant prior ``lambda _: 0.0``."""

  scale = 1.0

  def __call__(self, sample):
    return 0.0
