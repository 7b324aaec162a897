"""Closed-form targets for Bayesian linear regression (oracle only).

Model: y = X w + N(0, sigma^2), prior w ~ N(0, tau2 I).
Posterior: Lambda = X^T X / sigma^2 + I / tau2, mu = Lambda^{-1} X^T y / sigma^2.

SGLD at a fixed step size eps and temperature T with the minibatch gradient
g(w) = -(N/n) sum_{i in B} grad lik_i(w) + w / tau2 (n rows drawn with replacement)
is the linear recursion w' = w - eps g(w) + sqrt(2 T eps) xi. Its stationary mean is
exactly mu (g is unbiased) and its stationary covariance S solves
    S = (I - eps H) S (I - eps H)^T + 2 T eps I + eps^2 E_w[Cov(g | w)],
    E_w[Cov(g | w)] = (N^2 / n) (Cov_i[r_i(mu)] + E_i[A_i S A_i] - A S A),
with A_i = x_i x_i^T / sigma^2, A = E_i A_i, r_i(w) = x_i (y_i - x_i.w) / sigma^2 and
H = Lambda. This is the "matched posterior error" target at finite eps; S tends to
T Lambda^{-1} as eps -> 0.
"""
import numpy as np
from scipy import linalg


def linreg_posterior(X, y, sigma, tau2):
  X = np.asarray(X, np.float64)
  y = np.asarray(y, np.float64)
  D = X.shape[1]
  Lam = X.T @ X / sigma ** 2 + np.eye(D) / tau2
  mu = np.linalg.solve(Lam, X.T @ y / sigma ** 2)
  return mu, Lam


def sgld_stationary(X, y, sigma, tau2, eps, n, temperature=1.0, iters=200):
  X = np.asarray(X, np.float64)
  y = np.asarray(y, np.float64)
  N, D = X.shape
  mu, H = linreg_posterior(X, y, sigma, tau2)
  r = X * ((y - X @ mu) / sigma ** 2)[:, None]
  cov_r = np.cov(r.T, bias=True)
  A = X.T @ X / (N * sigma ** 2)
  M = np.eye(D) - eps * H
  S = temperature * np.linalg.inv(H)
  for _ in range(iters):
    XS = X @ S
    quad = np.einsum("ij,ij->i", XS, X)          # x_i^T S x_i
    EASA = (X * quad[:, None]).T @ X / (N * sigma ** 4)
    noise = (N * N / n) * (cov_r + EASA - A @ S @ A)
    Q = 2.0 * temperature * eps * np.eye(D) + eps * eps * noise
    S_new = linalg.solve_discrete_lyapunov(M, Q)
    if np.max(np.abs(S_new - S)) <= 1e-14 * np.max(np.abs(S_new)):
      S = S_new
      break
    S = S_new
  return mu, S
