"""One small run of the hot path on cuda:0, checked against the oracle (used by
__graft_entry__.smoke())."""
import numpy as np


def run():
  import torch
  from oracle import jaxrandom as jr
  from oracle import model as om
  from oracle import steps as ost
  from jax_sgmc_amd import _lib, fused, random as R, scheduler as SC
  _lib.require_gpu()
  torch.cuda.set_device(0)

  # RNG: bit-exact against the oracle
  key = R.key_from_words((0, 42))
  z = R.normal(key, (1001,)).cpu().numpy()
  ref = jr.normal(np.array([0, 42], np.uint32), (1001,))
  assert np.array_equal(z.view(np.uint32), ref.view(np.uint32)), "normals differ from oracle"

  # Fused SGHMC on logistic regression, 2 chains x 5 iterations x 3 leapfrog steps
  rng = np.random.default_rng(0)
  N, D, n = 3000, 64, 256
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  Xd = torch.as_tensor(X, device="cuda")
  yd = torch.as_tensor(y, device="cuda")
  its = 5
  eps = SC.polynomial_table(5e-3, 1e-3, its)
  chains = [dict(X=Xd, y=yd, D=D, n=n, theta0=np.zeros(D, np.float32), key=(0, c),
                 dkey=(1, c)) for c in range(2)]
  eng = fused.FusedGLM(chains, "logistic", "sghmc", eps, steps=3, friction=0.5)
  eng.run(its)
  torch.cuda.synchronize()
  fam = om.Family("logistic", X, y)
  for c in range(2):
    r = ost.sghmc(lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0), np.zeros(D), its,
                  eps, 3, 0.5, N, n, [D], key=np.array([0, c], np.uint32),
                  dkey=np.array([1, c], np.uint32))
    got = eng.theta(c).cpu().numpy().astype(np.float64)
    ref = r["theta"][-1]
    tol = 1e-5 * np.abs(ref) + 1e-5 * np.max(np.abs(ref))
    assert np.all(np.abs(got - ref) <= tol), f"chain {c}: fused SGHMC differs from the oracle"
  print("smoke ok: RNG bit-exact, fused SGHMC within 1e-5 of the oracle", flush=True)
