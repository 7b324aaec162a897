"""sgmc_glm_potential_noalloc (no workspace argument, static device scratch) against
sgmc_glm_potential with a caller-allocated workspace (GPU): bit-identical potential,
gradient, per-row likelihoods and var(lik) on two calls in a row, the second smaller
(fewer blocks, smaller D: its arrival counter must not land on a record the first call
wrote), then a larger call again; a call too large for the scratch is refused before
any launch. The oracle checks the first call's values."""
import numpy as np
import pytest
import torch

from oracle import model as om

from jax_sgmc_amd import _lib as L

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _problem(N, D, n, family, seed):
  rng = np.random.default_rng(seed)
  X = (rng.standard_normal((N, D)) / np.sqrt(D)).astype(np.float32)
  w = rng.standard_normal(D)
  if family == L.LIK_LOGISTIC:
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ w))).astype(np.float32)
  else:
    y = (X @ w + 0.5 * rng.standard_normal(N)).astype(np.float32)
  ldx = (D + 3) // 4 * 4
  Xp = np.zeros((N, ldx), np.float32)
  Xp[:, :D] = X
  theta = (0.3 * rng.standard_normal(D)).astype(np.float32)
  idx = rng.integers(0, N, n).astype(np.int32)
  return X, y, theta, idx, dict(X=torch.as_tensor(Xp, device=DEV), ldx=ldx,
                                y=torch.as_tensor(y, device=DEV),
                                idx=torch.as_tensor(idx, device=DEV),
                                th=torch.as_tensor(theta, device=DEV))


def _outputs(n, D):
  return [torch.full((k,), float("nan"), device=DEV) for k in (1, D, n, 1)]


def _call(noalloc, family, d, N, D, n, nblocks, outs):
  U, g, lik, var = outs
  args = [family, L.ptr(d["X"]), d["ldx"], L.ptr(d["y"]), L.ptr(d["idx"]), n, N, D,
          L.ptr(d["th"]), 0.5, 1, 2.0, 1.3, L.ptr(U), L.ptr(g), L.ptr(lik), L.ptr(var)]
  if noalloc:
    L.call("sgmc_glm_potential_noalloc", *args, nblocks, L.stream())
  else:
    ws = torch.zeros(int(L.query("sgmc_glm_workspace_floats", n, D, nblocks)), device=DEV)
    L.call("sgmc_glm_potential", *args, L.ptr(ws), ws.numel(), nblocks, L.stream())
  return torch.cat(outs)


CASES = [  # (family, N, D, n, nblocks): each call smaller than the one before, then larger
    (L.LIK_LOGISTIC, 20000, 512, 4096, 64),
    (L.LIK_LINEAR_GAUSSIAN, 5000, 37, 300, 5),
    (L.LIK_LOGISTIC, 8000, 1024, 2040, 255),
]


def test_noalloc_matches_workspace_entry_point():
  print("sgmc_glm_potential_noalloc vs sgmc_glm_potential", flush=True)
  for i, (family, N, D, n, nb) in enumerate(CASES):
    print("case", i, "family", family, "N", N, "D", D, "n", n, "blocks", nb, flush=True)
    X, y, theta, idx, d = _problem(N, D, n, family, seed=i)
    ref = _call(False, family, d, N, D, n, nb, _outputs(n, D))
    got = _call(True, family, d, N, D, n, nb, _outputs(n, D))
    torch.cuda.synchronize()
    assert torch.isfinite(ref).all()
    assert torch.equal(got, ref), (i, (got - ref).abs().max().item())
    kind = "logistic" if family == L.LIK_LOGISTIC else "linear_gaussian"
    Ur, gr, _ = om.stochastic_potential(om.Family(kind, X, y, sigma=0.5), theta, idx, 2.0, 1.3)
    assert abs(got[0].item() - Ur) <= 1e-5 * abs(Ur) + 1e-4, (got[0].item(), Ur)
    np.testing.assert_allclose(got[1:1 + D].cpu().numpy(), gr, rtol=1e-4,
                               atol=1e-5 * np.abs(gr).max())


def test_noalloc_refuses_oversized_calls():
  print("sgmc_glm_potential_noalloc size limit", flush=True)
  family, N, D, n = L.LIK_LOGISTIC, 8000, 1024, 2048
  _, _, _, _, d = _problem(N, D, n, family, seed=5)
  with pytest.raises(ValueError, match="static scratch"):
    _call(True, family, d, N, D, n, 256, _outputs(n, D))
