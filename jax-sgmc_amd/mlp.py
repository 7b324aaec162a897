"""Bayesian MLP classifier on MFMA: chain engine and autograd wrapper.

Model: x -> relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3 -> 10-class softmax
cross-entropy likelihood, Gaussian prior on every parameter. Parameters are a dict
{"W1","W2","W3","b1","b2","b3"} whose tree_flatten (sorted-key) order is W1, W2, W3,
b1, b2, b3; one flat float32 vector per chain in that order.

``FusedMLP`` advances C chains per call of ``sgmc_mlp_step`` (index draw, gather, MFMA
forward and hand-written backward in bf16, e4m3 or float32, update fused into the
weight-gradient GEMMs);
``MLPPotential`` exposes the same kernels as a torch.autograd.Function (forward: U,
backward: the kernel's gradient).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from . import tree as T

NCLS = 10
# sgmc_mlp_state.precision codes
PRECISIONS = {"bf16": 0, "fp8": 1, "fp32": 2}


MlpState = _lib.MlpState


def leaf_shapes(D_in, H1, H2):
  return {"W1": (D_in, H1), "W2": (H1, H2), "W3": (H2, NCLS), "b1": (H1,), "b2": (H2,),
          "b3": (NCLS,)}


def param_count(D_in, H1, H2):
  return D_in * H1 + H1 * H2 + H2 * NCLS + H1 + H2 + NCLS


def init_params(key_seed, D_in, H1, H2, scale=None):
  """Glorot-like random init as a flat vector (numpy, float32)."""
  rng = np.random.default_rng(key_seed)
  parts = []
  for name in ("W1", "W2", "W3", "b1", "b2", "b3"):
    shp = leaf_shapes(D_in, H1, H2)[name]
    if name.startswith("W"):
      s = scale if scale is not None else math.sqrt(2.0 / shp[0])
      parts.append((s * rng.standard_normal(shp)).astype(np.float32).ravel())
    else:
      parts.append(np.zeros(shp, np.float32).ravel())
  return np.concatenate(parts)


class MLPSoftmax:
  """likelihood(sample, obs) for one observation, torch (generic path)."""

  def __init__(self, x="x", y="y"):
    self.x, self.y = x, y

  def __call__(self, sample, obs):
    h = torch.relu(obs[self.x] @ sample["W1"] + sample["b1"])
    h = torch.relu(h @ sample["W2"] + sample["b2"])
    z = h @ sample["W3"] + sample["b3"]
    lab = obs[self.y].long().reshape(1)
    return torch.gather(torch.log_softmax(z, -1), -1, lab).reshape(())


class FusedMLP:
  """C chains on the MLP likelihood: SGLD (optionally rms_prop), OBABO
  (``integrator="obabo"``: ``steps`` OBABO steps per iteration, identity mass), or
  SGGMC (``integrator="sggmc"``: OBABO proposals with the full-data FP64 MH test).

  For SGGMC the keys are the solver's (split into the integrator key and the MH key
  as solver.py:478-487 does). precision "fp32" runs the GEMMs on float32 operands
  (the fp32 matrix instruction, activations in float32, weights read from theta);
  "bf16" on bf16 operands and a bf16 shadow of the parameters; "fp8" on e4m3 operands
  with per-tensor delayed scaling. State, noise and updates are float32 in all three.

  ``tune`` (five ints, see sgmc_mlp_state.tune) overrides the k-tile depth and split-K
  factor of each GEMM; autotune() chooses them by timing the live problem."""

  def __init__(self, X, labels, dims, n, theta0, eps_table, keys=None, dkeys=None, rms=False,
               alpha=0.9, lmbd=1e-5, prior_scale=1.0, has_prior=True, temperature=1.0,
               noise_temperature=1.0, save_slots=None, device="cuda", integrator="sgld",
               friction=1.0, steps=1, precision="bf16", tune=None):
    _lib.require_gpu()
    _lib.load()
    dev = torch.device(device)
    D_in, H1, H2 = dims
    self.dims = dims
    self.n = int(n)
    if X.dtype != torch.float32 or X.dim() != 2 or X.stride(1) != 1 or X.stride(0) % 4:
      raise ValueError("X must be float32 [N, ldx] with ldx % 4 == 0")
    self.X = X
    self.labels = labels.to(torch.int32).contiguous()
    th = np.asarray(theta0, np.float32)
    if th.ndim == 1:
      th = th[None]
    self.C = th.shape[0]
    self.P = param_count(D_in, H1, H2)
    assert th.shape[1] == self.P
    self.theta = torch.as_tensor(th, device=dev).contiguous()
    self.v = torch.ones_like(self.theta) if rms else None
    if precision not in PRECISIONS:
      raise ValueError(f"precision must be one of {tuple(PRECISIONS)}")
    self.precision = precision
    self.tune = [0] * 5 if tune is None else [int(t) for t in tune]
    self.gemm_plan()  # validates tune against the shapes
    self.shadow = (None if precision == "fp32" else
                   torch.empty(self.theta.shape, dtype=torch.int16, device=dev))
    self.shadow8 = self.q = None
    if precision == "fp8":
      self.shadow8 = torch.zeros(self.theta.shape, dtype=torch.uint8, device=dev)
      self.q = torch.zeros((self.C, 16), device=dev)
    kw = np.zeros((self.C, 2), np.uint32) if keys is None else np.asarray(keys, np.uint32)
    dk = np.zeros((self.C, 2), np.uint32) if dkeys is None else np.asarray(dkeys, np.uint32)
    self.keys = torch.as_tensor(kw.view(np.int32).copy(), device=dev)
    self.dkeys = torch.as_tensor(dk.view(np.int32).copy(), device=dev)
    nb = int(_lib.query("sgmc_mlp_workspace_bytes", self.C, self.n, D_in, H1, H2))
    self.work = torch.zeros(nb, dtype=torch.uint8, device=dev)
    self.U = torch.zeros(self.C, device=dev)
    eps = np.ascontiguousarray(np.asarray(eps_table, np.float32).reshape(-1))
    self.eps_table = torch.as_tensor(eps, device=dev)
    self.table_len = eps.size
    self.iter = torch.zeros(1, dtype=torch.int32, device=dev)
    self.counters = torch.zeros(self.C, dtype=torch.int32, device=dev)
    self.done = torch.zeros(1, dtype=torch.int32, device=dev)
    self.grad = None
    self.save_slot = None
    self.samples = None
    self.samples_U = None
    self.n_slots = 0
    if save_slots is not None:
      ss = np.asarray(save_slots, np.int32).reshape(-1)
      self.save_slot = torch.as_tensor(ss, device=dev)
      self.n_slots = int(ss.max()) + 1 if ss.size and ss.max() >= 0 else 0
      if self.n_slots:
        self.samples = torch.zeros((self.C, self.n_slots, self.P), device=dev)
        self.samples_U = torch.zeros((self.C, self.n_slots), device=dev)
    self.cfg = dict(rms=1 if rms else 0, alpha=float(alpha), lmbd=float(lmbd),
                    temperature=float(temperature), noise_temperature=float(noise_temperature),
                    has_prior=1 if has_prior else 0, prior_scale=float(prior_scale))
    if integrator not in ("sgld", "obabo", "sggmc"):
      raise ValueError("integrator must be 'sgld', 'obabo' or 'sggmc'")
    self.integrator = integrator
    self.steps = int(steps) if integrator != "sgld" else 1
    self.mom = self.ke = self.U_half = self.decay = self.noise = None
    self.U_full = self.U_prop = self.theta_old = self.mom_old = self.U_old = None
    self.mh_key = self.mh_stats = self.accepted = None
    if integrator != "sgld":
      if rms:
        raise ValueError("rms_prop applies to the SGLD integrator only")
      f32 = np.float32
      perm = np.exp(f32(-f32(friction)) * eps).astype(f32)
      self.decay = torch.as_tensor(np.sqrt(perm).astype(f32), device=dev)
      self.noise = torch.as_tensor(
          np.sqrt(((f32(1) - perm) * f32(noise_temperature)).astype(f32)).astype(f32),
          device=dev)
      self.mom = torch.zeros_like(self.theta)
      self.ke = torch.zeros((self.C, 2), dtype=torch.float64, device=dev)
      self.U_half = torch.zeros(self.C, device=dev)
    self.refresh_shadow()
    if integrator == "sggmc":
      # solver.py:478-487: (integrator key, MH key) = split(key)
      from . import random as R
      ks = np.stack([R.key_words(R.split(R.key_from_words(k, device=dev)))
                     for k in kw.reshape(self.C, 2)])
      self.keys.copy_(torch.as_tensor(ks[:, 0].copy().view(np.int32), device=dev))
      self.mh_key = torch.as_tensor(ks[:, 1].copy().view(np.int32), device=dev)
      self.U_full = torch.zeros(self.C, dtype=torch.float64, device=dev)
      self.U_prop = torch.zeros(self.C, dtype=torch.float64, device=dev)
      self.theta_old = self.theta.clone()
      self.mom_old = torch.zeros_like(self.theta)
      self.U_old = torch.zeros(self.C, device=dev)
      self.accepted = torch.zeros(self.C, dtype=torch.int32, device=dev)
      if self.n_slots:
        self.mh_stats = torch.zeros((self.C, self.n_slots, 3), device=dev)
      self.full_potential(out=self.U_full)

  def refresh_shadow(self):
    """Rewrites the GEMM copy of the parameters after theta changed outside the
    engine (bf16 shadow, or the e4m3 shadows and their scales; float32 GEMMs read
    theta itself)."""
    if self.precision == "fp32":
      return
    _lib.call("sgmc_mlp_shadow", _lib.ptr(self.theta), _lib.ptr(self.shadow),
              self.theta.numel(), _lib.stream())
    if self.precision == "fp8":
      st = self._state()
      _lib.call("sgmc_mlp_init_fp8", ctypes.byref(st), _lib.stream())
      # two gradient-mode evaluations calibrate the activation scales
      grad = self.grad
      if self.grad is None:
        self.grad = torch.empty_like(self.theta)
      U = self.U.clone()
      for _ in range(2):
        st = self._state()
        _lib.call("sgmc_mlp_step", ctypes.byref(st), 0, _lib.stream())
      self.U.copy_(U)
      self.grad = grad

  def _state(self):
    D_in, H1, H2 = self.dims
    st = MlpState()
    st.X, st.labels = self.X.data_ptr(), self.labels.data_ptr()
    st.N, st.ldx = self.X.shape[0], self.X.stride(0)
    st.D_in, st.H1, st.H2, st.n, st.C = D_in, H1, H2, self.n, self.C
    st.theta = self.theta.data_ptr()
    st.v = self.v.data_ptr() if self.v is not None else None
    st.shadow = self.shadow.data_ptr() if self.shadow is not None else None
    st.P = self.P
    st.key, st.dkey = self.keys.data_ptr(), self.dkeys.data_ptr()
    st.work, st.work_bytes = self.work.data_ptr(), self.work.numel()
    st.U = self.U.data_ptr()
    st.grad = self.grad.data_ptr() if self.grad is not None else None
    st.samples = self.samples.data_ptr() if self.samples is not None else None
    st.samples_U = self.samples_U.data_ptr() if self.samples_U is not None else None
    st.n_slots = self.n_slots
    st.eps_table = self.eps_table.data_ptr()
    st.save_slot = self.save_slot.data_ptr() if self.save_slot is not None else None
    st.iter, st.counters, st.done = (self.iter.data_ptr(), self.counters.data_ptr(),
                                     self.done.data_ptr())
    st.table_len = self.table_len
    for k, v in self.cfg.items():
      setattr(st, k, v)
    opt = lambda t: t.data_ptr() if t is not None else None
    st.mom, st.ke, st.U_half = opt(self.mom), opt(self.ke), opt(self.U_half)
    st.decay_table, st.noise_table = opt(self.decay), opt(self.noise)
    st.U_full, st.U_prop = opt(self.U_full), opt(self.U_prop)
    st.theta_old, st.mom_old, st.U_old = opt(self.theta_old), opt(self.mom_old), opt(self.U_old)
    st.mh_key, st.mh_stats, st.accepted = opt(self.mh_key), opt(self.mh_stats), opt(self.accepted)
    st.mode = 0
    st.precision = PRECISIONS[self.precision]
    st.shadow8, st.q = opt(self.shadow8), opt(self.q)
    st.tune[:] = self.tune
    return st

  def gemm_plan(self, update=True):
    """(k-tile depth, split-K factor) of the GEMMs G1 = Xb W1, G2 = h1 W2, G3 = dA2 W2^T,
    G4 = h1^T dA2, G5 = Xb^T dA1 as an iteration (update) or a gradient evaluation
    launches them."""
    D_in, H1, H2 = self.dims
    out = (ctypes.c_int32 * 5)()
    _lib.call("sgmc_mlp_gemm_plan", D_in, H1, H2, self.n, self.C, PRECISIONS[self.precision],
              1 if update else 0, (ctypes.c_int32 * 5)(*self.tune), out)
    return [(v >> 4, v & 15) for v in out]

  def tile_candidates(self, slot):
    """Codes 16 k + s that sgmc_mlp_step accepts for GEMM `slot` at these shapes."""
    D_in, H1, H2 = self.dims
    K = (D_in, H1, H2, self.n, self.n)[slot]
    depths = {"fp32": (32,), "fp8": (64, 128), "bf16": (32, 128)}[self.precision]
    splits = (1, 2, 3, 4) if slot < 3 else (1,)
    return [k * 16 + s for k in depths for s in splits if K % (k * s) == 0]

  def _mutable(self):
    names = ("theta", "v", "shadow", "shadow8", "q", "keys", "dkeys", "U", "iter", "counters",
             "done", "grad", "samples", "samples_U", "mom", "ke", "U_half", "U_full", "U_prop",
             "theta_old", "mom_old", "U_old", "mh_key", "mh_stats", "accepted", "work")
    return [t for t in (getattr(self, n) for n in names) if t is not None]

  def autotune(self, reps=20, min_gain=0.01):
    """Chooses the k-tile depth and split-K factor of each GEMM by timing the live
    problem (HIP events on the launch stream around `reps` back-to-back units after one
    warm-up unit): whole iterations for SGLD, one OBABO step (two evaluations, the unit
    OBABO and SGGMC iterations repeat) otherwise. The forward GEMMs one at a time, then the
    two parameter GEMMs jointly (they share a launch at equal depths), starting from the
    built-in choice; a candidate replaces the current one only if it is faster by more
    than `min_gain`. Each timing starts from a snapshot of the engine's state (the
    iteration counter runs at most reps + 1 past it; the schedule tables clamp it),
    which is restored at the end, so the chains continue where they were. Sets
    self.tune (and self.tune_ms: the time per unit of every plan tried, keyed (slot, code)
    and (3, gW2 code, gW1 code), self.tune_gain:
    (built-in ms, tuned ms)) and returns gemm_plan()."""
    unit = self.iteration if self.integrator == "sgld" else self.obabo_step
    tensors = self._mutable()
    snap = [t.clone() for t in tensors]

    def restore():
      for t, s0 in zip(tensors, snap):
        t.copy_(s0)

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed():
      restore()
      unit()  # warm-up
      e0.record()
      for _ in range(reps):
        unit()
      e1.record()
      e1.synchronize()
      return e0.elapsed_time(e1) / reps

    self.tune = [k * 16 + s for k, s in self.gemm_plan()]
    self.tune_ms = {}
    timed()  # the first timing after other work runs slow by a few percent
    best = builtin = timed()
    for slot in range(3):
      best_code = self.tune[slot]
      self.tune_ms[(slot, best_code)] = best
      for code in self.tile_candidates(slot):
        if code == best_code:
          continue
        self.tune[slot] = code
        t = timed()
        self.tune_ms[(slot, code)] = t
        if t < best * (1.0 - min_gain):
          best, best_code = t, code
      self.tune[slot] = best_code
    # the parameter GEMMs jointly: gW2 runs in gW1's launch when both take the same
    # k-tile, so their depths are not independent (tune_ms[(3, c4, c5)])
    best_pair = (self.tune[3], self.tune[4])
    self.tune_ms[(3,) + best_pair] = best
    for c4 in self.tile_candidates(3):
      for c5 in self.tile_candidates(4):
        if (c4, c5) == best_pair:
          continue
        self.tune[3], self.tune[4] = c4, c5
        t = timed()
        self.tune_ms[(3, c4, c5)] = t
        if t < best * (1.0 - min_gain):
          best, best_pair = t, (c4, c5)
    self.tune[3], self.tune[4] = best_pair
    self.tune_gain = (builtin, best)
    restore()
    torch.cuda.synchronize()
    return self.gemm_plan()

  def iteration(self):
    if self.integrator == "sgld":
      st = self._state()
      _lib.call("sgmc_mlp_step", ctypes.byref(st), 1, _lib.stream())
      return
    for _ in range(self.steps):
      self.obabo_step()
    if self.integrator == "sggmc":
      self.full_potential(out=self.U_prop)
      st = self._state()
      _lib.call("sgmc_mlp_mh", ctypes.byref(st), _lib.stream())
    else:
      self._end_iteration()

  def obabo_step(self):
    """One OBABO step (integrator.py:203-273) for every chain: two launches of the
    kernel sequence, each on a fresh minibatch."""
    st = self._state()
    st.mode = 1
    _lib.call("sgmc_mlp_step", ctypes.byref(st), 1, _lib.stream())
    st.mode = 2
    _lib.call("sgmc_mlp_step", ctypes.byref(st), 1, _lib.stream())

  def _end_iteration(self):
    # plain OBABO through solver.sgmc: save the kept sample, advance the counter
    if self.samples is not None:
      it = torch.clamp(self.iter, 0, self.table_len - 1).long()
      slot = self.save_slot[it]
      ok = (slot >= 0)
      sl = torch.clamp(slot, 0, self.n_slots - 1).long()
      cur = self.samples[:, sl[0]]
      self.samples[:, sl[0]] = torch.where(ok, self.theta, cur)
      self.samples_U[:, sl[0]] = torch.where(ok, self.U, self.samples_U[:, sl[0]])
    self.iter += 1

  def full_potential(self, out=None):
    """float64 full-data potential of every chain's current parameters."""
    if out is None:
      out = torch.empty(self.C, dtype=torch.float64, device=self.theta.device)
    st = self._state()
    _lib.call("sgmc_mlp_full_potential", ctypes.byref(st), _lib.ptr(out), _lib.stream())
    return out

  def potential_and_grad(self):
    """U [C] and grad U [C, P] at the current parameters for the minibatch the data
    keys select (the keys do not advance)."""
    if self.grad is None:
      self.grad = torch.empty_like(self.theta)
    st = self._state()
    _lib.call("sgmc_mlp_step", ctypes.byref(st), 0, _lib.stream())
    return self.U.clone(), self.grad.clone()

  def capture(self, iterations_per_graph=1):
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with _lib.capture_guard(), torch.cuda.stream(s):
      with torch.cuda.graph(g, stream=s):
        for _ in range(iterations_per_graph):
          self.iteration()
    torch.cuda.current_stream().wait_stream(s)
    self._graph, self._graph_iters = g, iterations_per_graph
    return g

  def run(self, iterations, graph_iterations=0):
    if graph_iterations and iterations >= graph_iterations:
      g = getattr(self, "_graph", None)
      if g is None or self._graph_iters != graph_iterations:
        g = self.capture(graph_iterations)
      reps, rest = divmod(iterations, graph_iterations)
      for _ in range(reps):
        g.replay()
      for _ in range(rest):
        self.iteration()
    else:
      for _ in range(iterations):
        self.iteration()

  def params(self, c):
    D_in, H1, H2 = self.dims
    flat = self.theta[c]
    out, off = {}, 0
    for name in ("W1", "W2", "W3", "b1", "b2", "b3"):
      shp = leaf_shapes(D_in, H1, H2)[name]
      size = int(np.prod(shp))
      out[name] = flat[off:off + size].reshape(shp)
      off += size
    return out


class MLPPotential(torch.autograd.Function):
  """U(theta) for every chain of a FusedMLP, differentiable: backward returns the
  hand-written kernel's gradient (no activations kept; the kernels recompute)."""

  @staticmethod
  def forward(ctx, theta, engine):
    with torch.no_grad():
      engine.theta.copy_(theta)
      engine.refresh_shadow()
      U, g = engine.potential_and_grad()
    ctx.save_for_backward(g)
    return U

  @staticmethod
  def backward(ctx, gU):
    (g,) = ctx.saved_tensors
    return gU[:, None] * g, None
