"""Fused multi-chain SG-MCMC for generalized-linear likelihoods.

Every chain of a chain table advances by one gradient evaluation per kernel launch
(``sgmc_glm_fused_step``): index draw, row gather, GLM forward/backward, reduction
and the integrator update with in-register noise, for all chains at once. This
replaces solver.mcmc's sequential per-chain loop (solver.py:162-171, strategy
'map') and its jit(lax.scan) over iterations (solver.py:152-160): the step size of
iteration t is read on the device from a schedule table, the iteration counter lives
on the device, so a captured HIP graph of one or more iterations replays without
host input.

Chains may differ in dataset, N, D (<= 1024), minibatch size and temperature
(tempering ladders); they share the family, the prior and the integrator. Per-chain
results equal the same chain run alone (chains are independent; every chain owns
its keys and its reduction order is fixed by its own block count). TableGroup runs
several tables side by side on their own streams; ReplicaExchange, AmagoldMH and
SggmcMH attach a swap or Metropolis-Hastings step to a table.
"""
import ctypes
import math
import time
import weakref

import numpy as np
import torch

from . import _lib

MODES = {"sgld": _lib.MODE_SGLD, "sgld_rms": _lib.MODE_SGLD_RMS, "sghmc": _lib.MODE_SGHMC,
         "amagold": _lib.MODE_AMAGOLD, "obabo": _lib.MODE_OBABO, "sggmc": _lib.MODE_SGGMC}
_STEPPED = (_lib.MODE_SGHMC, _lib.MODE_AMAGOLD, _lib.MODE_OBABO, _lib.MODE_SGGMC)
_OBABO = (_lib.MODE_OBABO, _lib.MODE_SGGMC)
# floats of per-chain energy bookkeeping: AMAGOLD's energy; OBABO's kinetic energies at
# the start and the end; SGGMC's as OBABO plus the minibatch potential of the start
_ENERGY = {_lib.MODE_AMAGOLD: 1, _lib.MODE_OBABO: 2, _lib.MODE_SGGMC: 3}
_KEEPS_START = (_lib.MODE_AMAGOLD, _lib.MODE_SGGMC)
FAMILIES = {"linear_gaussian": _lib.LIK_LINEAR_GAUSSIAN, "logistic": _lib.LIK_LOGISTIC}
NT = 256


def _align4(x):
  return (x + 3) // 4 * 4


# Rows per block whose likelihoods a fused block stages in LDS (kLdsRows in glm.hip).
LDS_ROWS = 4096


def blocks_for(n, D, target_bytes):
  """Grid blocks for one chain: about target_bytes of rows per block, at least 32
  rows per block, at most LDS_ROWS rows per block."""
  rows_bytes = n * 4 * _align4(D)
  nb = int(round(rows_bytes / max(1, target_bytes)))
  return max(1, -(-n // LDS_ROWS), min(nb, (n + 31) // 32))


class FusedGLM:
  """Chain table for ``sgmc_glm_fused_step``.

  chains: list of dicts with keys X (device float32 [N, ldx], ldx % 4 == 0, zero
  padding beyond D), y (device float32 [N]), D, n, theta0 (array-like [D]), key and
  dkey (two uint32 words each), temperature (SGLD noise temperature, default 1).
  Optional per-chain keys: N (observations, default X.shape[0]) and save (default
  True; False leaves the chain out of the sample buffer).
  keep_lik: also keep the per-row likelihoods of the last evaluation in ``lik``.
  mode: "sgld", "sgld_rms" (one launch per iteration), "sghmc", "amagold" (``steps``
  launches), "obabo", "sggmc" (two launches per leapfrog step); "amagold" and "sggmc"
  expect their Metropolis-Hastings test (AmagoldMH, SggmcMH) to be attached.
  inv_mass / sqrt_mass: a diagonal mass shared by all chains (device float32 [D]).
  """

  def __init__(self, chains, family, mode, eps_table, steps=1, friction=1.0, alpha=0.9,
               lmbd=1e-5, sigma=1.0, prior_scale=1.0, has_prior=True, pot_temperature=1.0,
               inv_mass=None, sqrt_mass=None, save_slots=None, n_slots=0, target_bytes=None,
               keep_lik=False, device="cuda"):
    _lib.require_gpu()
    self.device = torch.device(device)
    self.family = FAMILIES[family] if isinstance(family, str) else int(family)
    self.mode = MODES[mode] if isinstance(mode, str) else int(mode)
    self.steps = int(steps) if self.mode in _STEPPED else 1
    # OBABO modes take two launches per leapfrog step (O B A, then B O)
    self.launches = 2 * self.steps if self.mode in _OBABO else self.steps
    self.friction = float(friction)
    self.alpha, self.lmbd = float(alpha), float(lmbd)
    self.sigma, self.prior_scale = float(sigma), float(prior_scale)
    self.has_prior = 1 if has_prior else 0
    self.pot_temperature = float(pot_temperature)
    self.spec = chains
    self.C = len(chains)
    eps = np.ascontiguousarray(np.asarray(eps_table, np.float32).reshape(-1))
    self.table_len = eps.size
    self.eps_table = torch.as_tensor(eps, device=self.device)
    self.save_slot = None
    self.n_slots = int(n_slots)
    if save_slots is not None:
      ss = np.ascontiguousarray(np.asarray(save_slots, np.int32).reshape(-1))
      assert ss.size == self.table_len
      self.save_slot = torch.as_tensor(ss, device=self.device)
      self.n_slots = max(self.n_slots, int(ss.max()) + 1 if ss.size else 0)
    self.inv_mass = inv_mass
    self.sqrt_mass = sqrt_mass
    Ds = [int(c["D"]) for c in chains]
    for c in chains:
      X, y = c["X"], c["y"]
      D, n = int(c["D"]), int(c["n"])
      N = int(c.get("N", X.shape[0]))
      if not 1 <= D <= 1024:
        raise ValueError("fused GLM path supports 1 <= D <= 1024")
      if not (X.is_cuda and y.is_cuda and X.dtype == torch.float32 and
              y.dtype == torch.float32 and X.dim() == 2 and X.stride(1) == 1):
        raise ValueError("X must be a float32 [N, ldx] device tensor, y float32 [N]")
      if X.stride(0) % 4 or X.stride(0) < _align4(D) or X.data_ptr() % 16:
        raise ValueError("X rows must be 16-byte aligned and padded to a multiple of 4 floats")
      if not 1 <= N <= min(X.shape[0], y.numel(), 2 ** 31 - 1):
        raise ValueError("observation count outside the dataset")
      if not y.is_contiguous():
        raise ValueError("y must be contiguous")
      if n < 1:
        raise ValueError("minibatch size must be positive")
      if np.asarray(c["theta0"]).size != D:
        raise ValueError("theta0 must have D entries")
    dcl = {_lib.query("sgmc_glm_dclass", d) for d in Ds}
    self.uniform_dclass = dcl.pop() if len(dcl) == 1 else -1
    total_bytes = sum(int(c["n"]) * 4 * _align4(int(c["D"])) for c in chains)
    self.target_bytes = target_bytes or max(64 * 1024, total_bytes // 2048)
    self.keep_lik = bool(keep_lik)
    self.exchange = None
    self.mh = None
    # iterations launched through this object: the device counter's value as the
    # launches leave it, known on the host without a synchronization (a sparse
    # exchange and the loop launches are planned from it)
    self.host_iter = 0
    self._graph = None
    self._capture_stream = None
    # run / capture / replay default for loop launches: None = whenever the table
    # allows them
    self.prefer_loop = None
    self._allocate()
    self._build_table()

  # -------------------------------------------------------------- storage ----
  def _allocate(self):
    dev = self.device
    ch = self.spec
    self.D = np.array([int(c["D"]) for c in ch])
    self.n = np.array([int(c["n"]) for c in ch])
    Da = np.array([_align4(d) for d in self.D])
    self.poff = np.concatenate([[0], np.cumsum(Da)]).astype(np.int64)
    P = int(self.poff[-1])
    self.theta_flat = torch.zeros(P, device=dev)
    self.mom_flat = torch.zeros(P, device=dev) if self.mode in _STEPPED else None
    self.v_flat = torch.ones(P, device=dev) if self.mode == _lib.MODE_SGLD_RMS else None
    self.energy_width = _ENERGY.get(self.mode, 0)
    keeps = self.mode in _KEEPS_START
    self.energy = (torch.zeros(self.C * self.energy_width, device=dev)
                   if self.energy_width else None)
    self.theta_old_flat = torch.zeros(P, device=dev) if keeps else None
    self.mom_old_flat = torch.zeros(P, device=dev) if keeps else None
    for i, c in enumerate(ch):
      self.theta_flat[self.poff[i]:self.poff[i] + self.D[i]] = torch.as_tensor(
          np.asarray(c["theta0"], np.float32).reshape(-1), device=dev)
    keys = np.array([np.asarray(c.get("key", (0, 0)), np.uint32) for c in ch])
    dkeys = np.array([np.asarray(c.get("dkey", (0, 0)), np.uint32) for c in ch])
    self.keys = torch.as_tensor(keys.view(np.int32).copy(), device=dev)
    self.dkeys = torch.as_tensor(dkeys.view(np.int32).copy(), device=dev)
    self.U = torch.zeros(self.C, device=dev)
    self.var = torch.zeros(self.C, device=dev)
    self.loff = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
    self.lik = torch.zeros(int(self.loff[-1]), device=dev) if self.keep_lik else None
    self.counters = torch.zeros(self.C, dtype=torch.int32, device=dev)
    self.iter = torch.zeros(1, dtype=torch.int32, device=dev)
    self.done = torch.zeros(1, dtype=torch.int32, device=dev)
    self.samples = None
    self.samples_U = None
    if self.n_slots > 0:
      self.samples = torch.zeros(self.n_slots * P, device=dev)
      self.samples_U = torch.zeros(self.C * self.n_slots, device=dev)

  def _build_table(self):
    nbs = [blocks_for(int(n), int(d), self.target_bytes) for n, d in zip(self.n, self.D)]
    self.nblocks_chain = np.array(nbs)
    # Grid order: chains with the most bytes per block first, so that the longest
    # blocks start in the first wave of the dispatch (longest-processing-time first).
    per_block = self.n * 4 * ((self.D + 3) // 4 * 4) / np.maximum(1, self.nblocks_chain)
    order = np.argsort(-per_block, kind="stable")
    self.block_begin = np.zeros(self.C, np.int64)
    self.block_begin[order] = np.concatenate([[0], np.cumsum(self.nblocks_chain[order])[:-1]])
    self.nblocks = int(self.nblocks_chain.sum())
    D4 = (self.D + 3) // 4
    stride = 4 * D4 + 2
    pof = np.concatenate([[0], np.cumsum(self.nblocks_chain * stride)]).astype(np.int64)
    self.partial = torch.zeros(int(pof[-1]), device=self.device)
    arr = (_lib.GlmChain * self.C)()
    fs = 4
    P = int(self.poff[-1])
    for i, c in enumerate(self.spec):
      e = arr[i]
      e.X = c["X"].data_ptr()
      e.y = c["y"].data_ptr()
      e.theta = self.theta_flat.data_ptr() + fs * int(self.poff[i])
      e.mom = self.mom_flat.data_ptr() + fs * int(self.poff[i]) if self.mom_flat is not None else None
      e.v = self.v_flat.data_ptr() + fs * int(self.poff[i]) if self.v_flat is not None else None
      e.key = self.keys.data_ptr() + 8 * i
      e.dkey = self.dkeys.data_ptr() + 8 * i
      e.U = self.U.data_ptr() + fs * i
      e.var = self.var.data_ptr() + fs * i
      e.lik = (self.lik.data_ptr() + fs * int(self.loff[i])) if self.lik is not None else None
      e.partial = self.partial.data_ptr() + fs * int(pof[i])
      e.counter = self.counters.data_ptr() + 4 * i
      saves = self.samples is not None and c.get("save", True)
      e.samples = (self.samples.data_ptr() + fs * int(self.poff[i]) * self.n_slots
                   if saves else None)
      e.samples_U = (self.samples_U.data_ptr() + fs * i * self.n_slots if saves else None)
      e.N = int(c["X"].shape[0]) if "N" not in c else int(c["N"])
      e.ldx = int(c["X"].stride(0))
      e.D = int(self.D[i])
      e.n = int(self.n[i])
      e.block_begin = int(self.block_begin[i])
      e.nblocks = int(self.nblocks_chain[i])
      e.dclass = _lib.query("sgmc_glm_dclass", int(self.D[i]))
      e.temperature = float(c.get("temperature", 1.0))
      if self.energy is not None:
        e.energy = self.energy.data_ptr() + fs * self.energy_width * i
      if self.theta_old_flat is not None:
        e.theta_old = self.theta_old_flat.data_ptr() + fs * int(self.poff[i])
        e.mom_old = self.mom_old_flat.data_ptr() + fs * int(self.poff[i])
    bc = np.repeat(order.astype(np.int32), self.nblocks_chain[order])
    # a uniform table in chain order needs no block map (the kernel divides)
    self.uniform_blocks = bool(np.all(self.nblocks_chain == self.nblocks_chain[0]) and
                               np.array_equal(order, np.arange(self.C)))
    # The kernels trust the table: check it on the host before it goes to the device.
    bc_host = None if self.uniform_blocks else (ctypes.c_int32 * self.nblocks)(*bc.tolist())
    _lib.call("sgmc_glm_check_table", arr, self.C, bc_host, self.nblocks, 0)
    if self.loop_capable():
      _lib.call("sgmc_glm_check_table", arr, self.C, None, self.nblocks, 1)
    raw = bytes(arr)
    self.table = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(self.device)
    self.block_chain = torch.as_tensor(bc, device=self.device)
    self._arr = arr  # keep the host copy alive for inspection
  # ------------------------------------------------------------ launching ----
  def _launch(self, step):
    # A mixed-D table runs as one launch of the kernel with a block-uniform D-class
    # switch: per-class launches leave the device underfilled on small problems.
    bmap = None if self.uniform_blocks else _lib.ptr(self.block_chain)
    _lib.call("sgmc_glm_fused_step", _lib.ptr(self.table), self.C, bmap,
              self.nblocks, self.uniform_dclass, self.family, self.sigma, self.has_prior,
              self.prior_scale, self.pot_temperature, self.mode, step, self.steps,
              self.friction, self.alpha, self.lmbd, _lib.ptr(self.inv_mass),
              _lib.ptr(self.sqrt_mass), _lib.ptr(self.eps_table), self.table_len,
              _lib.ptr(self.save_slot), _lib.ptr(self.iter), _lib.ptr(self.done), 1,
              _lib.stream())

  def _launch_steps(self):
    for s in range(self.launches):
      self._launch(s)

  def loop_capable(self):
    """Chains of one block each, no Metropolis-Hastings test, and no exchange or one
    that runs only at its test iterations: many iterations can run in one launch
    (``sgmc_glm_fused_loop``), cut at the exchange's test iterations."""
    return (bool(np.all(self.nblocks_chain == 1)) and self.mh is None and
            (self.exchange is None or self.exchange.sparse) and
            self.mode in (_lib.MODE_SGLD, _lib.MODE_SGLD_RMS, _lib.MODE_SGHMC,
                          _lib.MODE_OBABO))

  def _launch_loop(self, iterations):
    _lib.call("sgmc_glm_fused_loop", _lib.ptr(self.table), self.C, self.uniform_dclass,
              self.family, self.sigma, self.has_prior, self.prior_scale, self.pot_temperature,
              self.mode, self.steps, self.friction, self.alpha, self.lmbd,
              _lib.ptr(self.inv_mass), _lib.ptr(self.sqrt_mass), _lib.ptr(self.eps_table),
              self.table_len, _lib.ptr(self.save_slot), _lib.ptr(self.iter),
              _lib.ptr(self.done), int(iterations), _lib.stream())

  def _plan(self, iterations, loop):
    """Launch sequence of ``iterations`` iterations starting after host_iter: ("it",)
    one iteration's evaluations, ("loop", k) k iterations in one loop launch, ("swap",)
    the exchange, ("mh",) the Metropolis-Hastings test."""
    ops, done, left = [], self.host_iter, int(iterations)
    ex = self.exchange
    while left > 0:
      if loop:
        k = min(1000, left)
        t = ex.iterations_to_next_launch(done) if ex is not None else None
        if t is not None:
          k = min(k, t)
        ops.append(("loop", k))
      else:
        k = 1
        ops.append(("it",))
      done += k
      left -= k
      if ex is not None and ex.needs_launch(done):
        ops.append(("swap",))
      if self.mh is not None:
        ops.append(("mh",))
    return ops

  def _execute(self, ops):
    for op in ops:
      if op[0] == "it":
        self._launch_steps()
        self.host_iter += 1
      elif op[0] == "loop":
        self._launch_loop(op[1])
        self.host_iter += op[1]
      elif op[0] == "swap":
        self.exchange.launch()
      else:
        self.mh.launch()

  def _loop_arg(self, loop):
    if loop is None:
      loop = self.prefer_loop
    if loop is None:
      return self.loop_capable()
    if loop and not self.loop_capable():
      raise ValueError("loop launches need single-block chains, no MH test, and no "
                       "exchange or one without saving pairs")
    return bool(loop)

  def iteration(self):
    """Advances every chain by one reference iteration (steps launches, two per step in
    the OBABO modes), followed by the replica-exchange step or the Metropolis-Hastings
    test when one is attached."""
    self._execute(self._plan(1, False))

  def capture(self, iterations_per_graph=1, loop=None):
    """Captures ``iterations_per_graph`` iterations into a HIP graph (loop launches
    when the table allows them, unless ``loop=False``). The launch sequence depends on
    where the iterations start relative to a sparse exchange's test iterations: replay
    the returned graph (or run with graph_iterations), which checks that it still
    fits, and captures again when it does not."""
    loop = self._loop_arg(loop)
    plan = self._plan(iterations_per_graph, loop)
    g = torch.cuda.CUDAGraph()
    # one capture stream per table, kept: streams created on every capture would move
    # the hardware queues that later streams get
    if self._capture_stream is None:
      self._capture_stream = torch.cuda.Stream(device=self.device)
    s = self._capture_stream
    s.wait_stream(torch.cuda.current_stream())
    start = self.host_iter
    with _lib.capture_guard(), torch.cuda.stream(s):
      with torch.cuda.graph(g, stream=s):
        self._execute(plan)
    torch.cuda.current_stream().wait_stream(s)
    self.host_iter = start  # capturing records the launches without running them
    self._graph = EngineGraph(self, g, int(iterations_per_graph), loop, plan)
    return self._graph

  def replay(self, iterations_per_graph, loop=None):
    """Replays the captured graph of ``iterations_per_graph`` iterations, capturing it
    first when there is none or when its launch sequence does not fit this start."""
    loop = self._loop_arg(loop)
    g = self._graph
    if (g is None or g.iterations != iterations_per_graph or g.loop != loop or
        g.plan != self._plan(iterations_per_graph, loop)):
      g = self.capture(iterations_per_graph, loop)
    g.replay()

  def run(self, iterations, graph_iterations=0, loop=None):
    """Runs ``iterations`` iterations (eagerly, or by replaying a captured graph of
    ``graph_iterations`` iterations; the remainder runs eagerly). A table of
    single-block chains with no test between iterations (or only a sparse exchange)
    runs them in loop launches of up to 1000 iterations (``loop=False`` disables
    that)."""
    loop = self._loop_arg(loop)
    rest = int(iterations)
    if graph_iterations and iterations >= graph_iterations:
      reps, rest = divmod(int(iterations), int(graph_iterations))
      for _ in range(reps):
        self.replay(graph_iterations, loop)
    self._execute(self._plan(rest, loop))

  # ------------------------------------------------------------- results -----
  def theta(self, i=None):
    if i is None:
      return [self.theta(j) for j in range(self.C)]
    return self.theta_flat[self.poff[i]:self.poff[i] + self.D[i]]

  def momentum(self, i):
    return self.mom_flat[self.poff[i]:self.poff[i] + self.D[i]]

  def kinetic_energies(self, i):
    """OBABO modes: chain i's (kinetic energy start, kinetic energy end) accumulators."""
    w = self.energy_width
    return self.energy[i * w:i * w + 2]

  def sample_trace(self, i):
    """[n_slots, D] samples saved for chain i."""
    P = int(self.poff[-1])
    base = int(self.poff[i]) * self.n_slots
    return self.samples[base:base + self.n_slots * int(self.D[i])].reshape(self.n_slots,
                                                                          int(self.D[i]))

  def sample_potentials(self, i):
    """[n_slots] potential stored with each saved sample of chain i."""
    return self.samples_U[i * self.n_slots:(i + 1) * self.n_slots]

  def iteration_count(self):
    return int(self.iter.item())

  # ------------------------------------------------------------ autotuner ----
  def autotune(self, candidates=None, reps=3):
    """Picks the bytes-per-block target with the shortest launch time; restores the
    chain state afterwards."""
    if candidates is None:
      base = self.target_bytes
      candidates = sorted({max(16 * 1024, int(base * f)) for f in (0.25, 0.5, 1, 2, 4, 8, 16)})
    snap = self._snapshot()
    best = None
    for tb in candidates:
      self.target_bytes = tb
      self._build_table()
      self._launch_steps()
      torch.cuda.synchronize()
      t0 = time.perf_counter()
      for _ in range(reps):
        self._launch_steps()
      torch.cuda.synchronize()
      dt = (time.perf_counter() - t0) / reps
      if best is None or dt < best[0]:
        best = (dt, tb)
    self.target_bytes = best[1]
    self._build_table()
    self._restore(snap)
    self._graph = None
    self.tuned = {"target_bytes": best[1], "seconds_per_iteration": best[0],
                  "candidates": candidates}
    return self.tuned

  def _snapshot(self):
    t = {k: getattr(self, k).clone() for k in ("theta_flat", "keys", "dkeys", "iter", "U", "var")}
    for k in ("mom_flat", "v_flat", "samples", "samples_U", "energy", "theta_old_flat",
              "mom_old_flat"):
      v = getattr(self, k)
      t[k] = v.clone() if v is not None else None
    return t

  def _restore(self, snap):
    for k, v in snap.items():
      if v is not None:
        getattr(self, k).copy_(v)
    torch.cuda.synchronize()


class EngineGraph:
  """A HIP graph of a FusedGLM's iterations with the launch sequence it recorded.
  replay() runs it if that sequence still fits the engine's position (a sparse
  exchange tests at fixed iteration counts) and raises otherwise."""

  def __init__(self, engine, graph, iterations, loop, plan):
    # a weak reference: the engine holds its graph, and a cycle would leave the graph
    # to the cyclic collector
    self._engine = weakref.ref(engine)
    self.graph = graph
    self.iterations, self.loop, self.plan = iterations, loop, plan

  def replay(self):
    e = self._engine()
    if e is None:
      raise RuntimeError("the engine of this graph no longer exists")
    if e._plan(self.iterations, self.loop) != self.plan:
      raise RuntimeError("this graph was recorded at another position relative to the "
                         "exchange's test iterations: replay through FusedGLM.replay / run, "
                         "which capture again when needed")
    self.graph.replay()
    e.host_iter += self.iterations


class TableGroup:
  """Several chain tables advanced side by side, each on its own HIP stream.

  A fused launch begins with a prologue (index draw, evaluation point) before its row
  gather saturates HBM and ends with a tail in which only each chain's last-arriving
  block works (the reduction records and the update with its noise). Tables on
  separate streams fill one table's prologue and tail with another table's row gather,
  so the device stays on the gather for nearly the whole step. Chains stay
  independent: a chain gives bit for bit what it gives in a single table with the same
  blocks per chain.

  engines: FusedGLM tables. A replica exchange or MH test attached to a table acts on
  that table's chains only and runs on its stream after each of its iterations. Results
  are indexed over the group's chains in table order.
  stream_of: the stream of each table (indices from 0; tables on one stream run one
  after the other). Default: one stream per table. The group's streams come from a
  per-device pool of streams with hardware queues of their own
  (``_lib.dedicated_stream``): HIP gives a process few hardware queues
  (GPU_MAX_HW_QUEUES, 4 by default) and shares them among ordinary streams, and the
  kernels of two streams on one queue run one after the other. Two groups used at the
  same time share the pool's streams.
  """

  def __init__(self, engines, stream_of=None):
    engines = list(engines)
    if not engines:
      raise ValueError("TableGroup needs at least one table")
    if len({str(e.device) for e in engines}) != 1:
      raise ValueError("the tables of a TableGroup share one device")
    self.engines = engines
    if stream_of is None:
      stream_of = list(range(len(engines)))
    stream_of = [int(i) for i in stream_of]
    if len(stream_of) != len(engines) or min(stream_of) < 0:
      raise ValueError("stream_of needs one stream index >= 0 per table")
    self.stream_of = stream_of
    self.streams = [_lib.dedicated_stream(engines[0].device, i)
                    for i in range(max(stream_of) + 1)]
    self.offsets = np.concatenate([[0], np.cumsum([e.C for e in engines])]).astype(np.int64)
    self.C = int(self.offsets[-1])
    self._graphs = None
    self._graph_iters = 0

  @classmethod
  def split(cls, chains, tables, *args, **kwargs):
    """``tables`` FusedGLM tables over consecutive, near-equal slices of ``chains``
    (the other arguments as for FusedGLM)."""
    k = max(1, min(int(tables), len(chains)))
    cut = np.linspace(0, len(chains), k + 1).round().astype(int)
    return cls([FusedGLM(chains[a:b], *args, **kwargs) for a, b in zip(cut[:-1], cut[1:])])

  @property
  def nblocks(self):
    return [e.nblocks for e in self.engines]

  def _locate(self, i):
    t = int(np.searchsorted(self.offsets, i, side="right") - 1)
    return self.engines[t], i - int(self.offsets[t])

  def _fork(self):
    cur = torch.cuda.current_stream()
    for s in self.streams:
      s.wait_stream(cur)

  def _join(self):
    cur = torch.cuda.current_stream()
    for s in self.streams:
      cur.wait_stream(s)

  def _side_stream(self):
    if getattr(self, "_side", None) is None:
      self._side = torch.cuda.Stream(device=self.engines[0].device)
    return self._side

  def _table_streams(self):
    return [self.streams[i] for i in self.stream_of]

  def _each(self, fn):
    self._fork()
    for e, s in zip(self.engines, self._table_streams()):
      with torch.cuda.stream(s):
        fn(e)
    self._join()

  def iteration(self):
    """One iteration of every table, the tables concurrently."""
    self._each(lambda e: e.iteration())

  def capture(self, iterations_per_graph=1):
    """One HIP graph of ``iterations_per_graph`` iterations per table, replayed on
    the table's stream."""
    self._graphs = [e.capture(iterations_per_graph) for e in self.engines]
    self._graph_iters = iterations_per_graph
    return self._graphs

  def run(self, iterations, graph_iterations=0):
    """Runs ``iterations`` iterations of every table (by replaying graphs of
    ``graph_iterations`` iterations when given, captured again where a table's launch
    sequence no longer fits; the remainder eagerly). The current stream waits for all
    tables."""
    rest = iterations
    if graph_iterations and iterations >= graph_iterations:
      reps, rest = divmod(iterations, graph_iterations)
      self._fork()
      for _ in range(reps):
        for e, s in zip(self.engines, self._table_streams()):
          with torch.cuda.stream(s):
            e.replay(graph_iterations)
      self._join()
    if rest:
      self._each(lambda e: e.run(rest))

  def autotune(self, candidates=None, reps=50, tables=None):
    """Chooses the bytes-per-block target of the tables listed in ``tables`` (indices;
    default all): first one target common to them, then each table's own, one table at
    a time. Each candidate is timed as a HIP graph in which every table of the group
    runs ``reps`` iterations' evaluations back to back on its stream (loop-capable
    tables not being tuned in one loop launch, as they run), with HIP events around a
    replay after a warm-up replay. Exchanges and tests are not launched; the chain
    states are restored afterwards."""
    tables = list(range(len(self.engines))) if tables is None else list(tables)
    if not tables:
      return None
    eng = [self.engines[i] for i in tables]
    if candidates is None:
      base = max(e.target_bytes for e in eng)
      candidates = sorted({max(16 * 1024, int(base * f)) for f in (0.25, 0.5, 1, 2, 4, 8, 16)})
    snaps = [e._snapshot() for e in self.engines]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    built = {}

    def block():
      cur = torch.cuda.current_stream()
      for s in self.streams:
        s.wait_stream(cur)
      for i, (e, s) in enumerate(zip(self.engines, self._table_streams())):
        with torch.cuda.stream(s):
          if i not in tables and e.loop_capable():
            e._launch_loop(reps)
          else:
            for _ in range(reps):
              e._launch_steps()
      for s in self.streams:
        cur.wait_stream(s)

    def timed(targets):
      for e, tb in zip(eng, targets):
        if built.get(id(e)) != tb:
          e.target_bytes = tb
          e._build_table()
          built[id(e)] = tb
      g = torch.cuda.CUDAGraph()
      side = self._side_stream()
      side.wait_stream(torch.cuda.current_stream())
      with _lib.capture_guard(), torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
          block()
      torch.cuda.current_stream().wait_stream(side)
      g.replay()
      e0.record()
      g.replay()
      e1.record()
      e1.synchronize()
      del g
      return e0.elapsed_time(e1) / 1e3 / reps

    best = None
    for tb in candidates:
      dt = timed([tb] * len(eng))
      if best is None or dt < best[0]:
        best = (dt, [tb] * len(eng))
    for j in range(len(eng)) if len(eng) > 1 else []:
      for tb in candidates:
        if tb == best[1][j]:
          continue
        trial = list(best[1])
        trial[j] = tb
        dt = timed(trial)
        if dt < best[0]:
          best = (dt, trial)
    for e, tb in zip(eng, best[1]):
      e.target_bytes = tb
      e._build_table()
      e._graph = None
    for e, snap in zip(self.engines, snaps):
      e._restore(snap)
    self._graphs = None
    self.tuned = {"target_bytes": best[1], "seconds_per_iteration": best[0],
                  "candidates": list(candidates), "tables": tables}
    return self.tuned

  def theta(self, i=None):
    if i is None:
      return [self.theta(j) for j in range(self.C)]
    e, j = self._locate(i)
    return e.theta(j)

  def momentum(self, i):
    e, j = self._locate(i)
    return e.momentum(j)

  def sample_trace(self, i):
    e, j = self._locate(i)
    return e.sample_trace(j)

  def sample_potentials(self, i):
    e, j = self._locate(i)
    return e.sample_potentials(j)

  def iteration_count(self):
    return min(e.iteration_count() for e in self.engines)


class ReplicaExchange:
  """parallel_tempering's swap step (solver.py:264-293) between pairs of chains of a
  FusedGLM, launched after every iteration (``sgmc_resgld_swap``).

  pairs: [(lo, hi)] chain indices, lo on the colder schedule; temperatures: [(t_lo,
  t_hi)] the slot temperatures; keys: one swap key (two uint32 words) per pair;
  swap_every: iterations between tests; parity: per pair -1 (tested every round) or
  0 / 1 (even / odd rounds of a ladder); save: per pair, keep the lo chain's samples at
  the engine's save slots (taken after the exchange, as the reference saves the normal
  chain). Pairs tested in one round must not share a chain.

  The swap kernel tests when the engine's iteration count (capped at the schedule's
  length, as the step size index is) is a multiple of swap_every. Without saving
  pairs the other launches do nothing, so the engine launches the kernel only at those
  iterations (``sparse``); saving pairs copy the lo chain after every iteration, which
  needs every launch.
  """

  def __init__(self, engine, pairs, temperatures, keys=None, swap_every=1, parity=None,
               save=None, F=1.0, ssq_init=0.0):
    dev = engine.device
    P = len(pairs)
    parity = [-1] * P if parity is None else [int(q) for q in parity]
    save = [False] * P if save is None else [bool(v) for v in save]
    if swap_every < 1:
      raise ValueError("swap_every must be >= 1")
    for (lo, hi), q in zip(pairs, parity):
      if not (0 <= lo < engine.C and 0 <= hi < engine.C and lo != hi):
        raise ValueError("pair chains out of range")
      if engine.D[lo] != engine.D[hi]:
        raise ValueError("paired chains must have the same D")
      if q not in (-1, 0, 1):
        raise ValueError("parity must be -1, 0 or 1")
    for r in (0, 1):  # chains tested together in an even / odd round are distinct
      used = [c for pr, q in zip(pairs, parity) if q in (-1, r) for c in pr]
      if len(used) != len(set(used)):
        raise ValueError("pairs tested in the same round share a chain")
    for p, ((lo, _), sv) in enumerate(zip(pairs, save)):
      if sv and any(lo in pr for j, pr in enumerate(pairs) if j != p):
        raise ValueError("a saving pair's lo chain must not belong to another pair")
    self.engine = engine
    self.P = P
    self.F = float(F)
    self.swap_every = int(swap_every)
    self.ssq = torch.full((P,), float(ssq_init), device=dev)
    self.step = torch.zeros(P, dtype=torch.int32, device=dev)
    kw = np.zeros((P, 2), np.uint32) if keys is None else np.asarray(keys, np.uint32)
    self.keys = torch.as_tensor(kw.reshape(P, 2).view(np.int32).copy(), device=dev)
    self.swaps = torch.zeros(P, dtype=torch.int32, device=dev)
    self.save = save
    self.n_slots = engine.n_slots
    nsave = sum(save)
    self.samples = None
    self.samples_U = None
    self.save_index = {}
    if nsave and self.n_slots:
      D = [int(engine.D[lo]) for (lo, _), sv in zip(pairs, save) if sv]
      self.soff = np.concatenate([[0], np.cumsum(D)]).astype(np.int64) * self.n_slots
      self.samples = torch.zeros(int(self.soff[-1]), device=dev)
      self.samples_U = torch.zeros(nsave * self.n_slots, device=dev)
    arr = (_lib.SwapPair * P)()
    k = 0
    for p, ((lo, hi), (tl, th), q, sv) in enumerate(zip(pairs, temperatures, parity, save)):
      e = arr[p]
      e.lo, e.hi, e.t_lo, e.t_hi, e.parity = int(lo), int(hi), float(tl), float(th), q
      e.ssq = self.ssq.data_ptr() + 4 * p
      e.step = self.step.data_ptr() + 4 * p
      e.key = self.keys.data_ptr() + 8 * p
      e.swaps = self.swaps.data_ptr() + 4 * p
      if sv and self.samples is not None:
        e.samples = self.samples.data_ptr() + 4 * int(self.soff[k])
        e.samples_U = self.samples_U.data_ptr() + 4 * k * self.n_slots
        self.save_index[p] = k
        k += 1
    self.table = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
    self.pairs = list(pairs)
    self.sparse = not any(save)
    engine.exchange = self
    engine._graph = None

  def needs_launch(self, done):
    """Whether the swap kernel has work after the engine's iteration ``done`` (count
    of completed iterations)."""
    if not self.sparse:
      return True
    d = min(int(done), self.engine.table_len)
    return d >= 1 and d % self.swap_every == 0

  def iterations_to_next_launch(self, done):
    """Iterations after ``done`` up to and including the next one that needs the swap
    kernel, or None when none does."""
    if not self.sparse:
      return 1
    S, L = self.swap_every, self.engine.table_len
    if done >= L:
      return 1 if L % S == 0 else None
    nxt = (int(done) // S + 1) * S
    return nxt - int(done) if nxt <= L else None

  def launch(self):
    eng = self.engine
    _lib.call("sgmc_resgld_swap", _lib.ptr(eng.table), eng.C, _lib.ptr(self.table), self.P,
              self.F, self.swap_every, _lib.ptr(eng.save_slot), eng.table_len,
              _lib.ptr(eng.iter), _lib.stream())

  def sample_trace(self, p):
    """[n_slots, D] samples of the lo chain of pair p (saved after the exchange)."""
    k = self.save_index[p]
    D = int(self.engine.D[self.pairs[p][0]])
    return self.samples[int(self.soff[k]):int(self.soff[k + 1])].reshape(self.n_slots, D)

  def sample_potentials(self, p):
    k = self.save_index[p]
    return self.samples_U[k * self.n_slots:(k + 1) * self.n_slots]


class _SharedDataTest:
  """Full-data Metropolis-Hastings test after every iteration of a FusedGLM whose
  chains share one dataset: the full potential of all positions runs as one pass over
  X (``sgmc_glm_full_potential_chains``, FP64 sums).

  keys: one key (two uint32 words) per chain, split as the reference's MH solvers do
  at init (solver.py:340-352, 489-490): split(key)[0] starts the integrator,
  split(key)[1] the test. full_potential: "split" (default) takes the row dot products
  on the fp16 matrix cores as hi + lo halves (about 1e-6 relative per dot product,
  ``sgmc_glm_full_potential_chains_split``), "fp32" on the fp32 matrix cores (6e-8);
  both sum in FP64."""

  mode = None

  def __init__(self, engine, keys=None, nblocks=None, full_potential="split"):
    if engine.mode != self.mode:
      raise ValueError(f"{type(self).__name__} needs a FusedGLM in mode "
                       f"{[k for k, v in MODES.items() if v == self.mode][0]!r}")
    spec = engine.spec
    c0 = spec[0]
    for c in spec:
      if (c["X"].data_ptr() != c0["X"].data_ptr() or c["y"].data_ptr() != c0["y"].data_ptr()
          or int(c["D"]) != int(c0["D"]) or int(c.get("N", c["X"].shape[0])) !=
          int(c0.get("N", c0["X"].shape[0]))):
        raise ValueError(f"{type(self).__name__} needs chains that share one dataset")
    from . import random as R
    dev = engine.device
    self.engine = engine
    self.X, self.y = c0["X"], c0["y"]
    self.N = int(c0.get("N", c0["X"].shape[0]))
    self.D = int(c0["D"])
    self.ld_theta = int(engine.poff[1] - engine.poff[0])
    if full_potential not in ("split", "fp32"):
      raise ValueError("full_potential must be 'split' or 'fp32'")
    self.split = full_potential == "split"
    # one scale for the dataset: max |X| 2^x_exp < 2^14
    amax = float(self.X[:self.N, :self.D].abs().max().item()) if self.split else 0.0
    self.x_exp = 14 - math.frexp(amax)[1] if math.isfinite(amax) and amax > 0 else 0
    self.x_exp = max(-64, min(64, self.x_exp))
    C = engine.C
    self.nblocks = int(nblocks or torch.cuda.get_device_properties(dev).multi_processor_count)
    self.ws = torch.zeros(int(_lib.query("sgmc_glm_full_chains_workspace_doubles",
                                         self.nblocks)), dtype=torch.float64, device=dev)
    self.U_full = torch.zeros(C, dtype=torch.float64, device=dev)
    self.U_prop = torch.zeros(C, dtype=torch.float64, device=dev)
    self.n_slots = engine.n_slots
    self.ratio_samples = torch.zeros(max(1, C * self.n_slots), device=dev)
    self.int_keys, mh_words = [], np.zeros((C, 2), np.uint32)
    for c in range(C):
      kw = (0, 0) if keys is None else tuple(int(v) for v in keys[c])
      k = R.split(R.key_from_words(kw, device=dev))
      mh_words[c] = R.key_words(k[1])
      self.int_keys.append(k[0].clone())
    self.keys = torch.as_tensor(mh_words.view(np.int32).copy(), device=dev)

  def _set_integrator_keys(self, words):
    self.engine.keys.copy_(torch.as_tensor(np.asarray(words, np.uint32).view(np.int32).copy(),
                                           device=self.engine.device))

  def _full(self, out):
    eng = self.engine
    head = (eng.family, _lib.ptr(self.X), int(self.X.stride(0)), _lib.ptr(self.y), self.N,
            self.D, _lib.ptr(eng.theta_flat), self.ld_theta, eng.C, eng.sigma, eng.has_prior,
            eng.prior_scale, eng.pot_temperature)
    tail = (_lib.ptr(out), _lib.ptr(self.ws), self.ws.numel(), self.nblocks, _lib.stream())
    if self.split:
      _lib.call("sgmc_glm_full_potential_chains_split", *head, self.x_exp, *tail)
    else:
      _lib.call("sgmc_glm_full_potential_chains", *head, *tail)

  def acceptance_ratios(self, i):
    """[n_slots] exp(log a) of the test at each saved iteration of chain i."""
    return self.ratio_samples[i * self.n_slots:(i + 1) * self.n_slots]


class AmagoldMH(_SharedDataTest):
  """AMAGOLD's amortized Metropolis-Hastings test (solver.py:301-432) for a FusedGLM in
  mode "amagold" whose chains share one dataset, launched after every iteration
  (``sgmc_glm_amagold_mh``). The integrator key draws the initial momentum
  sqrt_mass * xi (integrator.py:488-492)."""

  mode = _lib.MODE_AMAGOLD

  def __init__(self, engine, keys=None, nblocks=None, full_potential="split"):
    super().__init__(engine, keys, nblocks, full_potential)
    from . import random as R
    from . import tree as T
    dev = engine.device
    self.energy_kept = torch.zeros(engine.C, device=dev)
    # (k_int', sub) = split(k_int); p = sqrt_mass * random_tree(sub, theta0)
    table, nb = T.Layout([torch.zeros(self.D)]).leaf_table(dev, _lib.PAIRS_PER_BLOCK)
    int_words = np.zeros((engine.C, 2), np.uint32)
    for c, k in enumerate(self.int_keys):
      k_out = torch.empty(2, dtype=torch.int32, device=dev)
      _lib.call("sgmc_sghmc_resample", _lib.ptr(engine.momentum(c)),
                _lib.ptr(engine.sqrt_mass), _lib.ptr(table), nb, _lib.ptr(k), _lib.ptr(k_out),
                _lib.stream())
      int_words[c] = R.key_words(k_out)
    self._set_integrator_keys(int_words)
    self._full(self.U_full)
    engine.mh = self
    engine._graph = None

  def launch(self):
    eng = self.engine
    self._full(self.U_prop)
    _lib.call("sgmc_glm_amagold_mh", _lib.ptr(eng.table), eng.C, _lib.ptr(self.U_prop),
              _lib.ptr(self.U_full), _lib.ptr(self.energy_kept), _lib.ptr(self.keys),
              _lib.ptr(self.ratio_samples), self.n_slots, _lib.ptr(eng.save_slot),
              eng.table_len, _lib.ptr(eng.iter), _lib.stream())


class SggmcMH(_SharedDataTest):
  """SGGMC's Metropolis-Hastings test (solver.py:434-567) for a FusedGLM in mode
  "sggmc" whose chains share one dataset: OBABO proposals from a zero initial momentum
  (integrator.py:303), the full potential of every proposal, then
  ``sgmc_glm_sggmc_mh``. Besides the acceptance ratios it keeps, per saved iteration,
  the proposal's kinetic-energy difference and the kept full potential."""

  mode = _lib.MODE_SGGMC

  def __init__(self, engine, keys=None, nblocks=None, full_potential="split"):
    super().__init__(engine, keys, nblocks, full_potential)
    from . import random as R
    dev = engine.device
    self._set_integrator_keys([R.key_words(k) for k in self.int_keys])
    C = engine.C
    self.ke_samples = torch.zeros(max(1, C * self.n_slots), device=dev)
    self.U_samples = torch.zeros(max(1, C * self.n_slots), dtype=torch.float64, device=dev)
    self._full(self.U_full)
    engine.mh = self
    engine._graph = None

  def launch(self):
    eng = self.engine
    self._full(self.U_prop)
    _lib.call("sgmc_glm_sggmc_mh", _lib.ptr(eng.table), eng.C, _lib.ptr(self.U_prop),
              _lib.ptr(self.U_full), _lib.ptr(self.keys), _lib.ptr(self.ratio_samples),
              _lib.ptr(self.ke_samples), _lib.ptr(self.U_samples), self.n_slots,
              _lib.ptr(eng.save_slot), eng.table_len, _lib.ptr(eng.iter), _lib.stream())

  def kinetic_energies(self, i):
    """[n_slots] KE_end - KE_start of the proposal tested at each saved iteration."""
    return self.ke_samples[i * self.n_slots:(i + 1) * self.n_slots]

  def full_potentials(self, i):
    """[n_slots] full potential kept after the test at each saved iteration (FP64)."""
    return self.U_samples[i * self.n_slots:(i + 1) * self.n_slots]
