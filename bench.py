"""Headline benchmark: chain-steps/s of the fused multi-chain SGHMC step on one MI355X.

Workload (the multi-chain configuration that fits one GPU): SGHMC (friction_leapfrog,
5 leapfrog steps, friction 0.5) on Bayesian logistic regression, X ~ N(0,1)/sqrt(D)
float32 [N=1,000,000, D=512], y ~ Bernoulli(sigmoid(X w*)), w* ~ N(0, I), prior
N(0, I), minibatch 4096, 64 independent chains advanced together as two chain tables
on two HIP streams. A step is one reference iteration for every chain (5 gradient
evaluations each).
Data is synthetic, drawn on the GPU with the library's jax-compatible Threefry
normal/uniform from PRNGKey(1).

Also reported: the plain single-chain path (separate index-draw, gather, potential
and update kernels per gradient evaluation), the fused/plain ratio with a check that
both compute the same chain, the roofline of the dominant kernel, and the oracle's C
port timed on the host cores as the CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain]
       python bench.py --copy-calibration   (rocprofv3 counter scaling run)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chain-steps/sec on 1 MI355X, fused vs plain HIP, at matched posterior error"
HBM_PEAK_GBS = 8000.0
N_OBS, D, BATCH, CHAINS, LEAPFROG, FRICTION = 1_000_000, 512, 4096, 64, 5, 0.5
EPS_FIRST, EPS_LAST = 1e-2, 1e-3
# Algorithmic HBM bytes of one gradient evaluation of one chain: n rows of 4D bytes,
# n labels and n indices, theta and momentum read and written.
BYTES_PER_CHAIN_GRAD = BATCH * (4 * D + 8) + 16 * D
TABLES = 2  # chain tables of the headline run, one stream each


def log(msg):
  print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_data(torch, R):
  """X ~ N(0,1)/sqrt(D), w* ~ N(0,1), y = 1[u < sigmoid(X w*)] from PRNGKey(1)."""
  kx, kw, ku = R.split(R.PRNGKey(1), 3)
  X = R.normal(kx, (N_OBS * D,)).reshape(N_OBS, D)
  X.mul_(1.0 / np.sqrt(D))
  w = R.normal(kw, (D,))
  u = R.uniform(ku, (N_OBS,))
  y = (u < torch.sigmoid(X @ w)).to(torch.float32)
  return X, y, w


def step_sizes(iters):
  sys.path.insert(0, ROOT)
  from jax_sgmc_amd import scheduler
  return scheduler.polynomial_table(EPS_FIRST, EPS_LAST, iters)


def chain_keys(R, C):
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(2), C))
  return keys, dkeys


def run_fused(torch, fused, R, X, y, K, W, graph_iters=10, target_bytes=0):
  """The 64 chains as TABLES chain tables on separate streams (fused.TableGroup): each
  table's launch prologue and finalizer tail overlap the other table's row gather.
  Every chain gives what it gives in one table with the same blocks per chain."""
  C = CHAINS
  keys, dkeys = chain_keys(R, C)
  eps = step_sizes(K + W + graph_iters)
  chains = [dict(X=X, y=y, D=D, n=BATCH, theta0=np.zeros(D, np.float32), key=tuple(keys[c]),
                 dkey=tuple(dkeys[c])) for c in range(C)]
  grp = fused.TableGroup.split(chains, TABLES, "logistic", "sghmc", eps, steps=LEAPFROG,
                               friction=FRICTION, prior_scale=1.0)
  if target_bytes:
    # a fixed grid (counter collection serializes launches, which misleads the tuner)
    tuned = grp.autotune(candidates=[target_bytes], reps=1)
  else:
    tuned = grp.autotune(reps=2)
  log(f"autotune: {tuned}")
  grp.capture(graph_iters)
  # warmup (graph replays) then the timed region, bracketed by synchronizations; the
  # group forks from and joins into the current stream, whose events bracket it
  grp.run(max(1, W // graph_iters) * graph_iters, graph_iters)
  torch.cuda.synchronize()
  reps = max(1, K // graph_iters)
  stream = torch.cuda.current_stream()
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  t0 = time.perf_counter()
  e0.record(stream)
  grp.run(reps * graph_iters, graph_iters)
  e1.record(stream)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t0
  ms = e0.elapsed_time(e1)
  steps = reps * graph_iters
  ok = all(bool(torch.isfinite(e.theta_flat).all().item()) for e in grp.engines)
  return grp, dict(ms=ms, wall=wall, steps=steps, launches=steps * LEAPFROG, finite=ok,
                   nblocks=grp.nblocks, tuned=tuned)


def run_plain(torch, L, R, T, X, y, iters, key, dkey):
  """One chain through the modular kernels: per gradient evaluation an index draw,
  a row gather, the potential/gradient kernel and the elementwise update kernels."""
  dev = X.device
  lay = T.Layout(torch.zeros(D))
  table, nb = lay.leaf_table(dev, L.PAIRS_PER_BLOCK)
  theta = torch.zeros(D, device=dev)
  p = torch.zeros(D, device=dev)
  g = torch.empty(D, device=dev)
  U = torch.empty(1, device=dev)
  k = [R.key_from_words(key), torch.empty(2, dtype=torch.int32, device=dev)]
  dk = [R.key_from_words(dkey), torch.empty(2, dtype=torch.int32, device=dev)]
  idx = torch.empty(BATCH, dtype=torch.int32, device=dev)
  Xb = torch.empty(BATCH, D, device=dev)
  yb = torch.empty(BATCH, device=dev)
  blocks = L.query("sgmc_glm_blocks", BATCH, D)
  ws = torch.zeros(int(L.query("sgmc_glm_workspace_floats", BATCH, D, blocks)), device=dev)
  eps = step_sizes(iters)
  s = L.stream()

  def one(t):
    e = float(eps[t])
    L.call("sgmc_sghmc_resample", L.ptr(p), None, L.ptr(table), nb, L.ptr(k[0]), L.ptr(k[1]), s)
    k.reverse()
    for _ in range(LEAPFROG):
      L.call("sgmc_sghmc_position", L.ptr(theta), L.ptr(p), None, None, FRICTION, D, e, s)
      L.call("sgmc_draw_batch", L.ptr(dk[0]), L.ptr(dk[1]), L.ptr(idx), BATCH, N_OBS, s)
      dk.reverse()
      L.call("sgmc_gather_rows", L.ptr(X), N_OBS, L.ptr(idx), L.ptr(Xb), BATCH, 4 * D, 4 * D, s)
      L.call("sgmc_gather_rows", L.ptr(y), N_OBS, L.ptr(idx), L.ptr(yb), BATCH, 4, 4, s)
      L.call("sgmc_glm_potential", L.LIK_LOGISTIC, L.ptr(Xb), D, L.ptr(yb), None, BATCH, N_OBS,
             D, L.ptr(theta), 1.0, 1, 1.0, 1.0, L.ptr(U), L.ptr(g), None, None, L.ptr(ws),
             ws.numel(), blocks, s)
      L.call("sgmc_sghmc_momentum", L.ptr(p), L.ptr(g), None, FRICTION, L.ptr(table), nb,
             L.ptr(k[0]), L.ptr(k[1]), e, s)
      k.reverse()

  return one, theta


def cpu_baseline(X_host, y_host, keys, dkeys, target_s=12.0):
  from oracle import cport
  threads = max(1, min(16, os.cpu_count() or 1))
  C = min(CHAINS, threads)
  eps = step_sizes(1000)

  def go(iters):
    th = np.zeros((C, D), np.float32)
    kk = np.array(keys[:C], np.uint32).copy()
    dd = np.array(dkeys[:C], np.uint32).copy()
    t0 = time.perf_counter()
    cport.sghmc(1, X_host, y_host, BATCH, th, kk, dd, eps[:iters], LEAPFROG, FRICTION,
                threads=threads)
    return time.perf_counter() - t0

  t1 = go(1)
  iters = int(max(1, min(5000, target_s / max(t1, 1e-3))))
  dt = go(iters)
  return {"value": C * iters / dt, "unit": "chain-steps/s", "cores": threads, "kind": "port",
          "sample": f"{C} chains x {iters} iterations of the same workload "
                    f"(oracle/cport.c SGHMC logistic, OpenMP over chains, {dt:.1f} s)"}


# ---------------------------------------------------------------------------------
# Secondary workloads (python bench.py --workload ...): not the headline line.

LADDER_PROBLEMS, LADDER_RUNGS, LADDER_SWAP_EVERY, LADDER_BATCH = 32, 8, 50, 512
LADDER_DIMS = (16, 64, 256, 1024)
MLP_DIMS, MLP_N, MLP_BATCH, MLP_CHAINS = (3072, 512, 512), 50_000, 256, 16


def ladder_problems(torch, R):
  """32 logistic problems: N_k = round(exp(U[ln 2e3, ln 2e5])), D_k cycling 16, 64,
  256, 1024, X ~ N(0,1)/sqrt(D), w* ~ N(0,1), y ~ Bernoulli(sigmoid(X w*))."""
  ku, kp = R.split(R.PRNGKey(3), 2)
  u = R.uniform(ku, (LADDER_PROBLEMS,)).double().cpu().numpy()
  Ns = np.round(np.exp(np.log(2e3) + u * (np.log(2e5) - np.log(2e3)))).astype(np.int64)
  pk = R.split(kp, LADDER_PROBLEMS)
  out = []
  for k in range(LADDER_PROBLEMS):
    Dk = LADDER_DIMS[k % len(LADDER_DIMS)]
    kx, kw, ky = R.split(pk[k].clone(), 3)
    X = R.normal(kx, (int(Ns[k]) * Dk,)).reshape(int(Ns[k]), Dk).mul_(1.0 / np.sqrt(Dk))
    w = R.normal(kw, (Dk,))
    y = (R.uniform(ky, (int(Ns[k]),)) < torch.sigmoid(X @ w)).to(torch.float32)
    out.append((X, y))
  return out


def run_ladder(args):
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, fused, random as R, scheduler
  L.require_gpu()
  log("ladder data")
  probs = ladder_problems(torch, R)
  C = LADDER_PROBLEMS * LADDER_RUNGS
  temps = [20.0 ** (j / (LADDER_RUNGS - 1)) for j in range(LADDER_RUNGS)]
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(2), C))
  npairs = LADDER_PROBLEMS * (LADDER_RUNGS - 1)
  skeys = R.key_words(R.split(R.PRNGKey(4), npairs))
  chains, pairs, ptemps, parity = [], [], [], []
  for k, (X, y) in enumerate(probs):
    for j in range(LADDER_RUNGS):
      c = k * LADDER_RUNGS + j
      chains.append(dict(X=X, y=y, D=X.shape[1], n=LADDER_BATCH,
                         theta0=np.zeros(X.shape[1], np.float32), key=tuple(keys[c]),
                         dkey=tuple(dkeys[c]), temperature=temps[j], save=False))
      if j + 1 < LADDER_RUNGS:
        pairs.append((c, c + 1))
        ptemps.append((temps[j], temps[j + 1]))
        parity.append(j % 2)
  K, W = args.steps, args.warmup
  G = LADDER_SWAP_EVERY
  total = K + W + 2 * G
  eps = scheduler.polynomial_table(1e-3, 1e-4, total)
  eng = fused.FusedGLM(chains, "logistic", "sgld", eps, prior_scale=1.0)
  tuned = eng.autotune(reps=3)
  log(f"autotune: {tuned}")
  ex = fused.ReplicaExchange(eng, pairs, ptemps, keys=[tuple(k) for k in skeys], swap_every=G,
                             parity=parity)
  g = eng.capture(G)
  for _ in range(max(1, W // G)):
    g.replay()
  torch.cuda.synchronize()
  reps = max(1, K // G)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    g.replay()
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1)
  iters = reps * G
  value = C * iters / (ms / 1e3)
  bytes_it = sum(LADDER_RUNGS * (LADDER_BATCH * (4 * X.shape[1] + 8) + 8 * X.shape[1])
                 for X, _ in probs)
  achieved = bytes_it / (ms / 1e3 / iters) / 1e9
  out = {"metric": "chain-steps/sec, reSGLD ladders (fused, 1 MI355X)", "value": value,
         "unit": "chain-steps/s", "n_gpus": 1, "steps": iters, "warmup": W,
         "ms_per_step": ms / iters, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "fp32",
         "data": "synthetic (jax-compatible Threefry draws from PRNGKey(3))",
         "config": {"workload": "reSGLD: 32 logistic problems x 8 rungs (T_j = 20^(j/7)), "
                                "swap every 50 steps", "chains": C, "batch": LADDER_BATCH,
                    "N": [int(X.shape[0]) for X, _ in probs],
                    "D": [int(X.shape[1]) for X, _ in probs],
                    "step_size": "polynomial 1e-3 -> 1e-4",
                    "target_bytes_per_block": tuned["target_bytes"], "grid_blocks": eng.nblocks},
         "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                      "kernel": "glm_fused_kernel<-1> (mixed D classes) + swap_kernel",
                      "bytes_per_iteration": bytes_it},
         "exchanges": int(ex.swaps.sum().item()), "tests": int(ex.step.sum().item())}
  if not args.no_cpu:
    log("CPU baseline (oracle C port per chain)")
    out["cpu_baseline"] = ladder_cpu_baseline(probs, temps, keys, dkeys, eps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def ladder_cpu_baseline(probs, temps, keys, dkeys, eps, target_s=10.0):
  """The oracle's C port of SGLD for every chain of the ladders, one chain per task on a
  thread pool (ctypes releases the interpreter lock during the call). The swap tests,
  one per pair every 50 iterations, are a negligible share of the work and are left
  out."""
  from concurrent.futures import ThreadPoolExecutor
  from oracle import cport
  threads = max(1, min(16, os.cpu_count() or 1))
  Xs = [np.ascontiguousarray(X.cpu().numpy(), np.float32) for X, _ in probs]
  ys = [np.ascontiguousarray(y.cpu().numpy(), np.float32) for _, y in probs]
  chains = [(k, j) for k in range(len(probs)) for j in range(len(temps))]

  def one(c, iters):
    k, j = chains[c]
    th = np.zeros((1, Xs[k].shape[1]), np.float32)
    kk = np.array([keys[c]], np.uint32)
    dd = np.array([dkeys[c]], np.uint32)
    cport.sgld(1, Xs[k], ys[k], LADDER_BATCH, th, kk, dd, np.resize(eps, iters),
               temperature=temps[j], threads=1)

  def go(iters):
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
      list(ex.map(lambda c: one(c, iters), range(len(chains))))
    return time.perf_counter() - t0

  t1 = go(50)  # long enough that the pool's start-up does not dominate
  iters = int(max(50, min(20000, target_s * 50 / max(t1, 1e-6))))
  dt = go(iters)
  return {"value": len(chains) * iters / dt, "unit": "chain-steps/s", "cores": threads,
          "kind": "port",
          "sample": f"all {len(chains)} chains x {iters} iterations (oracle/cport.c SGLD, one "
                    f"chain per task on {threads} threads, {dt:.1f} s; the swap tests left out)"}


def mlp_data(torch, R, M):
  """Inputs U[0,1]^3072 [50000]; labels = argmax of a fixed random teacher MLP."""
  D_in, H1, H2 = MLP_DIMS
  kx, kt = R.split(R.PRNGKey(5), 2)
  X = R.uniform(kx, (MLP_N * D_in,)).reshape(MLP_N, D_in)
  t = torch.as_tensor(M.init_params(99, D_in, H1, H2), device=X.device)
  W1, W2, W3, b1, b2, b3 = torch.split(t, [D_in * H1, H1 * H2, H2 * 10, H1, H2, 10])
  h = torch.relu(X @ W1.view(D_in, H1) + b1)
  h = torch.relu(h @ W2.view(H1, H2) + b2)
  labels = torch.argmax(h @ W3.view(H2, 10) + b3, dim=1).to(torch.int32)
  return X, labels


def run_mlp(args):
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, mlp as M, random as R, scheduler
  L.require_gpu()
  D_in, H1, H2 = MLP_DIMS
  X, labels = mlp_data(torch, R, M)
  C = args.chains or MLP_CHAINS
  th0 = np.stack([M.init_params(10 + c, D_in, H1, H2) for c in range(C)])
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(2), C))
  K, W = args.steps, args.warmup
  G = 10
  eps = scheduler.polynomial_table(1e-4, 1e-5, K + W + 2 * G)
  prec = args.precision or "bf16"
  eng = M.FusedMLP(X, labels, MLP_DIMS, MLP_BATCH, th0, eps, keys=keys, dkeys=dkeys, rms=True,
                   prior_scale=1.0, precision=prec)
  g = eng.capture(G)
  for _ in range(max(1, W // G)):
    g.replay()
  torch.cuda.synchronize()
  reps = max(1, K // G)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    g.replay()
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1)
  iters = reps * G
  value = C * iters / (ms / 1e3)
  macs = D_in * H1 + H1 * H2 + H2 * 10
  flops = 2 * MLP_BATCH * macs * 2 + 2 * MLP_BATCH * (H1 * H2 + H2 * 10)
  achieved = C * flops / (ms / 1e3 / iters) / 1e12
  out = {"metric": "chain-steps/sec, pSGLD Bayesian MLP (fused bf16 MFMA, 1 MI355X)",
         "value": value, "unit": "chain-steps/s", "n_gpus": 1, "steps": iters, "warmup": W,
         "ms_per_step": ms / iters, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": prec,
         "data": "synthetic (U[0,1] inputs from PRNGKey(5), teacher-MLP labels)",
         "config": {"workload": "pSGLD (rms_prop 0.9, 1e-5) MLP 3072-512-512-10, N=50000, "
                                "batch 256, 16 chains", "chains": C, "P": eng.P,
                    "batch": MLP_BATCH, "N": MLP_N},
         "roofline": {"bound": "mfma", "achieved": achieved, "peak": 2500.0,
                      "unit": "TFLOP/s", "frac": achieved / 2500.0, "traffic": None,
                      "kernel": "all kernels of sgmc_mlp_step (GEMM flops per iteration)",
                      "flops_per_chain_step": flops}}
  if not args.no_plain:
    log("plain fp32 torch path (generic potential, one chain)")
    from jax_sgmc_amd import alias, data as DT, likelihoods as LK, potential as PT
    loader = DT.DeviceNumpyDataLoader(x=X, y=labels)
    pot = PT.minibatch_potential(LK.GaussianPrior(1.0), M.MLPSoftmax(), strategy="vmap")
    run = alias.sgld(pot, loader, cache_size=1, batch_size=MLP_BATCH, rms_prop=True,
                     first_step_size=1e-4, last_step_size=1e-5, accepted_samples=1,
                     save_to_numpy=False)
    p0 = {k: torch.as_tensor(v, device=X.device) for k, v in
          zip(("W1", "W2", "W3", "b1", "b2", "b3"),
              np.split(th0[0], np.cumsum([D_in * H1, H1 * H2, H2 * 10, H1, H2])))}
    p0 = {k: v.reshape(M.leaf_shapes(D_in, H1, H2)[k]) for k, v in p0.items()}
    run(p0, iterations=3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = 20
    run(p0, iterations=its)
    torch.cuda.synchronize()
    plain = its / (time.perf_counter() - t0)
    out["plain"] = {"value": plain, "unit": "chain-steps/s", "chains": 1,
                    "path": "generic torch fp32 potential (vmap + autograd) with the "
                            "alias.sgld rms_prop loop"}
    out["fused_over_plain"] = value / plain
  if not args.no_cpu:
    log("CPU baseline (numpy oracle, one chain)")
    out["cpu_baseline"] = mlp_cpu_baseline(X.cpu().numpy(), labels.cpu().numpy(), th0[0], eps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def mlp_cpu_baseline(X_host, labels_host, theta0, eps, target_s=10.0):
  """The numpy oracle's pSGLD (float64 MLP forward / backward, jax-compatible noise) on
  one chain of the same workload; numpy's BLAS threads do the matrix products."""
  from oracle import model as om
  from oracle import steps as ost
  D_in, H1, H2 = MLP_DIMS
  fam = om.Family("mlp", X_host, labels_host, dims=MLP_DIMS + (10,))
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  leaves = [D_in * H1, H1 * H2, H2 * 10, H1, H2, 10]

  def go(iters):
    t0 = time.perf_counter()
    ost.sgld(pot, np.asarray(theta0, np.float64), iters, eps[:iters], 1.0, MLP_N, MLP_BATCH,
             leaves, rms=(0.9, 1e-5))
    return time.perf_counter() - t0

  t1 = go(1)
  iters = int(max(1, min(len(eps), target_s / max(t1, 1e-6))))
  dt = go(iters)
  threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
  return {"value": iters / dt, "unit": "chain-steps/s", "cores": threads, "kind": "port",
          "sample": f"1 chain x {iters} iterations (oracle/model.py + oracle/steps.py pSGLD "
                    f"in float64 numpy, {dt:.1f} s)"}


def run_sggmc_mlp(args):
  """SGGMC on the MLP 3072-1024-1024-10: OBABO proposals of 100 steps (two minibatches
  of 512 per step), then the float64 full-data potential over 50,000 rows and the MH
  test; chains advance together."""
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, mlp as M, random as R, scheduler
  L.require_gpu()
  dims = (3072, 1024, 1024)
  X, labels = mlp_data(torch, R, M)
  C, n, steps = args.chains or 4, 512, 100
  th0 = np.stack([M.init_params(10 + c, *dims) for c in range(C)])
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(2), C))
  K, W = max(1, args.steps), args.warmup
  # a constant step size (a polynomial table with equal first and last step sizes has
  # NaN coefficients, as the reference's schedule does)
  eps = np.full(K + W + 2, 1e-4, np.float32)
  prec = args.precision or "fp8"
  eng = M.FusedMLP(X, labels, dims, n, th0, eps, keys=keys, dkeys=dkeys, prior_scale=1.0,
                   integrator="sggmc", friction=1.0, steps=steps, precision=prec)
  g = eng.capture(1)
  for _ in range(max(1, W)):
    g.replay()
  torch.cuda.synchronize()
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(K):
    g.replay()
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1)
  value = C * K / (ms / 1e3)
  D_in, H1, H2 = dims
  macs = D_in * H1 + H1 * H2 + H2 * 10
  grad_flops = 2 * n * macs * 2 + 2 * n * (H1 * H2 + H2 * 10)
  full_flops = 2 * MLP_N * macs
  flops = 2 * steps * grad_flops + full_flops
  achieved = C * flops / (ms / 1e3 / K) / 1e12
  peak = 5000.0 if prec == "fp8" else 2500.0
  out = {"metric": "chain-steps/sec, SGGMC Bayesian MLP (fused MFMA, 1 MI355X)",
         "value": value, "unit": "chain-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
         "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": prec,
         "data": "synthetic (U[0,1] inputs from PRNGKey(5), teacher-MLP labels)",
         "config": {"workload": "SGGMC: OBABO 100 steps per MH test, MLP 3072-1024-1024-10, "
                                "batch 512, N=50000, float64 full-data potential",
                    "chains": C, "P": eng.P, "batch": n, "N": MLP_N, "obabo_steps": steps},
         "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                      "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                      "kernel": "all kernels of one SGGMC iteration (GEMM flops)",
                      "flops_per_chain_step": flops},
         "accepted": eng.accepted.tolist()}
  if not args.no_cpu:
    log("CPU baseline (numpy oracle, sampled)")
    out["cpu_baseline"] = sggmc_mlp_cpu_baseline(X.cpu().numpy(), labels.cpu().numpy(), th0[0],
                                                 dims, n, steps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def sggmc_mlp_cpu_baseline(X_host, labels_host, theta0, dims, n, steps, sample_steps=2):
  """The numpy oracle's SGGMC iteration on one chain, timed on a bounded sample: two
  OBABO steps (four minibatch gradients and their noise) and one full-data potential,
  scaled to an iteration of ``steps`` OBABO steps plus the full-data pass."""
  from oracle import model as om
  from oracle import steps as ost
  D_in, H1, H2 = dims
  fam = om.Family("mlp", X_host, labels_host, dims=tuple(dims) + (10,))
  pot = lambda th, idx: om.stochastic_potential(fam, th, idx, 1.0)
  leaves = [D_in * H1, H1 * H2, H2 * 10, H1, H2, 10]
  th = np.asarray(theta0, np.float64)
  ones = np.ones(th.size)
  t0 = time.perf_counter()
  ost.obabo_integrate(pot, th, np.zeros(th.size), np.array([0, 1], np.uint32),
                      np.array([0, 2], np.uint32), np.float32(1e-4), 1.0, 1.0, sample_steps,
                      MLP_N, n, leaves, ones, ones)
  t_steps = time.perf_counter() - t0
  t0 = time.perf_counter()
  om.full_potential(fam, th, 1.0, 5000)
  t_full = time.perf_counter() - t0
  per_iter = t_steps * steps / sample_steps + t_full
  threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
  return {"value": 1.0 / per_iter, "unit": "chain-steps/s", "cores": threads, "kind": "port",
          "sample": f"1 chain: {sample_steps} OBABO steps ({t_steps:.1f} s) and one full-data "
                    f"potential ({t_full:.1f} s) of the numpy oracle in float64, scaled to an "
                    f"iteration of {steps} steps ({per_iter:.0f} s)"}


def run_sgld_linreg(args):
  """SGLD on Bayesian linear regression, N=10,000, D=32, batch 128, one chain, through
  the fused engine's loop launch (the chain's iterations back to back in one block)."""
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, fused, random as R, scheduler
  L.require_gpu()
  N, Dl, n = 10_000, 32, 128
  kx, kw, ke = R.split(R.PRNGKey(0), 3)
  X = R.normal(kx, (N * Dl,)).reshape(N, Dl)
  w = R.normal(kw, (Dl,))
  y = X @ w + 0.5 * R.normal(ke, (N,))
  K, W, G = args.steps, args.warmup, 100
  eps = scheduler.polynomial_table(1e-4, 1e-5, K + W + 2 * G)
  eng = fused.FusedGLM([dict(X=X, y=y.contiguous(), D=Dl, n=n, theta0=np.zeros(Dl, np.float32))],
                       "linear_gaussian", "sgld", eps, sigma=0.5, prior_scale=float(np.sqrt(10.0)))
  # one chain of one block: the loop launch runs all iterations back to back
  assert eng.loop_capable()
  eng.run(max(1, W), loop=True)
  torch.cuda.synchronize()
  reps = max(1, K // G)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  eng.run(reps * G, loop=True)
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1)
  iters = reps * G
  value = iters / (ms / 1e3)
  bytes_step = n * (4 * Dl + 8) + 8 * Dl
  out = {"metric": "chain-steps/sec, SGLD linear regression (one chain, 1 MI355X)",
         "value": value, "unit": "chain-steps/s", "n_gpus": 1, "steps": iters, "warmup": W,
         "ms_per_step": ms / iters, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "fp32",
         "data": "synthetic (jax-compatible Threefry draws from PRNGKey(0))",
         "config": {"workload": "SGLD linear regression N=10000, D=32, batch 128, 1 chain",
                    "step_size": "polynomial 1e-4 -> 1e-5",
                    "launch": "sgmc_glm_fused_loop, up to 1000 iterations per launch"},
         "roofline": {"bound": "hbm", "achieved": bytes_step / (ms / 1e3 / iters) / 1e9,
                      "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": bytes_step / (ms / 1e3 / iters) / 1e9 / HBM_PEAK_GBS,
                      "traffic": None, "kernel": "glm_loop_kernel (latency bound: one block)",
                      "bytes_per_chain_step": bytes_step}}
  if not args.no_cpu:
    log("CPU baseline (oracle C port, one chain)")
    out["cpu_baseline"] = linreg_cpu_baseline(X.cpu().numpy(), y.cpu().numpy(), eps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def linreg_cpu_baseline(X_host, y_host, eps_table, target_s=10.0):
  """The oracle's C port of the same single-chain SGLD loop on one host core (a single
  chain has nothing to spread over more)."""
  from oracle import cport
  N, Dl = X_host.shape

  def go(iters):
    theta = np.zeros((1, Dl), np.float32)
    keys = np.zeros((1, 2), np.uint32)
    dkeys = np.zeros((1, 2), np.uint32)
    eps = np.resize(np.asarray(eps_table, np.float32), iters)
    t0 = time.perf_counter()
    cport.sgld(0, X_host, y_host, 128, theta, keys, dkeys, eps, sigma=0.5,
               prior_scale=float(np.sqrt(10.0)), threads=1)
    return time.perf_counter() - t0

  t1 = go(2000)
  iters = int(max(2000, min(2_000_000, target_s * 2000 / max(t1, 1e-6))))
  dt = go(iters)
  return {"value": iters / dt, "unit": "chain-steps/s", "cores": 1, "kind": "port",
          "sample": f"1 chain x {iters} iterations of the same workload (oracle/cport.c "
                    f"SGLD linear regression, {dt:.1f} s)"}


# Full-data MH workloads on the headline's data: leapfrog steps per test, friction and
# the polynomial step-size range.
MH_SETTINGS = {"amagold": (5, 0.25, (1e-4, 5e-5)), "sggmc": (5, 1.0, (1e-4, 5e-5))}


def mh_cpu_baseline(kind, X_host, y_host, theta0, keys, dkeys, eps, target_s=12.0):
  """The oracle's C port of the same MH loop (reversible leapfrog or OBABO, FP64 full
  potential over all rows, MH), OpenMP over chains."""
  from oracle import cport
  from oracle import jaxrandom as jr
  steps, fric, _ = MH_SETTINGS[kind]
  threads = max(1, min(16, os.cpu_count() or 1))
  C = min(CHAINS, threads)
  Xh = np.ascontiguousarray(X_host, np.float32)
  yh = np.ascontiguousarray(y_host, np.float32)

  def go(iters):
    th = np.tile(np.asarray(theta0, np.float32), (C, 1))
    mom = np.zeros((C, D), np.float32)
    kk = np.zeros((C, 2), np.uint32)
    mk = np.zeros((C, 2), np.uint32)
    for c in range(C):
      if kind == "amagold":
        mom[c], kk[c], mk[c] = cport.amagold_init(th[c], np.asarray(keys[c], np.uint32))
      else:
        kk[c], mk[c] = jr.split(np.asarray(keys[c], np.uint32))
    dd = np.array(dkeys[:C], np.uint32).copy()
    U = np.full(C, cport.full_potential(1, Xh, yh, th[0]))
    run = cport.amagold if kind == "amagold" else cport.sggmc
    t0 = time.perf_counter()
    run(1, Xh, yh, BATCH, th, mom, kk, mk, dd, U, eps[:iters], steps, fric, threads=threads)
    return time.perf_counter() - t0

  t1 = go(1)
  iters = int(max(1, min(500, target_s / max(t1, 1e-3))))
  dt = go(iters)
  name = "AMAGOLD" if kind == "amagold" else "SGGMC"
  return {"value": C * iters / dt, "unit": "chain-steps/s", "cores": threads, "kind": "port",
          "sample": f"{C} chains x {iters} iterations of the same workload (oracle/cport.c "
                    f"{name} logistic with its FP64 full-data potential, OpenMP over chains, "
                    f"{dt:.1f} s)"}


def run_mh(args, kind):
  """AMAGOLD or SGGMC on the headline's logistic data (N=1e6, D=512, batch 4096), 64
  chains, 5 leapfrog steps (reversible leapfrog, or OBABO with two minibatches per
  step) per full-data MH test: per iteration 5 (10) fused step launches, one FP64
  full-data potential of all 64 proposals (fp32 matrix cores, X read once) and the MH
  kernel, captured in HIP graphs of 10 iterations."""
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, fused, random as R, scheduler
  L.require_gpu()
  steps, fric, eps_range = MH_SETTINGS[kind]
  name = "AMAGOLD" if kind == "amagold" else "SGGMC"
  log("generating data")
  X, y, w_true = make_data(torch, R)
  torch.cuda.synchronize()
  K, W, G = args.steps, args.warmup, 10
  total = K + W + 2 * G
  eps = scheduler.polynomial_table(eps_range[0], eps_range[1], total)
  keys, dkeys = chain_keys(R, CHAINS)
  # Chains start at the data-generating weights, inside the posterior's bulk: from a
  # distant start the stochastic-gradient energy error rejects nearly every proposal.
  theta0 = w_true.cpu().numpy().astype(np.float32)
  chains = [dict(X=X, y=y, D=D, n=BATCH, theta0=theta0, dkey=tuple(dkeys[c]))
            for c in range(CHAINS)]
  eng = fused.FusedGLM(chains, "logistic", kind, eps, steps=steps, friction=fric,
                       prior_scale=1.0, save_slots=np.arange(total, dtype=np.int32),
                       n_slots=total)
  tuned = eng.autotune(reps=2)
  log(f"autotune: {tuned}")
  test = fused.AmagoldMH if kind == "amagold" else fused.SggmcMH
  mh = test(eng, keys=[tuple(k) for k in keys])
  g = eng.capture(G)
  for _ in range(max(1, W // G)):
    g.replay()
  torch.cuda.synchronize()
  reps = max(1, K // G)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  for _ in range(reps):
    g.replay()
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1)
  iters = reps * G
  done = eng.iteration_count()
  ratios = torch.stack([mh.acceptance_ratios(c)[:done] for c in range(CHAINS)])
  accept_mean = float(ratios.mean().item())
  # per-kernel averages on the same stream, after the timed region
  def avg_us(fn, n=20):
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(n):
      fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1e3 / n
  full_us = avg_us(lambda: mh._full(mh.U_prop))
  step_us = avg_us(lambda: eng._launch(1))
  value = CHAINS * iters / (ms / 1e3)
  full_flops = 2.0 * N_OBS * D * CHAINS
  # the split pass reads X once (4 D + 4 bytes per row) and runs three fp16 products per
  # multiply-add on the matrix cores (about 80 us of matrix time): bound by the X stream
  full_bytes = N_OBS * (4 * D + 4)
  achieved = full_bytes / (full_us / 1e6) / 1e9
  step_bytes = CHAINS * BYTES_PER_CHAIN_GRAD
  proposal = ("reversible-leapfrog steps" if kind == "amagold" else
              "OBABO steps (two minibatch gradients each)")
  out = {"metric": f"chain-steps/sec, {name} logistic regression (fused, 1 MI355X)",
         "value": value, "unit": "chain-steps/s", "n_gpus": 1, "steps": iters, "warmup": W,
         "ms_per_step": ms / iters, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "fp32",
         "data": "synthetic (jax-compatible Threefry draws from PRNGKey(1); random-init chains)",
         "config": {"workload": f"{name} logistic regression, N=1e6, D=512, batch 4096, "
                                f"64 chains, {steps} {proposal} per MH test, "
                                f"friction {fric}, FP64 full-data potential every iteration",
                    "chains": CHAINS, "N": N_OBS, "D": D, "batch": BATCH,
                    "leapfrog_steps": steps, "step_launches": eng.launches, "friction": fric,
                    "step_size": f"polynomial {eps_range[0]} -> {eps_range[1]}",
                    "graph_iterations": G, "grid_blocks": eng.nblocks,
                    "target_bytes_per_block": tuned["target_bytes"],
                    "full_potential_blocks": mh.nblocks},
         "mean_acceptance_ratio": accept_mean,
         "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                      "kernel": "glm_full_chains_split_kernel<2, logistic> "
                                "(sgmc_glm_full_potential_chains_split)",
                      "bytes_per_launch": full_bytes, "avg_launch_us": full_us,
                      "algorithmic_TFLOPs": full_flops / (full_us / 1e6) / 1e12,
                      "fp16_matrix_TFLOPs_executed": 3 * full_flops / (full_us / 1e6) / 1e12},
         "step_kernel": {"kernel": f"glm_fused_kernel<3> ({name} mode)",
                         "avg_launch_us": step_us,
                         "achieved_GBs": step_bytes / (step_us / 1e6) / 1e9,
                         "bytes_per_launch": step_bytes}}
  if not args.no_cpu:
    log("CPU baseline (oracle C port)")
    out["cpu_baseline"] = mh_cpu_baseline(kind, X.cpu().numpy(), y.cpu().numpy(), theta0, keys,
                                          dkeys, eps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def main():
  ap = argparse.ArgumentParser()
  ap.add_argument("--gpus", type=int, default=1)
  ap.add_argument("--steps", type=int, default=1000)
  ap.add_argument("--warmup", type=int, default=50)
  ap.add_argument("--no-cpu", action="store_true")
  ap.add_argument("--no-plain", action="store_true")
  ap.add_argument("--copy-calibration", action="store_true",
                  help="only copy a 512 MiB device buffer 5 times (the known-size kernel "
                       "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
  ap.add_argument("--workload", default="sghmc-logistic",
                  choices=["sghmc-logistic", "resgld-ladder", "psgld-mlp", "sggmc-mlp",
                           "sgld-linreg", "amagold-logistic", "sggmc-logistic"])
  ap.add_argument("--chains", type=int, default=0, help="chains of the sggmc-mlp workload")
  ap.add_argument("--target-bytes", type=int, default=0,
                  help="headline: fixed bytes-per-block target instead of the autotuner "
                       "(for counter collection, which serializes launches)")
  ap.add_argument("--precision", choices=["fp32", "bf16", "fp8"], default=None,
                  help="GEMM operands of the MLP workloads (sggmc-mlp default fp8, psgld-mlp bf16)")
  args = ap.parse_args()
  if args.gpus != 1:
    print(f"bench.py: the benchmark runs on one GPU only; --gpus {args.gpus} is not supported "
          "(chains would shard across GPUs without communication)", file=sys.stderr)
    sys.exit(2)

  if args.workload == "resgld-ladder":
    return run_ladder(args)
  if args.workload == "psgld-mlp":
    return run_mlp(args)
  if args.workload == "sggmc-mlp":
    return run_sggmc_mlp(args)
  if args.workload == "sgld-linreg":
    return run_sgld_linreg(args)
  if args.workload == "amagold-logistic":
    return run_mh(args, "amagold")
  if args.workload == "sggmc-logistic":
    return run_mh(args, "sggmc")
  import torch
  if args.copy_calibration:
    n = 128 << 20
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
      b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n,
                      "launches": 5}), flush=True)
    return
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, fused, random as R, tree as T
  L.require_gpu()

  log("generating data")
  X, y, _ = make_data(torch, R)
  torch.cuda.synchronize()
  K, W = args.steps, args.warmup
  log(f"fused run: {CHAINS} chains, {K} timed iterations")
  eng, fr = run_fused(torch, fused, R, X, y, K, W, target_bytes=args.target_bytes)
  chain_steps = CHAINS * fr["steps"]
  value = chain_steps / (fr["ms"] / 1e3)
  # one leapfrog step of all chains = one launch per table, the tables concurrent
  step_s = fr["ms"] / 1e3 / fr["launches"]
  bytes_step = CHAINS * BYTES_PER_CHAIN_GRAD
  achieved = bytes_step / step_s / 1e9
  out = {
      "metric": METRIC, "value": value, "unit": "chain-steps/s", "n_gpus": 1,
      "steps": fr["steps"], "warmup": W, "ms_per_step": fr["ms"] / fr["steps"],
      "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
      "data": "synthetic (jax-compatible Threefry draws from PRNGKey(1); random-init chains)",
      "config": {"workload": "SGHMC logistic regression, N=1e6, D=512, batch 4096, "
                             "64 chains, 5 leapfrog steps, friction 0.5",
                 "chains": CHAINS, "N": N_OBS, "D": D, "batch": BATCH,
                 "leapfrog_steps": LEAPFROG, "friction": FRICTION,
                 "step_size": f"polynomial {EPS_FIRST} -> {EPS_LAST}",
                 "tables": TABLES, "grid_blocks_per_table": fr["nblocks"],
                 "graph_iterations": 10,
                 "target_bytes_per_block": fr["tuned"]["target_bytes"]},
      "gradient_evals_per_s": value * LEAPFROG,
  }
  traffic = None
  tp = os.path.join(ROOT, "profiles", "traffic_glm_fused_sghmc_logistic.json")
  if os.path.exists(tp):
    with open(tp) as f:
      traffic = json.load(f).get("hbm_bytes_per_step")
  out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"glm_fused_kernel<3> (sgmc_glm_fused_step), {TABLES} tables "
                               f"of {CHAINS // TABLES} chains on {TABLES} streams",
                     "bytes_per_step": bytes_step, "launches_per_step": TABLES,
                     "avg_step_us": step_s * 1e6}

  keys, dkeys = chain_keys(R, CHAINS)
  if not args.no_plain:
    log("plain single-chain path")
    one, th_plain = run_plain(torch, L, R, T, X, y, 64, tuple(keys[0]), tuple(dkeys[0]))
    one(0)
    torch.cuda.synchronize()
    it = 40
    t0 = time.perf_counter()
    for t in range(1, 1 + it):
      one(t)
    torch.cuda.synchronize()
    plain_rate = it / (time.perf_counter() - t0)
    # same chain through both paths: chain 0 of a fresh fused table vs the plain path
    eng2 = fused.FusedGLM([dict(X=X, y=y, D=D, n=BATCH, theta0=np.zeros(D, np.float32),
                                key=tuple(keys[0]), dkey=tuple(dkeys[0]))], "logistic",
                          "sghmc", step_sizes(64), steps=LEAPFROG, friction=FRICTION)
    one2, th2 = run_plain(torch, L, R, T, X, y, 64, tuple(keys[0]), tuple(dkeys[0]))
    for t in range(8):
      one2(t)
    eng2.run(8)
    torch.cuda.synchronize()
    a, b = eng2.theta(0).cpu().numpy(), th2.cpu().numpy()
    rel = float(np.max(np.abs(a - b)) / max(1e-30, float(np.max(np.abs(b)))))
    out["plain"] = {"value": plain_rate, "unit": "chain-steps/s", "chains": 1,
                    "path": "per gradient: draw_batch, 2x gather_rows, glm_potential, "
                            "sghmc_position, sghmc_momentum (eager launches)",
                    "same_chain_max_rel_diff_after_8_iterations": rel}
    out["fused_over_plain"] = value / plain_rate

  if not args.no_cpu:
    log("CPU baseline (oracle C port)")
    X_host = X.cpu().numpy()
    y_host = y.cpu().numpy()
    out["cpu_baseline"] = cpu_baseline(X_host, y_host, keys, dkeys)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


if __name__ == "__main__":
  main()
