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
import os

# The CPU baselines alternate numpy's BLAS thread pool with OpenMP loops of the C port;
# OpenMP threads that spin between parallel regions halve the BLAS speed. The policy is
# read when the first OpenMP runtime loads (torch's, whose symbols the port then binds),
# so it is set before any import.
os.environ.setdefault("OMP_WAIT_POLICY", "passive")

import argparse  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402

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


# ---------------------------------------------------------------- CPU baseline ----
# Protocol (100 warm-up iterations, then the median of 5 timed runs, bounded so that a
# default run stays within minutes): one calibration call, a warm-up of up to 100
# iterations (about warm_s seconds), further warm-up chunks while the time per iteration
# is still falling (at most settle_s seconds), then the median of 5 timed runs of about
# run_s seconds each, on the host's CPU share (OMP_NUM_THREADS, else os.cpu_count()).
# The C port used is the baseline build (vectorized float32 sums with FMA contraction,
# AVX-512 where the host has it).

def host_info():
  model = "unknown"
  try:
    for line in open("/proc/cpuinfo"):
      if line.startswith("model name"):
        model = line.split(":", 1)[1].strip()
        break
  except OSError:
    pass
  from oracle import cport
  return {"cpu": model, "os_cpu_count": os.cpu_count(), "threads": cpu_threads(),
          "isa": "avx512" if cport.host_has_avx512() else "avx2"}


def cpu_threads():
  n = os.cpu_count() or 1
  env = os.environ.get("OMP_NUM_THREADS", "")
  return max(1, min(n, int(env))) if env.isdigit() else n


def cpu_protocol(step, warm_s=3.0, run_s=3.0, runs=5, max_warm=100, min_iters=1,
                 max_iters=1_000_000, settle_s=10.0, chunk_s=1.0, settle_tol=0.03):
  """step(k) runs k more iterations of the workload. Returns the iterations per timed
  run, the warm-up iterations and seconds, and the median and all run times in seconds.

  The warm-up runs chunks of about chunk_s seconds until it has run max_warm
  iterations or warm_s seconds. After it, the host keeps running chunks of about chunk_s seconds while
  a chunk's time per iteration is more than settle_tol below the previous chunk's (the
  host's clocks and page placement are still settling), for at most settle_s seconds,
  so that the timed runs measure the settled host and not its ramp."""
  w0 = time.perf_counter()
  step(1)
  prev = max(time.perf_counter() - w0, 1e-6)
  warm = 0
  while warm < max_warm and time.perf_counter() - w0 < warm_s:
    k = int(max(1, min(max_warm - warm, chunk_s / prev)))
    t0 = time.perf_counter()
    step(k)
    prev = max((time.perf_counter() - t0) / k, 1e-6)
    warm += k
  settle = 0
  while time.perf_counter() - w0 < warm_s + settle_s:
    k = int(max(1, chunk_s / prev))
    t0 = time.perf_counter()
    step(k)
    cur = max((time.perf_counter() - t0) / k, 1e-6)
    settle += k
    falling = cur < prev * (1.0 - settle_tol)
    prev = cur
    if not falling:
      break
  warm_total = time.perf_counter() - w0
  per = int(min(max_iters, max(min_iters, run_s / prev)))
  times = []
  for _ in range(runs):
    t0 = time.perf_counter()
    step(per)
    times.append(time.perf_counter() - t0)
  return {"warmup_iterations": 1 + warm + settle, "warmup_s": round(warm_total, 2),
          "iterations_per_run": per, "runs": runs,
          "median_s": float(np.median(times)), "run_s": [round(t, 4) for t in times]}


def baseline_line(chains_per_iter, prot, what, threads):
  """cpu_baseline object: chain-steps/s of the median run."""
  value = chains_per_iter * prot["iterations_per_run"] / prot["median_s"]
  return {"value": value, "unit": "chain-steps/s", "cores": threads, "kind": "port",
          "sample": f"{what}; warm-up {prot['warmup_iterations']} iterations ({prot['warmup_s']:.1f} s, "
                    f"until the time per iteration stopped falling), then the median of "
                    f"{prot['runs']} timed runs of {prot['iterations_per_run']} iterations "
                    f"({prot['median_s']:.2f} s) on {threads} threads",
          "protocol": prot, "host": host_info()}


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
    tuned = grp.autotune()
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


def run_plain(torch, L, R, T, X, y, iters, key, dkey, theta0=None, eps=None):
  """One chain through the modular kernels: per gradient evaluation an index draw,
  a row gather, the potential/gradient kernel and the elementwise update kernels."""
  dev = X.device
  lay = T.Layout(torch.zeros(D))
  table, nb = lay.leaf_table(dev, L.PAIRS_PER_BLOCK)
  theta = (torch.zeros(D, device=dev) if theta0 is None else
           torch.as_tensor(np.asarray(theta0, np.float32), device=dev).clone())
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
  eps = step_sizes(iters) if eps is None else eps
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


def laplace_logistic(torch, X, y, newton=10):
  """MAP and covariance of the Laplace approximation of the logistic posterior with prior
  N(0, I): Newton's method on the log posterior, its Hessian X^T diag(p (1 - p)) X + I
  formed on the GPU in float32 and solved in float64. At N = 1e6 rows and D = 512 the
  posterior is close to this Gaussian (Bernstein-von Mises)."""
  dev = X.device
  Dx = X.shape[1]
  w = torch.zeros(Dx, dtype=torch.float64, device=dev)
  eye = torch.eye(Dx, dtype=torch.float64, device=dev)
  for _ in range(newton + 1):
    p = torch.sigmoid(X @ w.float())
    g = (X.T @ (y - p)).double() - w
    H = (X.T @ (X * (p * (1 - p))[:, None])).double() + eye
    step = torch.linalg.solve(H, g)
    w = w + step
  return w, torch.linalg.inv(H), float(step.abs().max().item())


def c2_posterior(torch, fused, L, R, T, X, y, keys, dkeys, its=400, burn=200):
  """Posterior moments of the fused 64-chain run and of the plain single-chain path (the
  same chain 0, same keys), each against the Laplace approximation, in posterior standard
  deviations. Chains start at the MAP; the step-size schedule runs over the check's
  iterations; samples after the burn-in are kept."""
  mu, cov, last_step = laplace_logistic(torch, X, y)
  sd = torch.sqrt(torch.diagonal(cov))
  th0 = mu.float().cpu().numpy()
  eps = step_sizes(its)
  slots = np.where(np.arange(its) >= burn, np.arange(its) - burn, -1).astype(np.int32)
  chains = [dict(X=X, y=y, D=D, n=BATCH, theta0=th0, key=tuple(keys[c]), dkey=tuple(dkeys[c]))
            for c in range(CHAINS)]
  eng = fused.FusedGLM(chains, "logistic", "sghmc", eps, steps=LEAPFROG, friction=FRICTION,
                       prior_scale=1.0, save_slots=slots)
  eng.run(its)
  S = torch.stack([eng.sample_trace(c) for c in range(CHAINS)]).double()   # [C, its-burn, D]
  one, th_plain = run_plain(torch, L, R, T, X, y, its, tuple(keys[0]), tuple(dkeys[0]),
                            theta0=th0, eps=eps)
  plain = []
  for t in range(its):
    one(t)
    if t >= burn:
      plain.append(th_plain.clone())
  P = torch.stack(plain).double()
  torch.cuda.synchronize()
  cm = S.mean(1)                                  # chain means [C, D]
  pooled = cm.mean(0)
  se = cm.std(0) / np.sqrt(CHAINS)
  within = ((S - cm[:, None]) ** 2).mean((0, 1))
  rms = lambda v: float(torch.sqrt(torch.mean(v * v)).item())
  z_f = (pooled - mu) / sd
  z_p = (P.mean(0) - mu) / sd
  return {"reference": "Laplace approximation of the logistic posterior (Newton MAP, inverse "
                       "Hessian; last Newton step %.1e)" % last_step,
          "iterations": its, "burn_in": burn, "start": "MAP",
          "step_size": f"polynomial {EPS_FIRST} -> {EPS_LAST} over the {its} iterations",
          "fused": {"chains": CHAINS, "mean_error_rms_sd": rms(z_f),
                    "mean_error_max_sd": float(z_f.abs().max().item()),
                    "mc_standard_error_rms_sd": rms(se / sd),
                    "variance_over_laplace_median": float((within / sd ** 2).median().item())},
          "plain": {"chains": 1, "mean_error_rms_sd": rms(z_p),
                    "mc_standard_error_rms_sd": rms(cm.std(0) / sd)},
          "chain0_fused_vs_plain_mean_max_diff_sd": float(((cm[0] - P.mean(0)) / sd).abs()
                                                          .max().item())}


def cpu_baseline(X_host, y_host, keys, dkeys):
  """The C port of the same SGHMC loop for all 64 chains (OpenMP over chains, each chain's
  row dot products and gradient sums vectorized)."""
  from oracle import cport
  threads = cpu_threads()
  C = CHAINS
  eps = step_sizes(100_000)
  th = np.zeros((C, D), np.float32)
  kk = np.array(keys[:C], np.uint32).copy()
  dd = np.array(dkeys[:C], np.uint32).copy()
  pos = [0]

  def step(k):
    e = eps[pos[0]:pos[0] + k]
    pos[0] += k
    cport.sghmc(1, X_host, y_host, BATCH, th, kk, dd, e, LEAPFROG, FRICTION, threads=threads,
                fast=True)

  prot = cpu_protocol(step, warm_s=10.0, run_s=3.0, max_warm=1_000_000)
  return baseline_line(C, prot, f"all {C} chains of the same workload (oracle/cport.c SGHMC "
                       "logistic, baseline build, OpenMP over chains)", threads)


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


LADDER_GROUPS = ((16,), (64,), (256,), (1024,))
LADDER_STREAMS = (0, 0, 1, 2)


def ladder_tables(torch, fused, R, probs, eps, G, groups=LADDER_GROUPS,
                  stream_of=LADDER_STREAMS, theta0s=None):
  """The 32 ladders as chain tables (fused.TableGroup), one per group of D classes, each
  table with its own replica exchange. Default: one table per D class, so each class's
  kernel is compiled for its row width only (fewer registers than the mixed-class
  kernel); the D = 16 and 64 tables keep one block per chain and run their iterations in
  loop launches cut at the exchange's test iterations, one after the other on one
  stream, beside the D = 256 and 1024 tables (one launch per iteration, a stream each).
  Keys are those of the chains' positions in a single table ordered by problem, so every
  chain and exchange gives what it gives there."""
  C = LADDER_PROBLEMS * LADDER_RUNGS
  temps = [20.0 ** (j / (LADDER_RUNGS - 1)) for j in range(LADDER_RUNGS)]
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(2), C))
  skeys = R.key_words(R.split(R.PRNGKey(4), LADDER_PROBLEMS * (LADDER_RUNGS - 1)))
  engines, exchanges, dims = [], [], []
  for group in groups:
    # a group is a tuple of D classes, or (classes, part, parts): every parts-th problem
    # of those classes from the part-th on
    classes, part, parts = (group, 0, 1) if isinstance(group[0], int) else group
    chains, pairs, ptemps, parity, sk = [], [], [], [], []
    seen = 0
    for k, (X, y) in enumerate(probs):
      D = X.shape[1]
      if D not in classes:
        continue
      seen += 1
      if (seen - 1) % parts != part:
        continue
      for j in range(LADDER_RUNGS):
        c = k * LADDER_RUNGS + j
        th0 = np.zeros(D, np.float32) if theta0s is None else theta0s[k]
        chains.append(dict(X=X, y=y, D=D, n=LADDER_BATCH, theta0=th0,
                           key=tuple(keys[c]), dkey=tuple(dkeys[c]), temperature=temps[j],
                           save=False))
        if j + 1 < LADDER_RUNGS:
          pairs.append((len(chains) - 1, len(chains)))
          ptemps.append((temps[j], temps[j + 1]))
          parity.append(j % 2)
          sk.append(tuple(skeys[k * (LADDER_RUNGS - 1) + j]))
    small = max(classes) <= 64
    eng = fused.FusedGLM(chains, "logistic", "sgld", eps, prior_scale=1.0,
                         target_bytes=(1 << 30) if small else None)
    exchanges.append(fused.ReplicaExchange(eng, pairs, ptemps, keys=sk, swap_every=G,
                                           parity=parity))
    engines.append(eng)
    dims.append(max(classes))
  return fused.TableGroup(engines, stream_of=list(stream_of)), exchanges, dims


def ladder_posterior(torch, fused, R, probs, G, its=2000, burn=500):
  """Posterior check of the cold (T = 1) rungs: each problem's Laplace approximation
  (Newton MAP, inverse Hessian) is the reference. The ladders start at the MAPs and run
  the workload's kernels, exchanges included, with the step sizes 1e-3 -> 1e-4 over
  the check's iterations; each cold chain's state is read after every exchange block of
  G iterations past the burn-in. Errors are in posterior standard deviations."""
  laps = [laplace_logistic(torch, X, y) for X, y in probs]
  from jax_sgmc_amd import scheduler
  eps = scheduler.polynomial_table(1e-3, 1e-4, its + G)
  grp, _, _ = ladder_tables(torch, fused, R, probs, eps, G,
                            theta0s=[mu.float().cpu().numpy() for mu, _, _ in laps])
  grp.capture(G)
  # cold chain of problem k: position of rung 0 of k in its class's table
  where = {}
  for e, eng in enumerate(grp.engines):
    ks = [k for k, (X, _) in enumerate(probs) if int(X.shape[1]) in
          {int(d) for d in eng.D}]
    for m, k in enumerate(ks):
      where[k] = (e, m * LADDER_RUNGS)
  snaps = [[] for _ in probs]
  for b in range(its // G):
    grp.run(G, G)
    torch.cuda.synchronize()
    if (b + 1) * G > burn:
      for k, (e, c) in where.items():
        snaps[k].append(grp.engines[e].theta(c).double().clone())
  z_all, var_ratio, worst = [], [], []
  finite = True
  for k, (mu, cov, _) in enumerate(laps):
    S = torch.stack(snaps[k])
    finite = finite and bool(torch.isfinite(S).all().item())
    sd = torch.sqrt(torch.diagonal(cov))
    z = (S.mean(0) - mu) / sd
    z_all.append(z)
    var_ratio.append(float((S.var(0) / sd ** 2).median().item()))
    worst.append(float(z.abs().max().item()))
  z = torch.cat(z_all)
  n_s = len(snaps[0])
  return {"reference": "Laplace approximation of each problem's logistic posterior",
          "iterations": its, "burn_in": burn, "start": "MAP",
          "samples_per_problem": n_s,
          "step_size": f"polynomial 1e-3 -> 1e-4 over the {its} iterations",
          "cold_chain_states_finite": finite,
          "mean_error_rms_sd": float(torch.sqrt(torch.mean(z * z)).item()),
          "mean_error_max_sd_per_problem": [round(w, 3) for w in worst],
          "independent_sample_standard_error_sd": 1.0 / float(np.sqrt(n_s)),
          "variance_over_laplace_median_per_problem": [round(v, 3) for v in var_ratio]}


def run_ladder(args):
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, fused, random as R, scheduler
  L.require_gpu()
  log("ladder data")
  probs = ladder_problems(torch, R)
  C = LADDER_PROBLEMS * LADDER_RUNGS
  K, W = args.steps, args.warmup
  G = LADDER_SWAP_EVERY
  warm_cap = 5000  # time-bounded warm-up (below): at most this many extra iterations
  total = K + W + warm_cap + 2 * G
  eps = scheduler.polynomial_table(1e-3, 1e-4, total)
  grp, exchanges, dims = ladder_tables(torch, fused, R, probs, eps, G)
  big = [i for i, e in enumerate(grp.engines) if not e.loop_capable()]
  tuned = grp.autotune(tables=big)
  log(f"autotune: {tuned}")
  assert all(e.loop_capable() == (d <= 64) for e, d in zip(grp.engines, dims))
  grp.capture(G)
  grp.run(max(1, W // G) * G, G)
  torch.cuda.synchronize()
  # warm-up continued for at least half a second (the device's clocks settle under
  # load), in whole iteration blocks, within the schedule
  t0, warm = time.perf_counter(), 0
  while time.perf_counter() - t0 < 0.5 and warm + G <= warm_cap:
    grp.run(G, G)
    torch.cuda.synchronize()
    warm += G
  reps = max(1, K // G)
  e0 = torch.cuda.Event(enable_timing=True)
  e1 = torch.cuda.Event(enable_timing=True)
  e0.record()
  grp.run(reps * G, G)
  e1.record()
  torch.cuda.synchronize()
  ms = e0.elapsed_time(e1)
  iters = reps * G
  value = C * iters / (ms / 1e3)
  bytes_it = sum(LADDER_RUNGS * (LADDER_BATCH * (4 * X.shape[1] + 8) + 8 * X.shape[1])
                 for X, _ in probs)
  achieved = bytes_it / (ms / 1e3 / iters) / 1e9
  tables = [{"D": d, "chains": e.C, "grid_blocks": e.nblocks,
             "launch": "loop (50 iterations, then the exchange)" if e.loop_capable() else
                       "one per iteration, exchange every 50",
             "kernel": (f"glm_loop_kernel<{e.uniform_dclass}>" if e.loop_capable() else
                        f"glm_fused_kernel<{e.uniform_dclass}>"),
             "target_bytes_per_block": e.target_bytes} for e, d in zip(grp.engines, dims)]
  out = {"metric": "chain-steps/sec, reSGLD ladders (fused, 1 MI355X)", "value": value,
         "unit": "chain-steps/s", "n_gpus": 1, "steps": iters, "warmup": W,
         "ms_per_step": ms / iters, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "fp32", "warmup_extra_iterations": warm,
         "data": "synthetic (jax-compatible Threefry draws from PRNGKey(3))",
         "config": {"workload": "reSGLD: 32 logistic problems x 8 rungs (T_j = 20^(j/7)), "
                                "swap every 50 steps", "chains": C, "batch": LADDER_BATCH,
                    "N": [int(X.shape[0]) for X, _ in probs],
                    "D": [int(X.shape[1]) for X, _ in probs],
                    "step_size": "polynomial 1e-3 -> 1e-4", "tables": tables},
         "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                      "kernel": "all tables of one iteration (glm_fused_kernel<4>, <2>, "
                                "glm_loop_kernel<1>, <0>, swap_kernel)",
                      "bytes_per_iteration": bytes_it},
         "exchanges": int(sum(int(ex.swaps.sum().item()) for ex in exchanges)),
         "tests": int(sum(int(ex.step.sum().item()) for ex in exchanges)),
         "chains_finite_after_timed_run": int(sum(
             sum(bool(torch.isfinite(e.theta(c)).all().item()) for c in range(e.C))
             for e in grp.engines))}
  if not args.no_posterior:
    log("posterior check (cold rungs against each problem's Laplace approximation)")
    out["posterior"] = ladder_posterior(torch, fused, R, probs, G)
  if not args.no_cpu:
    log("CPU baseline (oracle C port per chain)")
    temps = [20.0 ** (j / (LADDER_RUNGS - 1)) for j in range(LADDER_RUNGS)]
    keys = R.key_words(R.split(R.PRNGKey(0), C))
    dkeys = R.key_words(R.split(R.PRNGKey(2), C))
    out["cpu_baseline"] = ladder_cpu_baseline(probs, temps, keys, dkeys, eps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def ladder_cpu_baseline(probs, temps, keys, dkeys, eps):
  """The C port of SGLD for every chain of the ladders (baseline build), one chain per
  task on a thread pool of the host's CPU share (ctypes releases the interpreter lock
  during the call). The swap tests, one per pair every 50 iterations, are a negligible
  share of the work and are left out."""
  from concurrent.futures import ThreadPoolExecutor
  from oracle import cport
  threads = cpu_threads()
  Xs = [np.ascontiguousarray(X.cpu().numpy(), np.float32) for X, _ in probs]
  ys = [np.ascontiguousarray(y.cpu().numpy(), np.float32) for _, y in probs]
  chains = [(k, j) for k in range(len(probs)) for j in range(len(temps))]
  th = [np.zeros((1, Xs[k].shape[1]), np.float32) for k, _ in chains]
  kk = [np.array([keys[c]], np.uint32) for c in range(len(chains))]
  dd = [np.array([dkeys[c]], np.uint32) for c in range(len(chains))]
  pool = ThreadPoolExecutor(threads)

  def one(c, iters):
    k, j = chains[c]
    cport.sgld(1, Xs[k], ys[k], LADDER_BATCH, th[c], kk[c], dd[c], np.resize(eps, iters),
               temperature=temps[j], threads=1, fast=True)

  def step(iters):
    list(pool.map(lambda c: one(c, iters), range(len(chains))))

  # calibration and warm-up on blocks of 10 iterations: the pool's start-up per call
  # stays a small share
  prot = cpu_protocol(lambda k: step(10 * k), warm_s=2.0, run_s=3.0, max_warm=10)
  pool.shutdown()
  prot = dict(prot, warmup_iterations=10 * prot["warmup_iterations"],
              iterations_per_run=10 * prot["iterations_per_run"])
  return baseline_line(len(chains), prot, f"all {len(chains)} chains (oracle/cport.c SGLD, "
                       "baseline build, one chain per task; the swap tests left out)", threads)


def mlp_teacher(torch, M, dims, device):
  """The fixed random teacher MLP whose argmax labels the inputs."""
  D_in, H1, H2 = dims
  t = torch.as_tensor(M.init_params(99, D_in, H1, H2), device=device)
  return torch.split(t, [D_in * H1, H1 * H2, H2 * 10, H1, H2, 10])


def mlp_forward(torch, params, dims, X):
  """Class probabilities of the MLP (float32 torch, for the posterior checks)."""
  D_in, H1, H2 = dims
  W1, W2, W3, b1, b2, b3 = params
  h = torch.relu(X @ W1.view(D_in, H1) + b1)
  h = torch.relu(h @ W2.view(H1, H2) + b2)
  return torch.softmax(h @ W3.view(H2, 10) + b3, dim=1)


def mlp_data(torch, R, M, key=5, rows=MLP_N):
  """Inputs U[0,1]^3072 [rows]; labels = argmax of a fixed random teacher MLP
  3072-512-512-10 (both MLP workloads use these data)."""
  kx, _ = R.split(R.PRNGKey(key), 2)
  X = R.uniform(kx, (rows * MLP_DIMS[0],)).reshape(rows, MLP_DIMS[0])
  p = mlp_forward(torch, mlp_teacher(torch, M, MLP_DIMS, X.device), MLP_DIMS, X)
  return X, torch.argmax(p, dim=1).to(torch.int32)


def split_flat(torch, flat, dims):
  D_in, H1, H2 = dims
  return torch.split(flat, [D_in * H1, H1 * H2, H2 * 10, H1, H2, 10])


def mlp_posterior(torch, M, R, X, labels, th0, keys, dkeys, dims, n, prec, integrator, its, burn,
                  thin, steps=1, eps_range=(1e-4, 1e-5)):
  """Posterior moments of the fused engine in `prec` against the same chains (same
  starts, keys and minibatches) in float32, after the burn-in, every `thin` iterations:
  the posterior predictive of 1024 held-out inputs (teacher labels), the mean potential
  and the per-chain parameter means. Monte Carlo standard errors come from the spread
  over chains."""
  from jax_sgmc_amd import scheduler
  C = th0.shape[0]
  if eps_range[0] == eps_range[1]:
    eps = np.full(its, eps_range[0], np.float32)
  else:
    eps = scheduler.polynomial_table(eps_range[0], eps_range[1], its)
  keep = (np.arange(its) >= burn) & ((np.arange(its) - burn) % thin == 0)
  slots = np.full(its, -1, np.int32)
  slots[keep] = np.arange(int(keep.sum()), dtype=np.int32)
  Xp, yp = mlp_data(torch, R, M, key=6, rows=1024)
  res = {}
  for p in (prec, "fp32"):
    kw = dict(rms=True) if integrator == "sgld" else dict(integrator=integrator, friction=1.0,
                                                          steps=steps)
    eng = M.FusedMLP(X, labels, dims, n, th0, eps, keys=keys, dkeys=dkeys, prior_scale=1.0,
                     save_slots=slots, precision=p, **kw)
    eng.run(its)
    torch.cuda.synchronize()
    S = eng.samples
    assert bool(torch.isfinite(S).all().item())
    probs = torch.stack([torch.stack([mlp_forward(torch, split_flat(torch, S[c, k], dims), dims,
                                                  Xp) for k in range(S.shape[1])]).mean(0)
                         for c in range(C)])             # [C, 1024, 10]
    res[p] = dict(probs=probs, U=eng.samples_U.double(), mean=S.mean(1),
                  accepted=eng.accepted.tolist() if eng.accepted is not None else None)
    del eng, S
  f, q = res[prec], res["fp32"]
  th0_t = torch.as_tensor(th0, device=X.device)
  pf, pq = f["probs"].mean(0), q["probs"].mean(0)
  se = (q["probs"].std(0) / np.sqrt(C)).mean().item() if C > 1 else None
  acc = lambda pr: float((pr.argmax(1) == yp.long()).float().mean().item())
  Uf, Uq = f["U"].mean(1), q["U"].mean(1)
  drift = [float((torch.norm(f["mean"][c] - q["mean"][c]) /
                  torch.norm(q["mean"][c] - th0_t[c])).item()) for c in range(C)]
  out = {"reference": "the same chains with float32 GEMMs (precision fp32)",
         "iterations": its, "burn_in": burn, "thinning": thin, "samples_per_chain":
         int(keep.sum()), "chains": C,
         "predictive_mean_abs_diff": float((pf - pq).abs().mean().item()),
         "predictive_mc_standard_error": se,
         "probe_accuracy": {prec: acc(pf), "fp32": acc(pq)},
         "potential_mean": {prec: float(Uf.mean().item()), "fp32": float(Uq.mean().item()),
                            "mc_standard_error": (float(Uq.std().item() / np.sqrt(C))
                                                  if C > 1 else None)},
         "parameter_mean_diff_over_displacement": float(np.mean(drift))}
  if f["accepted"] is not None:
    out["accepted"] = {prec: f["accepted"], "fp32": q["accepted"]}
  return out


def tile_label(names, key):
  """'G1 k128 s2' for a (slot, code) timing, 'G4+G5 k32 k128' for a joint one."""
  if len(key) == 2:
    i, c = key
    return f"{names[i].split()[0]} k{c >> 4} s{c & 15}"
  _, c4, c5 = key
  return f"G4+G5 k{c4 >> 4} k{c5 >> 4}"


def tune_tiles(eng, args):
  """Times the GEMM tilings on the live problem (FusedMLP.autotune) unless disabled;
  returns the plan and its provenance for the bench line's config."""
  names = ("G1 Xb.W1", "G2 h1.W2", "G3 dA2.W2^T", "G4 h1^T.dA2", "G5 Xb^T.dA1")
  if args.tiles:
    # a fixed plan: "k x s" per GEMM, e.g. 128x2,128x1,32x1,32x1,128x1
    pairs = [t.split("x") for t in args.tiles.split(",")]
    if len(pairs) != 5:
      raise SystemExit("--tiles needs five k-tile x split entries (G1..G5)")
    eng.tune = [int(k) * 16 + int(sp) for k, sp in pairs]
    plan, how = eng.gemm_plan(), {"source": "fixed (--tiles " + args.tiles + ")"}
  elif args.no_autotune:
    plan, how = eng.gemm_plan(), {"source": "built-in"}
  else:
    log("autotune: GEMM k-tile depth and split-K on the live problem")
    plan = eng.autotune()
    b, t = eng.tune_gain
    how = {"source": "autotune", "unit_ms_builtin": b, "unit_ms_tuned": t,
           "unit_ms": {tile_label(names, key): round(ms, 5)
                       for key, ms in sorted(eng.tune_ms.items())}}
  tiles = {nm: {"BM": 64, "BN": 128, "BK": k, "split_k": s} for nm, (k, s) in zip(names, plan)}
  # gW2 runs in gW1's launch when both take the same k-tile
  return dict(how, tiles=tiles, gw1_gw2_one_launch=plan[3] == plan[4])


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

  def timed(eng, iters):
    eng.gemm_tiles = tune_tiles(eng, args)
    g = eng.capture(G)
    for _ in range(max(1, W // G)):
      g.replay()
    torch.cuda.synchronize()
    reps = max(1, iters // G)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
      g.replay()
    e1.record()
    torch.cuda.synchronize()
    assert bool(torch.isfinite(eng.theta).all().item())
    return e0.elapsed_time(e1), reps * G

  eng = M.FusedMLP(X, labels, MLP_DIMS, MLP_BATCH, th0, eps, keys=keys, dkeys=dkeys, rms=True,
                   prior_scale=1.0, precision=prec)
  ms, iters = timed(eng, K)
  gemm_tiles = eng.gemm_tiles
  del eng
  value = C * iters / (ms / 1e3)
  macs = D_in * H1 + H1 * H2 + H2 * 10
  flops = 2 * MLP_BATCH * macs * 2 + 2 * MLP_BATCH * (H1 * H2 + H2 * 10)
  achieved = C * flops / (ms / 1e3 / iters) / 1e12
  peak = {"bf16": 2500.0, "fp8": 5000.0, "fp32": 157.3}[prec]
  out = {"metric": "chain-steps/sec, pSGLD Bayesian MLP (fused MFMA, 1 MI355X)",
         "value": value, "unit": "chain-steps/s", "n_gpus": 1, "steps": iters, "warmup": W,
         "ms_per_step": ms / iters, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": prec,
         "data": "synthetic (U[0,1] inputs from PRNGKey(5), teacher-MLP labels)",
         "config": {"workload": f"pSGLD (rms_prop 0.9, 1e-5) MLP 3072-512-512-10, N=50000, "
                                f"batch 256, {C} chains", "chains": C, "P": int(th0.shape[1]),
                    "batch": MLP_BATCH, "N": MLP_N, "gemm_tiles": gemm_tiles},
         "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                      "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                      "kernel": "all kernels of sgmc_mlp_step (GEMM flops per iteration)",
                      "flops_per_chain_step": flops}}
  if not args.no_plain:
    log("plain path: one chain, float32 GEMMs")
    eng = M.FusedMLP(X, labels, MLP_DIMS, MLP_BATCH, th0[:1], eps, keys=keys[:1],
                     dkeys=dkeys[:1], rms=True, prior_scale=1.0, precision="fp32")
    pms, pits = timed(eng, min(K, 300))
    plain_tiles = eng.gemm_tiles
    del eng
    plain = pits / (pms / 1e3)
    out["plain"] = {"value": plain, "unit": "chain-steps/s", "chains": 1,
                    "path": "the same engine with float32 GEMMs (precision fp32) on one chain",
                    "gemm_tiles": plain_tiles}
    out["fused_over_plain"] = value / plain
    log(f"posterior moments: {prec} against fp32 ({C} chains)")
    out["posterior"] = mlp_posterior(torch, M, R, X, labels, th0, keys, dkeys, MLP_DIMS,
                                     MLP_BATCH, prec, "sgld", its=300, burn=200, thin=10)
  if not args.no_cpu:
    log("CPU baseline (float32 port: BLAS GEMMs, C port noise and updates)")
    out["cpu_baseline"] = mlp_cpu_baseline(X.cpu().numpy(), labels.cpu().numpy(), th0[0],
                                           tuple(keys[0]), tuple(dkeys[0]), eps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def mlp_cpu_baseline(X_host, labels_host, theta0, key, dkey, eps):
  """The float32 port of the same pSGLD iteration (oracle/mlp_port.py: numpy's float32
  BLAS for the matrix products on all threads, the C port's OpenMP loops for the noise
  and the update), one chain at a time: a chain-step uses every core already."""
  from oracle import mlp_port as mp
  threads = cpu_threads()
  port = mp.MLPPort(X_host, labels_host, MLP_DIMS, MLP_BATCH, threads=threads)
  theta = np.array(theta0, np.float32)
  v = np.ones_like(theta)
  k = np.array(key, np.uint32)
  dk = np.array(dkey, np.uint32)
  pos = [0]

  def step(m):
    e = np.resize(eps, pos[0] + m)[pos[0]:]
    pos[0] += m
    port.psgld(theta, v, k, dk, e)

  prot = cpu_protocol(step, warm_s=2.0, run_s=3.0, max_warm=100)
  return baseline_line(1, prot, "1 chain of the same workload (oracle/mlp_port.py pSGLD in "
                       "float32: numpy BLAS GEMMs, C port noise and rms_prop update)", threads)


def run_sggmc_mlp(args):
  """SGGMC on the MLP 3072-1024-1024-10: OBABO proposals of 100 steps (two minibatches
  of 512 per step), then the float64 full-data potential over 50,000 rows and the MH
  test; chains advance together."""
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, mlp as M, random as R
  L.require_gpu()
  dims = (3072, 1024, 1024)
  X, labels = mlp_data(torch, R, M)
  C, n, steps = args.chains or 16, 512, 100
  th0 = np.stack([M.init_params(10 + c, *dims) for c in range(C)])
  keys = R.key_words(R.split(R.PRNGKey(0), C))
  dkeys = R.key_words(R.split(R.PRNGKey(2), C))
  # one iteration of 16 chains takes about 0.18 s: the timed work stays within seconds
  K, W = max(1, min(args.steps, 30)), max(1, min(args.warmup, 3))
  # a constant step size (a polynomial table with equal first and last step sizes has
  # NaN coefficients, as the reference's schedule does)
  eps = np.full(K + W + 2, 1e-4, np.float32)
  prec = args.precision or "fp8"

  def timed(eng, iters):
    eng.gemm_tiles = tune_tiles(eng, args)
    g = eng.capture(1)
    for _ in range(W):
      g.replay()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
      g.replay()
    e1.record()
    torch.cuda.synchronize()
    # the state after the timed iterations is finite (the line checks its result)
    assert bool(torch.isfinite(eng.theta).all().item()) and \
        bool(torch.isfinite(eng.U_full).all().item())
    return e0.elapsed_time(e1)

  eng = M.FusedMLP(X, labels, dims, n, th0, eps, keys=keys, dkeys=dkeys, prior_scale=1.0,
                   integrator="sggmc", friction=1.0, steps=steps, precision=prec)
  ms = timed(eng, K)
  value = C * K / (ms / 1e3)
  D_in, H1, H2 = dims
  macs = D_in * H1 + H1 * H2 + H2 * 10
  grad_flops = 2 * n * macs * 2 + 2 * n * (H1 * H2 + H2 * 10)
  full_flops = 2 * MLP_N * macs
  flops = 2 * steps * grad_flops + full_flops
  achieved = C * flops / (ms / 1e3 / K) / 1e12
  peak = {"bf16": 2500.0, "fp8": 5000.0, "fp32": 157.3}[prec]
  out = {"metric": "chain-steps/sec, SGGMC Bayesian MLP (fused MFMA, 1 MI355X)",
         "value": value, "unit": "chain-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
         "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": prec,
         "data": "synthetic (U[0,1] inputs from PRNGKey(5), teacher-MLP labels)",
         "config": {"workload": "SGGMC: OBABO 100 steps per MH test, MLP 3072-1024-1024-10, "
                                "batch 512, N=50000, float64 full-data potential",
                    "chains": C, "P": eng.P, "batch": n, "N": MLP_N, "obabo_steps": steps,
                    "gemm_tiles": eng.gemm_tiles},
         "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                      "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                      "kernel": "all kernels of one SGGMC iteration (GEMM flops)",
                      "flops_per_chain_step": flops},
         "accepted": eng.accepted.tolist()}
  log("float64 oracle errors of chain 0")
  out["fp64_oracle_error"] = sggmc_mlp_oracle_errors(torch, M, X, labels, dims, n, eng)
  del eng
  if not args.no_plain:
    log("plain path: one chain, float32 GEMMs")
    eng = M.FusedMLP(X, labels, dims, n, th0[:1], eps, keys=keys[:1], dkeys=dkeys[:1],
                     prior_scale=1.0, integrator="sggmc", friction=1.0, steps=steps,
                     precision="fp32")
    pms = timed(eng, min(K, 5))
    plain_tiles = eng.gemm_tiles
    del eng
    plain = min(K, 5) / (pms / 1e3)
    out["plain"] = {"value": plain, "unit": "chain-steps/s", "chains": 1,
                    "path": "the same engine with float32 GEMMs (precision fp32) on one chain",
                    "gemm_tiles": plain_tiles}
    out["fused_over_plain"] = value / plain
    log(f"posterior moments: {prec} against fp32 ({C} chains)")
    out["posterior"] = mlp_posterior(torch, M, R, X, labels, th0, keys, dkeys, dims, n, prec,
                                     "sggmc", its=12, burn=2, thin=1, steps=steps,
                                     eps_range=(1e-4, 1e-4))
  if not args.no_cpu:
    log("CPU baseline (float32 port, sampled)")
    out["cpu_baseline"] = sggmc_mlp_cpu_baseline(X.cpu().numpy(), labels.cpu().numpy(), th0[0],
                                                 dims, n, steps)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def sggmc_mlp_oracle_errors(torch, M, X, labels, dims, n, eng):
  """Chain 0 of the engine against the float64 oracle: the full-data potential the MH test
  kept (float64 sum of the engine's likelihoods) against the oracle's at the same
  parameters, and the potential and gradient of the minibatch the chain's data key draws
  next."""
  from oracle import model as om
  from oracle import steps as ost
  Xh, yh = X.cpu().numpy(), labels.cpu().numpy()
  th = eng.theta[0].double().cpu().numpy()
  fam = om.Family("mlp", Xh, yh, dims=tuple(dims) + (10,))
  U_ref = om.full_potential(fam, th, 1.0, mb=5000)
  U_kept = float(eng.U_full[0].item())
  U_now = float(eng.full_potential()[0].item())
  dk = eng.dkeys[0].cpu().numpy().view(np.uint32).copy()
  _, idx = ost.draw_indices(dk, Xh.shape[0], n)
  Ur, gr, _ = om.stochastic_potential(fam, th, idx, 1.0)
  U_mb, g = eng.potential_and_grad()
  g0 = g[0].double().cpu().numpy()
  return {"full_potential_rel": abs(U_now - U_ref) / abs(U_ref),
          "kept_full_potential_rel": abs(U_kept - U_ref) / abs(U_ref),
          "minibatch_potential_rel": abs(float(U_mb[0].item()) - Ur) / abs(Ur),
          "gradient_rel_norm": float(np.linalg.norm(g0 - gr) / np.linalg.norm(gr)),
          "full_potential_oracle": U_ref}


def sggmc_mlp_cpu_baseline(X_host, labels_host, theta0, dims, n, steps):
  """The float32 port's SGGMC iteration on one chain, timed on a bounded sample: OBABO
  steps (two minibatch gradients and their noise each) by the protocol, the full-data
  potential (float64 sum of float32 batch likelihoods) three times; an iteration is
  `steps` OBABO steps plus one full-data pass."""
  from oracle import mlp_port as mp
  threads = cpu_threads()
  port = mp.MLPPort(X_host, labels_host, dims, n, threads=threads)
  theta = np.array(theta0, np.float32)
  p = np.zeros_like(theta)
  key = np.array([0, 1], np.uint32)
  dkey = np.array([0, 2], np.uint32)
  prot = cpu_protocol(lambda k: port.obabo_steps(theta, p, key, dkey, 1e-4, 1.0, k),
                      warm_s=1.0, run_s=1.0, max_warm=2)
  t_step = prot["median_s"] / prot["iterations_per_run"]
  tf = []
  for _ in range(3):
    t0 = time.perf_counter()
    port.full_potential(theta, 5000)
    tf.append(time.perf_counter() - t0)
  t_full = float(np.median(tf))
  per_iter = t_step * steps + t_full
  return {"value": 1.0 / per_iter, "unit": "chain-steps/s", "cores": threads, "kind": "port",
          "sample": f"1 chain of the float32 port (oracle/mlp_port.py): OBABO steps by the "
                    f"protocol ({t_step * 1e3:.0f} ms each: median of "
                    f"{prot['runs']} runs of {prot['iterations_per_run']}), the full-data "
                    f"potential ({t_full:.2f} s, median of 3); an iteration of {steps} steps "
                    f"and one full-data pass takes {per_iter:.1f} s",
          "protocol": prot, "host": host_info()}


def run_sgld_linreg(args):
  """SGLD on Bayesian linear regression, N=10,000, D=32, batch 128, one chain, through
  the fused engine's loop launch (the chain's iterations back to back in one block);
  the plain path runs the modular kernels (index draw, gathers, potential, SGLD step)
  with one launch each per iteration."""
  import torch
  import __graft_entry__
  __graft_entry__.load()
  from jax_sgmc_amd import _lib as L, fused, random as R, scheduler, tree as T
  L.require_gpu()
  N, Dl, n = 10_000, 32, 128
  kx, kw, ke = R.split(R.PRNGKey(0), 3)
  X = R.normal(kx, (N * Dl,)).reshape(N, Dl)
  w = R.normal(kw, (Dl,))
  y = (X @ w + 0.5 * R.normal(ke, (N,))).contiguous()
  K, W, G = args.steps, args.warmup, 100
  sigma, tau = 0.5, float(np.sqrt(10.0))
  eps = scheduler.polynomial_table(1e-4, 1e-5, K + W + 2 * G)
  eng = fused.FusedGLM([dict(X=X, y=y, D=Dl, n=n, theta0=np.zeros(Dl, np.float32))],
                       "linear_gaussian", "sgld", eps, sigma=sigma, prior_scale=tau)
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
  assert bool(torch.isfinite(eng.theta_flat).all().item())
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
  if not args.no_plain:
    log("plain path (modular kernels) and posterior moments against the analytic posterior")
    one, _ = linreg_plain(torch, L, R, T, X, y, eps, sigma, tau)
    for t in range(20):
      one(t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = 400
    for t in range(its):
      one(t)
    torch.cuda.synchronize()
    plain = its / (time.perf_counter() - t0)
    out["plain"] = {"value": plain, "unit": "chain-steps/s", "chains": 1,
                    "path": "per iteration: draw_batch, 2x gather_rows, glm_potential, "
                            "sgld_step (eager launches)"}
    out["fused_over_plain"] = value / plain
    out["posterior"] = c1_posterior(torch, fused, L, R, T, X, y, sigma, tau)
  if not args.no_cpu:
    log("CPU baseline (oracle C port, one chain)")
    out["cpu_baseline"] = linreg_cpu_baseline(X.cpu().numpy(), y.cpu().numpy(), sigma, tau)
    out["vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
  else:
    out["cpu_baseline"] = None
  print(json.dumps(out), flush=True)


def linreg_plain(torch, L, R, T, X, y, eps, sigma, tau):
  """One chain of langevin_diffusion (integrator.py:859-922) through the modular kernels,
  one launch each: index draw, row and label gathers, potential / gradient, SGLD step."""
  dev = X.device
  N, Dl = X.shape
  n = 128
  lay = T.Layout(torch.zeros(Dl))
  table, nb = lay.leaf_table(dev, L.PAIRS_PER_BLOCK)
  theta = torch.zeros(Dl, device=dev)
  g = torch.empty(Dl, device=dev)
  U = torch.empty(1, device=dev)
  k = [R.key_from_words((0, 0)), torch.empty(2, dtype=torch.int32, device=dev)]
  dk = [R.key_from_words((0, 0)), torch.empty(2, dtype=torch.int32, device=dev)]
  idx = torch.empty(n, dtype=torch.int32, device=dev)
  Xb = torch.empty(n, Dl, device=dev)
  yb = torch.empty(n, device=dev)
  blocks = L.query("sgmc_glm_blocks", n, Dl)
  ws = torch.zeros(int(L.query("sgmc_glm_workspace_floats", n, Dl, blocks)), device=dev)
  s = L.stream()

  def one(t):
    L.call("sgmc_draw_batch", L.ptr(dk[0]), L.ptr(dk[1]), L.ptr(idx), n, N, s)
    dk.reverse()
    L.call("sgmc_gather_rows", L.ptr(X), N, L.ptr(idx), L.ptr(Xb), n, 4 * Dl, 4 * Dl, s)
    L.call("sgmc_gather_rows", L.ptr(y), N, L.ptr(idx), L.ptr(yb), n, 4, 4, s)
    L.call("sgmc_glm_potential", L.LIK_LINEAR_GAUSSIAN, L.ptr(Xb), Dl, L.ptr(yb), None, n, N,
           Dl, L.ptr(theta), sigma, 1, tau, 1.0, L.ptr(U), L.ptr(g), None, None, L.ptr(ws),
           ws.numel(), blocks, s)
    L.call("sgmc_sgld_step", L.ptr(theta), L.ptr(g), None, Dl, L.ptr(table), nb, L.ptr(k[0]),
           L.ptr(k[1]), float(eps[t]), 1.0, 0.9, 1e-5, s)
    k.reverse()

  return one, theta


def c1_posterior(torch, fused, L, R, T, X, y, sigma, tau, its=10_000, burn=2_000, thin=10):
  """The full 10,000-iteration run of the configuration (polynomial 1e-4 -> 1e-5) through
  the fused loop launch and through the plain modular kernels, samples after the burn-in
  (every 10th) against the analytic Gaussian posterior: Lambda = X^T X / s^2 + I / t^2,
  mu = Lambda^-1 X^T y / s^2; errors in posterior standard deviations. The first step
  sizes exceed 2 / lambda_max (the configuration's 1e-4 start), so the chain first grows
  and then contracts once the schedule falls below it; the burn-in covers that."""
  from jax_sgmc_amd import scheduler
  Xh = X.double().cpu().numpy()
  yh = y.double().cpu().numpy()
  Lam = Xh.T @ Xh / sigma ** 2 + np.eye(Xh.shape[1]) / tau ** 2
  cov = np.linalg.inv(Lam)
  mu = cov @ (Xh.T @ yh) / sigma ** 2
  sd = np.sqrt(np.diag(cov))
  eps = scheduler.polynomial_table(1e-4, 1e-5, its)
  keep = (np.arange(its) >= burn) & ((np.arange(its) - burn) % thin == 0)
  slots = np.full(its, -1, np.int32)
  slots[keep] = np.arange(int(keep.sum()), dtype=np.int32)
  eng = fused.FusedGLM([dict(X=X, y=y, D=X.shape[1], n=128, theta0=np.zeros(X.shape[1],
                                                                          np.float32))],
                       "linear_gaussian", "sgld", eps, sigma=sigma, prior_scale=tau,
                       save_slots=slots)
  eng.run(its, loop=True)
  Sf = eng.sample_trace(0).double().cpu().numpy()
  one, theta = linreg_plain(torch, L, R, T, X, y, eps, sigma, tau)
  plain = []
  for t in range(its):
    one(t)
    if keep[t]:
      plain.append(theta.clone())
  Sp = torch.stack(plain).double().cpu().numpy()

  def stats(S):
    z = (S.mean(0) - mu) / sd
    # batch-means standard error of the mean (10 batches of consecutive samples)
    bm = S[: len(S) // 10 * 10].reshape(10, -1, S.shape[1]).mean(1)
    se = bm.std(0, ddof=1) / np.sqrt(10) / sd
    return {"mean_error_rms_sd": float(np.sqrt(np.mean(z * z))),
            "mean_error_max_sd": float(np.abs(z).max()),
            "mc_standard_error_rms_sd": float(np.sqrt(np.mean(se * se))),
            "variance_over_posterior_median": float(np.median(S.var(0) / sd ** 2))}

  return {"reference": "analytic Gaussian posterior of the linear regression",
          "iterations": its, "burn_in": burn, "thinning": thin, "samples": int(keep.sum()),
          "fused": stats(Sf), "plain": stats(Sp),
          "fused_vs_plain_mean_max_diff_sd": float(np.abs((Sf.mean(0) - Sp.mean(0)) / sd).max())}


def linreg_cpu_baseline(X_host, y_host, sigma, tau):
  """The C port of the same single-chain SGLD loop on one host core (one chain of 128
  rows x 32 features per step has too little work to split over threads): the full
  10,000-iteration run per timed run."""
  from oracle import cport
  from jax_sgmc_amd import scheduler
  Dl = X_host.shape[1]
  eps_run = scheduler.polynomial_table(1e-4, 1e-5, 10_000)
  theta = np.zeros((1, Dl), np.float32)
  keys = np.zeros((1, 2), np.uint32)
  dkeys = np.zeros((1, 2), np.uint32)

  def step(k):
    cport.sgld(0, X_host, y_host, 128, theta, keys, dkeys, np.resize(eps_run, k), sigma=sigma,
               prior_scale=tau, threads=1, fast=True)

  prot = cpu_protocol(step, warm_s=1.0, run_s=1.0, min_iters=10_000, max_iters=10_000)
  return baseline_line(1, prot, "1 chain of the same workload (oracle/cport.c SGLD linear "
                       "regression, baseline build)", 1)


# Full-data MH workloads on the headline's data: leapfrog steps per test, friction and
# the polynomial step-size range.
MH_SETTINGS = {"amagold": (5, 0.25, (1e-4, 5e-5)), "sggmc": (5, 1.0, (1e-4, 5e-5))}


def mh_cpu_baseline(kind, X_host, y_host, theta0, keys, dkeys, eps):
  """The C port of the same MH loop (reversible leapfrog or OBABO, FP64 full potential
  over all rows, MH) for all 64 chains, baseline build, OpenMP over chains."""
  from oracle import cport
  from oracle import jaxrandom as jr
  steps, fric, _ = MH_SETTINGS[kind]
  threads = cpu_threads()
  C = CHAINS
  Xh = np.ascontiguousarray(X_host, np.float32)
  yh = np.ascontiguousarray(y_host, np.float32)
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
  pos = [0]

  def step(k):
    e = np.resize(eps, pos[0] + k)[pos[0]:]
    pos[0] += k
    run(1, Xh, yh, BATCH, th, mom, kk, mk, dd, U, e, steps, fric, threads=threads, fast=True)

  prot = cpu_protocol(step, warm_s=1.0, run_s=1.0, max_warm=1)
  name = "AMAGOLD" if kind == "amagold" else "SGGMC"
  return baseline_line(C, prot, f"all {C} chains of the same workload (oracle/cport.c {name} "
                       "logistic with its FP64 full-data potential, baseline build, OpenMP "
                       "over chains)", threads)


def mh_posterior(torch, eng, X, y, done, theta0):
  """Posterior moments of the saved post-test states (every iteration of warm-up and
  timed region, the first half dropped as burn-in), pooled over the chains, against the
  Laplace approximation of the logistic posterior, in posterior standard deviations."""
  mu, cov, last_step = laplace_logistic(torch, X, y)
  sd = torch.sqrt(torch.diagonal(cov))
  burn = done // 2
  S = torch.stack([eng.sample_trace(c)[burn:done] for c in range(eng.C)]).double()
  cm = S.mean(1)
  pooled = cm.mean(0)
  z = (pooled - mu) / sd
  within = ((S - cm[:, None]) ** 2).mean((0, 1))
  rms = lambda v: float(torch.sqrt(torch.mean(v * v)).item())
  z0 = (torch.as_tensor(theta0, dtype=torch.float64, device=mu.device) - mu) / sd
  return {"reference": "Laplace approximation of the logistic posterior (Newton MAP, inverse "
                       "Hessian; last Newton step %.1e)" % last_step,
          "iterations": int(done), "burn_in": int(burn), "chains": int(eng.C),
          "start_error_rms_sd": rms(z0),
          "states_finite": bool(torch.isfinite(S).all().item()),
          "mean_error_rms_sd": rms(z), "mean_error_max_sd": float(z.abs().max().item()),
          "mc_standard_error_rms_sd": rms(cm.std(0) / np.sqrt(eng.C) / sd),
          "variance_over_laplace_median": float((within / sd ** 2).median().item())}


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
  if not args.no_posterior:
    log("posterior check (saved post-test states against the Laplace approximation)")
    out["posterior"] = mh_posterior(torch, eng, X, y, done, theta0)
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
  ap.add_argument("--no-posterior", action="store_true",
                  help="resgld-ladder, amagold-logistic, sggmc-logistic: skip the "
                       "posterior check")
  ap.add_argument("--tiles", default=None,
                  help="MLP workloads: a fixed GEMM plan instead of the tuner, five 'k x s' "
                       "entries for G1..G5 (e.g. 128x2,128x1,32x1,32x1,128x1)")
  ap.add_argument("--no-autotune", action="store_true",
                  help="MLP workloads: keep the built-in GEMM tiling instead of timing the "
                       "candidates on the live problem before the timed region")
  ap.add_argument("--copy-calibration", action="store_true",
                  help="only copy a 512 MiB device buffer 5 times (the known-size kernel "
                       "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
  ap.add_argument("--workload", default="sghmc-logistic",
                  choices=["sghmc-logistic", "resgld-ladder", "psgld-mlp", "sggmc-mlp",
                           "sgld-linreg", "amagold-logistic", "sggmc-logistic"])
  ap.add_argument("--chains", type=int, default=0,
                  help="chains of the MLP workloads (psgld-mlp default 16, sggmc-mlp 16)")
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
    log("posterior moments (fused 64 chains and plain chain 0 against the Laplace posterior)")
    out["posterior"] = c2_posterior(torch, fused, L, R, T, X, y, keys, dkeys)

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
