"""Step-size schedules, SG-MCMC integrators and solvers, restated (oracle only).

The vector arithmetic runs in float64; every scalar coefficient the reference forms
from float32 schedule values (step size, temperature) is formed in float32 exactly
as the reference forms it, and the noise is the float32 jax.random stream
(``oracle.jaxrandom``), so a float32 implementation with the same keys must agree
with these trajectories to float32 rounding.

``pot(theta, idx) -> (U, grad, lik)`` is any potential from ``oracle.model``.
"""
import numpy as np

from oracle import jaxrandom as jr

F32 = np.float32


# ---------------------------------------------------------------- schedules ----

def polynomial_step_size(a, b, gamma, iterations):
  """scheduler.py:469-481: eps_i = a * (b + i)^(-gamma), float32."""
  n = np.arange(iterations, dtype=np.int32)
  unscaled = np.power(F32(b) + n.astype(np.float32), F32(-gamma)).astype(np.float32)
  return (F32(a) * unscaled).astype(np.float32)


def polynomial_step_size_first_last(first, last, gamma, iterations):
  """scheduler.py:512-532: solve a, b so that eps_0 = first, eps_{I-1} = last."""
  assert gamma > 0
  assert first >= last
  gamma = F32(gamma)
  first = F32(first)
  last = F32(last)
  with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
    ginv = np.power(gamma, F32(-1.0))
    fpow = np.power(first, -ginv)
    lpow = np.power(last, -ginv)
    apow = F32((lpow - fpow) / F32(iterations - 1))
    a = np.power(apow, -gamma)
    b = np.power(F32(first / a), -ginv)
    n = np.arange(iterations).astype(np.float32)
    return (a * np.power(b + n, -gamma)).astype(np.float32)


def initial_burn_in(n, iterations):
  """scheduler.py:578-587: 1.0 once iteration >= n."""
  return np.where(n <= np.arange(iterations), F32(1.0), F32(0.0))


# ------------------------------------------------------------------- data ----

def draw_indices(dkey, N, n):
  """DeviceNumpyDataLoader.get_random_data (data/numpy_loader.py:132-134)."""
  dkey, sub = jr.split(dkey)
  return dkey, jr.randint(sub, (n,), 0, N)


# ------------------------------------------------------------ integrators ----

def sgld(pot, theta0, iterations, eps, temperature, N, n, leaf_sizes,
         key=None, dkey=None, rms=None):
  """langevin_diffusion (integrator.py:859-922), optional rms_prop
  (adaption.py:254-291, decorator at 120-222; v0 = ones, update before get)."""
  key = jr.PRNGKey(0) if key is None else np.asarray(key, np.uint32)
  dkey = jr.PRNGKey(0) if dkey is None else np.asarray(dkey, np.uint32)
  theta = np.asarray(theta0, np.float64).copy()
  eps = np.broadcast_to(np.asarray(eps, np.float32), (iterations,))
  temp = np.broadcast_to(np.asarray(temperature, np.float32), (iterations,))
  v = np.ones_like(theta) if rms is not None else None
  out = {"theta": [], "U": [], "variance": [], "idx": [], "noise": [], "grad": []}
  for t in range(iterations):
    key, sub = jr.split(key)
    dkey, idx = draw_indices(dkey, N, n)
    xi = jr.random_tree_flat(sub, leaf_sizes).astype(np.float64)
    U, g, lik = pot(theta, idx)
    var = np.var(lik)
    e = eps[t]
    sg = float(-e) * g
    sn = float(np.sqrt(F32(F32(2) * temp[t]) * e)) * xi
    if rms is None:
      theta = theta + (sg + sn)
    else:
      alpha, lmbd = rms
      v = alpha * v + (1.0 - alpha) * g * g
      G = 1.0 / (lmbd + np.sqrt(v))
      theta = theta + ((float(e) * 0.0 + G * sg) + np.sqrt(G) * sn)
    for k, val in (("theta", theta), ("U", U), ("variance", var), ("idx", idx),
                   ("noise", xi), ("grad", g)):
      out[k].append(np.copy(val))
  res = {k: np.asarray(v_) for k, v_ in out.items()}
  res["key"], res["dkey"] = key, dkey
  if rms is not None:
    res["v"] = v
  return res


def fisher_noise(G, g, N, n, friction, eps, diagonal=True):
  """adaption.fisher_information get (adaption.py:403-455): G [n, P] gradients of the
  observations' likelihoods, g the potential's gradient. diff = G - g / N;
  v = sum diff^2 / (n - 1) (outer products when not diagonal); b = 0.5 eps v.
  Diagonal: c = friction - b with entries c <= 0 replaced by the smallest positive
  entry; returns (sqrt(c), sqrt(friction - c)). Full: (U_cb sqrt(s_cb), U sqrt(s)) from
  the SVDs of diag(friction) - b and of b."""
  G = np.asarray(G, np.float64)
  diff = G - np.asarray(g, np.float64) / N
  e = float(F32(eps))
  friction = np.broadcast_to(np.asarray(friction, np.float64), (G.shape[1],))
  if diagonal:
    b = 0.5 * e * (np.sum(diff * diff, axis=0) / (n - 1))
    corr = friction - b
    pos = corr[corr > 0]
    smallest = pos.min() if pos.size else np.inf
    c = np.where(corr <= 0, smallest, corr)
    return np.sqrt(c), np.sqrt(friction - c)
  b = 0.5 * e * (diff.T @ diff / (n - 1))
  u_cb, s_cb, _ = np.linalg.svd(np.diag(friction) - b)
  u, s_, _ = np.linalg.svd(b)
  return u_cb @ np.sqrt(s_cb), u @ np.sqrt(s_)


def sghmc(pot, theta0, iterations, eps, steps, friction, N, n, leaf_sizes,
          key=None, dkey=None, inv_mass=None, sqrt_mass=None, obs_grads=None):
  """friction_leapfrog.integrate (integrator.py:715-757) with _body_fun (603-666):
  momentum resampled each iteration, then ``steps`` x {position, friction decay,
  gradient, noise}. The noise is C * sqrt(2 eps) xi, or, with
  ``obs_grads(theta, idx) -> [n, P]`` (the diagonal Fisher noise model,
  adaption.py:371-457), sqrt(2 eps) (cb_diff_sqrt xi) with cb_diff_sqrt from
  ``fisher_noise`` at the new position."""
  key = jr.PRNGKey(0) if key is None else np.asarray(key, np.uint32)
  dkey = jr.PRNGKey(0) if dkey is None else np.asarray(dkey, np.uint32)
  theta = np.asarray(theta0, np.float64).copy()
  P = theta.size
  inv_mass = np.ones(P) if inv_mass is None else np.asarray(inv_mass, np.float64)
  sqrt_mass = np.ones(P) if sqrt_mass is None else np.asarray(sqrt_mass, np.float64)
  C = np.broadcast_to(np.asarray(friction, np.float32), (P,)).astype(np.float64)
  Cf = np.broadcast_to(np.asarray(friction, np.float32), (P,))
  eps = np.broadcast_to(np.asarray(eps, np.float32), (iterations,))
  out = {"theta": [], "p": [], "U": [], "idx": []}
  for t in range(iterations):
    e = eps[t]
    key, sub = jr.split(key)
    p = sqrt_mass * jr.random_tree_flat(sub, leaf_sizes).astype(np.float64)
    neg_eC = (F32(-e) * Cf).astype(np.float64)
    noise_scale = float(np.sqrt(F32(2) * e))
    U = 0.0
    for _ in range(steps):
      s = inv_mass * p
      theta = theta + float(e) * s
      p = p + neg_eC * s
      dkey, idx = draw_indices(dkey, N, n)
      U, g, _ = pot(theta, idx)
      p = p + float(-e) * g
      key, sub = jr.split(key)
      xi = jr.random_tree_flat(sub, leaf_sizes).astype(np.float64)
      if obs_grads is None:
        p = p + C * (noise_scale * xi)
      else:
        cb, _ = fisher_noise(obs_grads(theta, idx), g, N, n, C, e)
        p = p + noise_scale * (cb * xi)
      out["idx"].append(idx)
    out["theta"].append(theta.copy())
    out["p"].append(p.copy())
    out["U"].append(U)
  res = {k: np.asarray(v_) for k, v_ in out.items()}
  res["key"], res["dkey"] = key, dkey
  return res


def obabo_integrate(pot, theta, p, key, dkey, eps, temperature, friction, steps,
                    N, n, leaf_sizes, inv_mass, sqrt_mass, ke_start=0.0, ke_end=0.0):
  """obabo.integrate (integrator.py:317-338) over _leapfrog_steps (203-273)."""
  e = F32(eps)
  T = F32(temperature)
  perm = np.exp(F32(-F32(friction) * e)).astype(np.float32)
  a_noise = float(np.sqrt(F32((F32(1) - perm) * T)))
  a_decay = float(np.sqrt(perm))
  half = float(F32(0.5) * e)
  U = 0.0
  for _ in range(steps):
    key, s1, s2 = jr.split(key, 3)
    xi1 = jr.random_tree_flat(s1, leaf_sizes).astype(np.float64)
    p1 = a_decay * p + a_noise * (sqrt_mass * xi1)
    ke_start = ke_start + 0.5 * np.dot(p1, inv_mass * p1)
    dkey, idx0 = draw_indices(dkey, N, n)
    U0, g0, _ = pot(theta, idx0)
    p2 = (-half) * g0 + p1
    theta = theta + float(e) * (inv_mass * p2)
    dkey, idx1 = draw_indices(dkey, N, n)
    U1, g1, _ = pot(theta, idx1)
    p3 = (-half) * g1 + p2
    xi2 = jr.random_tree_flat(s2, leaf_sizes).astype(np.float64)
    p = a_decay * p3 + a_noise * (sqrt_mass * xi2)
    ke_end = ke_end + 0.5 * np.dot(p3, inv_mass * p3)
    U = 0.5 * (U0 + U1)
  return theta, p, key, dkey, U, ke_start, ke_end


def obabo(pot, theta0, iterations, eps, steps, friction, N, n, leaf_sizes,
          temperature=1.0, key=None, dkey=None, inv_mass=None, sqrt_mass=None):
  """alias.obabo through solver.sgmc: p0 = 0 (integrator.py:303), KE carried."""
  key = jr.PRNGKey(0) if key is None else np.asarray(key, np.uint32)
  dkey = jr.PRNGKey(0) if dkey is None else np.asarray(dkey, np.uint32)
  theta = np.asarray(theta0, np.float64).copy()
  P = theta.size
  inv_mass = np.ones(P) if inv_mass is None else np.asarray(inv_mass, np.float64)
  sqrt_mass = np.ones(P) if sqrt_mass is None else np.asarray(sqrt_mass, np.float64)
  eps = np.broadcast_to(np.asarray(eps, np.float32), (iterations,))
  p = np.zeros(P)
  ks, ke = 0.0, 0.0
  out = {"theta": [], "p": [], "U": [], "ke_start": [], "ke_end": []}
  for t in range(iterations):
    theta, p, key, dkey, U, ks, ke = obabo_integrate(
        pot, theta, p, key, dkey, eps[t], temperature, friction, steps, N, n,
        leaf_sizes, inv_mass, sqrt_mass, ks, ke)
    for k, val in (("theta", theta), ("p", p), ("U", U), ("ke_start", ks),
                   ("ke_end", ke)):
      out[k].append(np.copy(val))
  res = {k: np.asarray(v_) for k, v_ in out.items()}
  res["key"], res["dkey"] = key, dkey
  return res


# ---------------------------------------------------------------- solvers ----

def sggmc(pot, full_pot, theta0, iterations, eps, steps, friction, N, n, leaf_sizes,
          temperature=1.0, key=None, dkey=None):
  """solver.sggmc (solver.py:462-567): OBABO proposal, full-data MH test
  log a = min(0, -(U_full' - U_full + KE_end - KE_start) / T), accept iff
  log(u) < log a; on reject theta, p and U revert while keys and data state advance;
  kinetic-energy accumulators reset after every test."""
  key = jr.PRNGKey(0) if key is None else np.asarray(key, np.uint32)
  dkey = jr.PRNGKey(0) if dkey is None else np.asarray(dkey, np.uint32)
  theta = np.asarray(theta0, np.float64).copy()
  P = theta.size
  ones = np.ones(P)
  eps = np.broadcast_to(np.asarray(eps, np.float32), (iterations,))
  U_full = full_pot(theta)
  ikey, key = jr.split(key)
  p = np.zeros(P)
  out = {"theta": [], "accept": [], "log_alpha": [], "U_full": []}
  for t in range(iterations):
    key, sub = jr.split(key)
    th_new, p_new, ikey, dkey, _, ks, ke = obabo_integrate(
        pot, theta, p, ikey, dkey, eps[t], temperature, friction, steps, N, n,
        leaf_sizes, ones, ones)
    U_new = full_pot(th_new)
    log_alpha = -1.0 / float(temperature) * (U_new - U_full + ke - ks)
    # jnp.where(log_alpha <= 0, log_alpha, 0.0) (solver.py:526): NaN becomes 0
    log_alpha = log_alpha if log_alpha <= 0.0 else 0.0
    u = jr.uniform(sub)
    acc = float(np.log(np.float64(u))) < log_alpha
    if acc:
      theta, p, U_full = th_new, p_new, U_new
    for k, val in (("theta", theta), ("accept", acc), ("log_alpha", log_alpha),
                   ("U_full", U_full)):
      out[k].append(np.copy(val))
  return {k: np.asarray(v_) for k, v_ in out.items()}


def resgld_swap_decision(U_normal, U_hot, T_normal, T_hot, ssq, F, key):
  """solver.py:279-291: returns (swap?, new key). The reference keeps the pair when
  log u < log s and swaps otherwise."""
  temps = F32(F32(1.0) / F32(T_normal) - F32(1.0) / F32(T_hot))
  correction = F32(temps * F32(ssq) / F32(F))
  log_s = F32(temps * F32(F32(U_normal) - F32(U_hot) - correction))
  key, sub = jr.split(key)
  log_u = jr.logf_f32(jr.uniform(sub).reshape(1))[0]
  return not (log_u < log_s), key


def resgld(pot, theta_normal, theta_hot, iterations, eps, T_hot, N, n, leaf_sizes,
           key=None, dkey=None, T_normal=1.0, F=1.0, ssq_init=0.0, keep=None):
  """parallel_tempering (solver.py:244-293) over langevin_diffusion as alias.re_sgld
  builds it (alias.py:178-207): init splits key into (swap key, normal key, hot key);
  both chains start their data loader from the same dkey; every iteration updates the
  normal chain (temperature T_normal) and the hot chain (T_hot), refreshes ssq with step
  size 1/step from the normal chain's var(lik), then keeps the pair when log u < log s
  and exchanges the two states otherwise. Returns the final states, the decision of
  each iteration and the normal chain's state after each kept iteration."""
  key = jr.PRNGKey(0) if key is None else np.asarray(key, np.uint32)
  dkey = jr.PRNGKey(0) if dkey is None else np.asarray(dkey, np.uint32)
  ks = jr.split(key, 3)
  skey = ks[0]
  st = [dict(theta=np.asarray(theta_normal, np.float64).copy(), key=ks[1], dkey=dkey.copy(),
             U=None, var=None),
        dict(theta=np.asarray(theta_hot, np.float64).copy(), key=ks[2], dkey=dkey.copy(),
             U=None, var=None)]
  eps = np.broadcast_to(np.asarray(eps, np.float32), (iterations,))
  temps_slot = (F32(T_normal), F32(T_hot))
  ssq = F32(ssq_init)
  swaps, saved, saved_U = [], [], []
  for t in range(iterations):
    for i in (0, 1):
      r = sgld(pot, st[i]["theta"], 1, eps[t:t + 1], temps_slot[i], N, n, leaf_sizes,
               key=st[i]["key"], dkey=st[i]["dkey"])
      st[i] = dict(theta=r["theta"][-1], key=r["key"], dkey=r["dkey"], U=F32(r["U"][-1]),
                   var=F32(r["variance"][-1]))
    step = t + 1
    lr = F32(F32(1.0) / F32(step))
    ssq = F32(F32(F32(1.0) - lr) * ssq + F32(lr * st[0]["var"]))
    sw, skey = resgld_swap_decision(st[0]["U"], st[1]["U"], temps_slot[0], temps_slot[1],
                                    ssq, F, skey)
    if sw:
      st[0], st[1] = st[1], st[0]
    swaps.append(bool(sw))
    if keep is None or keep[t]:
      saved.append(st[0]["theta"].copy())
      saved_U.append(st[0]["U"])
  return {"normal": st[0], "hot": st[1], "swaps": np.asarray(swaps), "ssq": ssq,
          "samples": np.asarray(saved), "samples_U": np.asarray(saved_U), "key": skey}


# --------------------------------------------------------------- AMAGOLD ----

def reversible_leapfrog_init(theta, key, leaf_sizes, sqrt_mass=None):
  """reversible_leapfrog init_fn (integrator.py:465-501): key, split = split(key);
  p = M^{1/2} * random_tree(split, theta); the state keeps ``key``."""
  key, sub = jr.split(np.asarray(key, np.uint32))
  xi = jr.random_tree_flat(sub, leaf_sizes).astype(np.float64)
  p = xi if sqrt_mass is None else np.asarray(sqrt_mass, np.float64) * xi
  return p, key


def reversible_leapfrog_integrate(pot, theta, p, key, dkey, eps, steps, friction, N, n,
                                  leaf_sizes, inv_mass=None, sqrt_mass=None, grad_fn=None):
  """reversible_leapfrog.integrate (integrator.py:503-553) over _body_fun (:397-453):
  half position step; ``steps`` x {position step (not at step 0), split key, noise
  sqrt(4 f eps) M^{1/2} xi, minibatch gradient g, p' = ((1 - eps f) p + (-eps) g +
  noise) / (1 + eps f), energy += 0.5 eps (p + p') . (M^-1 g)}; final half position
  step. Coefficients are formed in float32 as the reference forms them.
  ``grad_fn(theta) -> g`` replaces the minibatch gradient (no index draw) when given.
  Returns theta, p, key, dkey, energy and the last minibatch potential."""
  P = theta.size
  inv_mass = np.ones(P) if inv_mass is None else np.asarray(inv_mass, np.float64)
  sqrt_mass = np.ones(P) if sqrt_mass is None else np.asarray(sqrt_mass, np.float64)
  e = F32(eps)
  fr = F32(friction)
  half = float(F32(F32(0.5) * e))
  noise_scale = float(np.sqrt(F32(F32(4.0 * friction) * e)))
  decay = float(F32(F32(1.0) - F32(e * fr)))
  rescale = float(F32(F32(1.0) / F32(F32(1.0) + F32(e * fr))))
  theta = theta + half * (inv_mass * p)
  energy = 0.0
  U = 0.0
  for step in range(steps):
    if step > 0:
      theta = theta + float(e) * (inv_mass * p)
    key, sub = jr.split(key)
    xi = jr.random_tree_flat(sub, leaf_sizes).astype(np.float64)
    noise = noise_scale * (sqrt_mass * xi)
    if grad_fn is None:
      dkey, idx = draw_indices(dkey, N, n)
      U, g, _ = pot(theta, idx)
    else:
      g = grad_fn(theta)
    p_new = rescale * ((decay * p + float(-e) * g) + noise)
    energy = half * np.dot(p + p_new, inv_mass * g) + energy
    p = p_new
  theta = theta + half * (inv_mass * p)
  return theta, p, key, dkey, energy, U


def amagold(pot, full_pot, theta0, iterations, eps, steps, friction, N, n, leaf_sizes,
            key=None, dkey=None, inv_mass=None, sqrt_mass=None, mass_burn_in=None):
  """solver.amagold (solver.py:301-432) over reversible_leapfrog: init splits the
  solver key (integrator key, MH key); every iteration splits the MH key, integrates,
  evaluates the full-data potential at the proposal and accepts iff
  log u < min(0, U_full - U_full' + energy). A rejected proposal keeps theta and U_full
  and flips the momentum; keys and the data state advance either way. With
  ``mass_burn_in`` the diagonal mass adapts (adaption.mass_matrix with that burn_in,
  updated with every post-test position): from then on M^-1 is the biased variance of
  the first mass_burn_in post-test positions and M^{1/2} its inverse square root."""
  key = jr.PRNGKey(0) if key is None else np.asarray(key, np.uint32)
  dkey = jr.PRNGKey(0) if dkey is None else np.asarray(dkey, np.uint32)
  theta = np.asarray(theta0, np.float64).copy()
  eps = np.broadcast_to(np.asarray(eps, np.float32), (iterations,))
  U_full = full_pot(theta)
  kept = []
  ikey, key = jr.split(key)
  p, ikey = reversible_leapfrog_init(theta, ikey, leaf_sizes, sqrt_mass)
  out = {"theta": [], "p": [], "accept": [], "log_alpha": [], "U_full": [], "energy": []}
  for t in range(iterations):
    key, sub = jr.split(key)
    th_new, p_new, ikey, dkey, energy, _ = reversible_leapfrog_integrate(
        pot, theta, p, ikey, dkey, eps[t], steps, friction, N, n, leaf_sizes, inv_mass,
        sqrt_mass)
    U_new = full_pot(th_new)
    log_alpha = min(U_full - U_new + energy, 0.0)
    u = jr.uniform(sub)
    acc = float(np.log(np.float64(u))) < log_alpha
    if acc:
      theta, p, U_full = th_new, p_new, U_new
    else:
      p = -p
    if mass_burn_in is not None:
      kept.append(theta.copy())
      if len(kept) == mass_burn_in:
        var = np.var(np.asarray(kept), axis=0)
        inv_mass, sqrt_mass = var, 1.0 / np.sqrt(var)
    for k_, val in (("theta", theta), ("p", p), ("accept", acc), ("log_alpha", log_alpha),
                    ("U_full", U_full), ("energy", energy)):
      out[k_].append(np.copy(val))
  res = {k_: np.asarray(v_) for k_, v_ in out.items()}
  res["key"], res["ikey"], res["dkey"] = key, ikey, dkey
  return res


def adaptive_step_size(acceptance, burn_in, initial_step_size, stabilization_constant,
                       decay_constant, speed_constant, target_acceptance_rate):
  """scheduler.adaptive_step_size (scheduler.py:376-444), dual averaging in float32:
  m = i + 1; h = (1 - 1/(m + t0)) h + 1/(m + t0) (target - a_i);
  x = mu - sqrt(m)/gamma h; x_bar = (1 - m^-kappa) x_bar + m^-kappa x while i < burn_in;
  mu = log(10 eps0). Returns the step size used at every iteration (exp(x_bar) before
  the update of that iteration)."""
  x_bar = np.log(F32(initial_step_size))
  h_bar = F32(0.0)
  mu = np.log(F32(10.0 * initial_step_size))
  t0 = stabilization_constant
  out = []
  for i, a in enumerate(np.asarray(acceptance, np.float32)):
    out.append(np.exp(F32(x_bar)))
    m = i + 1
    w = F32(1.0) / F32(m + t0)
    h_bar = F32(h_bar * F32(F32(1.0) - w))
    h_bar = F32(h_bar + w * F32(F32(target_acceptance_rate) - a))
    x = F32(mu - F32(np.sqrt(F32(m)) / F32(speed_constant)) * h_bar)
    lr = np.power(F32(m), F32(-decay_constant))
    xb = F32(F32(x_bar * F32(F32(1.0) - lr)) + F32(lr * x))
    if i < burn_in:
      x_bar = xb
  return np.asarray(out, np.float32)


def mass_adaption(samples, burn_in, init_cov=None):
  """adaption.mass_matrix, diagonal (adaption.py:296-369), over a sequence of samples:
  returns (M^-1, M^{1/2}) after each update. Until burn_in updates: init_cov and
  1 / sqrt(init_cov); from then on the biased sample variance of the first burn_in
  samples and its inverse square root."""
  samples = np.asarray(samples, np.float64)
  cov = np.ones(samples.shape[1]) if init_cov is None else np.asarray(init_cov, np.float64)
  out = []
  inv, sq = cov, 1.0 / np.sqrt(cov)
  for t in range(len(samples)):
    if t + 1 == burn_in:
      var = np.var(samples[:burn_in], axis=0)
      inv, sq = var, 1.0 / np.sqrt(var)
    out.append((inv.copy(), sq.copy()))
  return out
