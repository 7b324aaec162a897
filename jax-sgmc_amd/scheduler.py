"""Step-size, temperature, burn-in and thinning schedules (host side).

Mirrors jax_sgmc.scheduler (scheduler.py:93-666): each specific scheduler is an
(init, update, get) triplet; init_scheduler combines them into the (init_fn,
update_fn, get_fn) the solver drives once per iteration, and ``get`` returns a
``schedule(step_size, temperature, burn_in, accept)``. The schedules are static
tables computed once in float32 (the reference precomputes them too,
scheduler.py:469-481); the fused path uploads the step-size table to the device.
"""
from collections import namedtuple

import numpy as np

specific_scheduler = namedtuple("specific_scheduler", ["init", "update", "get"])
schedule = namedtuple("schedule", ["step_size", "temperature", "burn_in", "accept"])
scheduler_state = namedtuple("scheduler_state", ["state", "step_size_state",
                                                 "temperature_state", "burn_in_state",
                                                 "thinning_state"])
static_information = namedtuple("static_information", ["samples_collected"])

F32 = np.float32


def _static_update(state, iteration, **kw):
  """Update of a schedule fixed at init (tables, constants)."""
  return state


def polynomial_table(first, last, iterations, gamma=0.33):
  """eps_i = a (b + i)^(-gamma) with eps_0 = first and eps_{I-1} = last, float32
  (scheduler.py:512-532)."""
  assert gamma > 0, f"Gamma must be bigger than 0, is {gamma}"
  assert first >= last, f"The first step size must be larger than the last: {first} !>= {last}"
  g = F32(gamma)
  first = F32(first)
  last = F32(last)
  with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
    ginv = np.power(g, F32(-1.0))
    fpow = np.power(first, -ginv)
    lpow = np.power(last, -ginv)
    apow = F32((lpow - fpow) / F32(iterations - 1))
    a = np.power(apow, -g)
    b = np.power(F32(first / a), -ginv)
    n = np.arange(iterations).astype(np.float32)
    return (a * np.power(b + n, -g)).astype(np.float32)


def polynomial_step_size(a=1.0, b=1.0, gamma=0.33):
  """scheduler.py:447-491: eps_i = a (b + i)^(-gamma)."""

  def init_fn(iterations, a=a, b=b, gamma=gamma):
    assert gamma >= 0, f"Gamma must be positive: gamma = {gamma}"
    assert a > 0, f"a must be positive: a = {a}"
    assert b > 0, f"b must be greater than zero: b = {b}"
    n = np.arange(iterations).astype(np.float32)
    return (F32(a) * np.power(F32(b) + n, F32(-gamma))).astype(np.float32)

  return specific_scheduler(init_fn, _static_update, lambda s, i, **kw: s[i])


def polynomial_step_size_first_last(first=1.0, last=1.0, gamma=0.33):
  """scheduler.py:494-542."""

  def init_fn(iterations, first=first, last=last, gamma=gamma):
    return polynomial_table(first, last, iterations, gamma)

  return specific_scheduler(init_fn, _static_update, lambda s, i, **kw: s[i])


def adaptive_step_size(burn_in=0, initial_step_size=0.05, stabilization_constant=100,
                       decay_constant=0.75, speed_constant=0.05, target_acceptance_rate=0.02):
  """scheduler.py:376-444: dual averaging of log(eps) towards a target acceptance rate
  during burn-in, in float32. The state is (burn_in, x_bar, h_bar, target, t0, kappa,
  gamma, mu) with mu = log(10 eps0); update(state, i, acceptance_ratio) with m = i + 1:
  h = (1 - 1/(m + t0)) h + 1/(m + t0) (target - a); x = mu - sqrt(m)/gamma h;
  x_bar = (1 - m^-kappa) x_bar + m^-kappa x while i < burn_in; get = exp(x_bar). As in
  the reference, the update uses the target given to this constructor."""

  def init_fn(iterations, burn_in=burn_in, initial_step_size=initial_step_size,
              stabilization_constant=stabilization_constant, decay_constant=decay_constant,
              speed_constant=speed_constant, target_acceptance_rate=target_acceptance_rate):
    del iterations
    return (burn_in, np.log(F32(initial_step_size)), F32(0.0), target_acceptance_rate,
            stabilization_constant, decay_constant, speed_constant,
            np.log(F32(10 * initial_step_size)))

  def update_fn(state, iteration, acceptance_ratio=0.0, **kw):
    del kw
    burn, x_bar, h_bar, alpha, t0, kappa, gamma, mu = state
    m = iteration + 1
    w = F32(1.0) / F32(m + t0)
    h_bar = F32(h_bar * F32(F32(1.0) - w))
    a = F32(float(acceptance_ratio))
    h_bar = F32(h_bar + w * F32(F32(target_acceptance_rate) - a))
    x = F32(mu - F32(np.sqrt(F32(m)) / F32(gamma)) * h_bar)
    lr = np.power(F32(m), F32(-kappa))
    new_x_bar = F32(F32(x_bar * F32(F32(1.0) - lr)) + F32(lr * x))
    x_bar = new_x_bar if iteration < burn else x_bar
    return burn, x_bar, h_bar, alpha, t0, kappa, gamma, mu

  def get_fn(state, iteration, **kw):
    return np.exp(F32(state[1]))

  return specific_scheduler(init_fn, update_fn, get_fn)


def constant_temperature(tau=1.0):
  """scheduler.py:245-276."""
  return specific_scheduler(lambda iterations, tau=tau: F32(tau), lambda s, i, **kw: s,
                            lambda s, i, **kw: s)


def initial_burn_in(n=0):
  """scheduler.py:567-589: 1.0 from iteration n on."""
  return specific_scheduler(lambda iterations, n=n: (n, iterations - n),
                            lambda s, i, **kw: s,
                            lambda s, i, **kw: F32(1.0) if s <= i else F32(0.0))


def gumbel_topk(uniforms, probs, k):
  """The draw of jax.random.choice(key, arange(n), (k,), replace=False, p=probs) given
  u = uniform(key, (n,), minval=tiny): the Gumbel top-k construction, the k smallest
  entries of -gumbel - log p with gumbel = -log(-log u), in float32 (zero-probability
  entries rank last). Returned in increasing order (the scheduler only tests
  membership)."""
  u = np.asarray(uniforms, np.float32)
  with np.errstate(divide="ignore"):
    gumbel = -np.log(-np.log(u))
    g = -gumbel - np.log(np.asarray(probs, np.float32))
  return np.sort(np.argsort(g, kind="stable")[:k])


def random_thinning(step_size_schedule, burn_in_schedule, selections, key=None):
  """scheduler.py:599-666: ``selections`` iterations drawn without replacement with
  probability proportional to step_size * burn_in, by the Gumbel top-k construction
  jax.random.choice(replace=False, p=...) uses: argsort(-gumbel - log p)[:k], with
  gumbel = -log(-log(uniform(key, minval=tiny))). Uniforms come from the device
  Threefry generator with key PRNGKey(0) by default."""

  def init_fn(iterations, step_size_schedule=step_size_schedule,
              burn_in_schedule=burn_in_schedule, selections=selections, key=key):
    from . import random as R
    steps = step_size_schedule.init(iterations)
    bstate, _ = burn_in_schedule.init(iterations)
    probs = np.array([steps[i] * burn_in_schedule.get(bstate, i) for i in range(iterations)],
                     np.float32)
    assert np.count_nonzero(probs) >= selections, "Cannot select enough values"
    k = R.PRNGKey(0) if key is None else key
    u = R.uniform(k, (iterations,), np.finfo(np.float32).tiny, 1.0).cpu().numpy()
    return gumbel_topk(u, probs, selections), selections

  def get_fn(state, iteration, **kw):
    return bool(np.any(state == iteration))

  return specific_scheduler(init_fn, lambda s, i, **kw: s, get_fn)


def init_scheduler(step_size=None, temperature=None, burn_in=None, thinning=None,
                   progress_bar=False, progress_bar_steps=20):
  """scheduler.py:93-237 (the progress bar is not reproduced)."""
  del progress_bar, progress_bar_steps
  if step_size is None:
    step_size = polynomial_step_size(a=1.0, b=1.0, gamma=0.0)
  if temperature is None:
    temperature = constant_temperature(1.0)
  if burn_in is None:
    burn_in = initial_burn_in(0)
  if thinning is None:
    thinning = specific_scheduler(lambda iterations: (None, iterations),
                                  lambda *a, **k: None, lambda *a, **k: True)

  def init_fn(iterations, **kw):
    th_state, total = thinning.init(iterations, **kw.get("thinning", {}))
    bi_state, collected = burn_in.init(iterations, **kw.get("burn_in", {}))
    total = min(total, collected)
    st = scheduler_state(state=(0, iterations),
                         step_size_state=step_size.init(iterations, **kw.get("step_size", {})),
                         temperature_state=temperature.init(iterations,
                                                            **kw.get("temperature", {})),
                         burn_in_state=bi_state, thinning_state=th_state)
    return st, static_information(samples_collected=total)

  def update_fn(state, **kw):
    it, total = state.state
    return state._replace(
        state=(it + 1, total),
        step_size_state=step_size.update(state.step_size_state, it, **kw),
        temperature_state=temperature.update(state.temperature_state, it, **kw),
        burn_in_state=burn_in.update(state.burn_in_state, it, **kw),
        thinning_state=thinning.update(state.thinning_state, it, **kw))

  def get_fn(state, **kw):
    it, _ = state.state
    return schedule(step_size=F32(step_size.get(state.step_size_state, it)),
                    temperature=F32(temperature.get(state.temperature_state, it)),
                    burn_in=F32(burn_in.get(state.burn_in_state, it)),
                    accept=bool(thinning.get(state.thinning_state, it)))

  def tables(state):
    """Whole-run tables (step size, temperature, keep flags) for device-driven loops."""
    if step_size.update.__qualname__.startswith("adaptive_step_size."):
      raise NotImplementedError("a device-driven loop needs a step-size table fixed in "
                                "advance; this step-size schedule adapts while running")
    _, total = state.state
    eps = np.array([step_size.get(state.step_size_state, i) for i in range(total)], np.float32)
    temp = np.array([temperature.get(state.temperature_state, i) for i in range(total)],
                    np.float32)
    keep = np.array([bool(burn_in.get(state.burn_in_state, i)) and
                     bool(thinning.get(state.thinning_state, i)) for i in range(total)])
    return eps, temp, keep

  init_fn.tables = tables
  return init_fn, update_fn, get_fn
