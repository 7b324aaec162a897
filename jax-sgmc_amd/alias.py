"""Pre-built samplers with the reference's signatures (alias.py:30-619).

Each returns ``run_fn(*init_samples, init_model_state=None, iterations=1000)`` whose
result is a list (one entry per chain) of ``{"sample_count": k, "samples":
{"variables": ..., "likelihood" | "energy": ..., "model_state": ...}}``.

Defaults reproduce the reference: every chain starts from the integrator key
PRNGKey(0) and the device data key PRNGKey(0) (integrator.py:804,
data/numpy_loader.py:124), so chains started from the same sample coincide. As an
extension, ``run_fn`` accepts ``keys=`` and ``data_keys=`` (one jax key per chain)
and ``fused=`` (None: automatic).

Execution of an MLP classifier (likelihood ``mlp.MLPSoftmax`` with a Gaussian or no
prior, over a DeviceNumpyDataLoader, identity mass): alias.sgld, obabo and sggmc run
every chain on the MLP engine (``mlp.FusedMLP``: hand-written forward / backward GEMMs on
the matrix cores with the update fused in) unless ``fused=False``; ``precision=``
selects its GEMM operands: "fp32" (default), "bf16" or "fp8". Parameters, momenta,
noise and updates are float32 in every precision. ``fused=False`` runs the generic torch
path (autograd over the likelihood callable).

Execution: when the potential is a recognized GLM family with a Gaussian (or zero)
prior, the data loader is device-resident, and the integrator is SGLD (with or
without rms_prop), SGHMC or OBABO with unit mass and scalar friction, all chains run
together in the fused multi-chain kernel (``fused.FusedGLM``); AMAGOLD and SGGMC do
too when the full potential is the same family, followed by one full-data pass for
all chains and the Metropolis-Hastings kernel. The result is the same chain by chain.
Otherwise the chains run one after another through the modular kernels (solver.mcmc
with strategy 'map', as in the reference).
"""
from functools import partial

import numpy as np
import torch

from . import adaption as AD
from . import data as DT
from . import integrator as IG
from . import io as IO
from . import likelihoods as LK
from . import potential as PT
from . import random as R
from . import scheduler as SC
from . import solver as SV
from . import tree as T


def _schedule(first_step_size, last_step_size, burn_in, accepted_samples, temperature=None):
  step = SC.polynomial_step_size_first_last(first=first_step_size, last=last_step_size)
  burn = SC.initial_burn_in(burn_in)
  thin = SC.random_thinning(step, burn, selections=accepted_samples) \
      if accepted_samples is not None else None
  return SC.init_scheduler(step_size=step, burn_in=burn, thinning=thin,
                           temperature=temperature)


def _mh_schedule(adaptive, first_step_size, last_step_size, burn_in, accepted_samples,
                 stabilization_constant, decay_constant, speed_constant,
                 target_acceptance_rate):
  """Schedules of the MH solvers (alias.py:287-311, 410-434): dual-averaging step size
  during burn-in when ``adaptive`` (no thinning then), the polynomial first/last
  schedule otherwise."""
  if not adaptive:
    return _schedule(first_step_size, last_step_size, burn_in, accepted_samples)
  assert accepted_samples is None, ("Thinning currently not supported for"
                                    " adaptive step size.")
  step = SC.adaptive_step_size(burn_in=burn_in, initial_step_size=first_step_size,
                               stabilization_constant=stabilization_constant,
                               decay_constant=decay_constant, speed_constant=speed_constant,
                               target_acceptance_rate=target_acceptance_rate)
  return SC.init_scheduler(step_size=step, burn_in=SC.initial_burn_in(burn_in))


def _fused_eligible(potential_fn, data_loader, init_samples, fused):
  if fused is False:
    return False
  ok = (isinstance(potential_fn, PT.StochasticPotential) and potential_fn.fast
        and isinstance(data_loader, DT.DeviceNumpyDataLoader))
  if ok:
    for s in init_samples:
      leaves, _ = T.flatten(s)
      if len(leaves) != 1 or np.ndim(np.asarray(leaves[0].cpu() if isinstance(
          leaves[0], torch.Tensor) else leaves[0])) != 1:
        ok = False
    X = data_loader.reference_data.get(potential_fn.likelihood.x)
    if X is None or X.dim() != 2 or X.shape[1] > 1024 or X.stride(0) % 4:
      ok = False
  if fused and not ok:
    raise ValueError("fused=True needs a GLM family potential, a DeviceNumpyDataLoader and "
                     "single-leaf 1-D samples")
  return ok


# chains from which alias runs of multi-block chains use two tables on two streams
_TABLE_SPLIT_CHAINS = 16


def _key_words(k):
  if k is None:
    return (0, 0)
  if isinstance(k, torch.Tensor):
    return tuple(int(v) for v in R.key_words(k))
  return tuple(int(v) for v in np.asarray(k, np.uint32))


def _run_fused(potential_fn, data_loader, batch_size, init_samples, iterations, sched, mode,
               keys, data_keys, steps=1, friction=1.0, alpha=0.9, lmbd=1e-5, to_numpy=True,
               result_key="likelihood"):
  init_fn = sched[0]
  st, _ = init_fn(iterations)
  eps, temp, keep = init_fn.tables(st)
  slots = np.full(iterations, -1, np.int32)
  kept = np.nonzero(keep)[0]
  slots[kept] = np.arange(kept.size, dtype=np.int32)
  lik = potential_fn.likelihood
  X = data_loader.reference_data[lik.x]
  y = data_loader.reference_data[lik.y].to(torch.float32).contiguous()
  has_prior, pscale = PT._prior_spec(potential_fn.prior)
  chains = []
  for c, s in enumerate(init_samples):
    leaf = T.flatten(s)[0][0]
    th0 = leaf.detach().cpu().numpy() if isinstance(leaf, torch.Tensor) else np.asarray(leaf)
    chains.append(dict(X=X, y=y, D=X.shape[1], n=batch_size, theta0=th0.astype(np.float32),
                       key=_key_words(keys[c] if keys else None),
                       dkey=_key_words(data_keys[c] if data_keys else None),
                       temperature=float(temp[0]) if temp.size else 1.0))
  from . import fused as FU
  kw = dict(steps=steps, friction=friction, alpha=alpha, lmbd=lmbd, sigma=lik.sigma,
            prior_scale=pscale, has_prior=bool(has_prior),
            pot_temperature=potential_fn.temperature, save_slots=slots,
            n_slots=max(1, kept.size))
  total = sum(int(c["n"]) * 4 * FU._align4(int(c["D"])) for c in chains)
  tb = max(64 * 1024, total // 2048)
  multi = any(FU.blocks_for(int(c["n"]), int(c["D"]), tb) > 1 for c in chains)
  if len(chains) >= _TABLE_SPLIT_CHAINS and multi:
    # many multi-block chains: two tables on two streams, so that one table's launch
    # tail overlaps the other's row gather (same results chain by chain)
    eng = FU.TableGroup.split(chains, 2, lik.family, mode, eps, **kw)
  else:
    eng = FU.FusedGLM(chains, lik.family, mode, eps, **kw)
  eng.run(iterations)
  torch.cuda.synchronize()
  out = []
  for c, s in enumerate(init_samples):
    tr = eng.sample_trace(c)[:kept.size]
    Us = eng.sample_potentials(c)[:kept.size]
    _, td = T.flatten(s)
    var = T.unflatten(td, [tr.clone()])
    val = -Us if result_key == "likelihood" else Us
    samples = {"variables": var, result_key: val.clone(), "model_state": None}
    if to_numpy:
      samples = T.tree_map(lambda x: x.cpu().numpy() if isinstance(x, torch.Tensor) else x,
                           samples)
    out.append({"sample_count": int(kept.size), "samples": samples})
  return out


def _run_fused_resgld(potential_fn, data_loader, batch_size, init_samples, iterations, sched,
                      tau, keys, data_keys, to_numpy=True):
  """alias.re_sgld on the fused engine: chain 2c runs the normal schedule of pair c,
  chain 2c + 1 the hot one (temperature tau); after every iteration the swap kernel
  tests each pair and saves the normal slot (solver.py:244-293)."""
  init_fn = sched[0]
  st, _ = init_fn(iterations)
  eps, temp, keep = init_fn.tables(st)
  slots = np.full(iterations, -1, np.int32)
  kept = np.nonzero(keep)[0]
  slots[kept] = np.arange(kept.size, dtype=np.int32)
  lik = potential_fn.likelihood
  X = data_loader.reference_data[lik.x]
  y = data_loader.reference_data[lik.y].to(torch.float32).contiguous()
  has_prior, pscale = PT._prior_spec(potential_fn.prior)
  t_normal = float(temp[0]) if temp.size else 1.0
  dev = X.device
  chains, pairs, temps, swap_keys = [], [], [], []
  for c, (normal, hot) in enumerate(init_samples):
    key = keys[c] if keys else None
    key = R.key_from_words(_key_words(key), device=dev)
    ks = R.key_words(R.split(key, 3))  # (swap key, normal key, hot key), solver.py:254
    dk = _key_words(data_keys[c] if data_keys else None)
    for s, k, t in ((normal, ks[1], t_normal), (hot, ks[2], float(tau))):
      leaf = T.flatten(s)[0][0]
      th0 = leaf.detach().cpu().numpy() if isinstance(leaf, torch.Tensor) else np.asarray(leaf)
      chains.append(dict(X=X, y=y, D=X.shape[1], n=batch_size,
                         theta0=np.asarray(th0, np.float32).reshape(-1), key=tuple(k),
                         dkey=dk, temperature=t, save=False))
    pairs.append((2 * c, 2 * c + 1))
    temps.append((t_normal, float(tau)))
    swap_keys.append(tuple(ks[0]))
  from . import fused as FU
  eng = FU.FusedGLM(chains, lik.family, "sgld", eps, sigma=lik.sigma, prior_scale=pscale,
                    has_prior=bool(has_prior), pot_temperature=potential_fn.temperature,
                    save_slots=slots, n_slots=max(1, kept.size))
  ex = FU.ReplicaExchange(eng, pairs, temps, keys=swap_keys, save=[True] * len(pairs))
  eng.run(iterations)
  torch.cuda.synchronize()
  out = []
  for c, (normal, _) in enumerate(init_samples):
    tr = ex.sample_trace(c)[:kept.size]
    Us = ex.sample_potentials(c)[:kept.size]
    _, td = T.flatten(normal)
    samples = {"variables": T.unflatten(td, [tr.clone()]), "likelihood": (-Us).clone(),
               "model_state": None}
    if to_numpy:
      samples = T.tree_map(lambda x: x.cpu().numpy() if isinstance(x, torch.Tensor) else x,
                           samples)
    out.append({"sample_count": int(kept.size), "samples": samples})
  return out


def _mh_fused_eligible(stochastic_potential_fn, full_potential_fn, data_loader,
                       init_samples, mass, adaptive, fused):
  """The fused AMAGOLD and SGGMC engines apply when the stochastic and the full potential are
  the same GLM family, prior and temperature over a DeviceNumpyDataLoader, the mass is
  the identity and the step-size schedule is fixed in advance."""
  if fused is False:
    return False
  sp, fp = stochastic_potential_fn, full_potential_fn
  ok = (not adaptive and mass is None and isinstance(fp, PT.FullPotential) and fp.fast
        and _fused_eligible(sp, data_loader, init_samples, None)
        and type(fp.likelihood) is type(sp.likelihood)
        and getattr(fp.likelihood, "sigma", None) == getattr(sp.likelihood, "sigma", None)
        and fp.likelihood.x == sp.likelihood.x and fp.likelihood.y == sp.likelihood.y
        and PT._prior_spec(fp.prior) == PT._prior_spec(sp.prior)
        and fp.temperature == sp.temperature)
  if fused and not ok:
    raise ValueError("fused=True needs matching GLM stochastic and full potentials over a "
                     "DeviceNumpyDataLoader, identity mass and a fixed step-size schedule")
  return ok


def _run_fused_mh(potential_fn, data_loader, batch_size, init_samples, iterations, sched,
                  mode, steps, friction, keys, data_keys, to_numpy=True):
  """alias.amagold / alias.sggmc on the fused engine: every chain's proposal
  (reversible leapfrog in ``steps`` launches, or OBABO in 2 ``steps``) runs for all
  chains at once, then one pass over the data gives every proposal's full potential
  and the MH kernel accepts or rejects each chain (solver.py:301-432, 434-567)."""
  init_fn = sched[0]
  st, _ = init_fn(iterations)
  eps, temp, keep = init_fn.tables(st)
  slots = np.full(iterations, -1, np.int32)
  kept = np.nonzero(keep)[0]
  slots[kept] = np.arange(kept.size, dtype=np.int32)
  lik = potential_fn.likelihood
  X = data_loader.reference_data[lik.x]
  y = data_loader.reference_data[lik.y].to(torch.float32).contiguous()
  has_prior, pscale = PT._prior_spec(potential_fn.prior)
  chains = []
  for c, s in enumerate(init_samples):
    leaf = T.flatten(s)[0][0]
    th0 = leaf.detach().cpu().numpy() if isinstance(leaf, torch.Tensor) else np.asarray(leaf)
    chains.append(dict(X=X, y=y, D=X.shape[1], n=batch_size,
                       theta0=np.asarray(th0, np.float32).reshape(-1),
                       dkey=_key_words(data_keys[c] if data_keys else None),
                       temperature=float(temp[0]) if temp.size else 1.0))
  from . import fused as FU
  eng = FU.FusedGLM(chains, lik.family, mode, eps, steps=steps, friction=friction,
                    sigma=lik.sigma, prior_scale=pscale, has_prior=bool(has_prior),
                    pot_temperature=potential_fn.temperature, save_slots=slots,
                    n_slots=max(1, kept.size))
  test = FU.AmagoldMH if mode == "amagold" else FU.SggmcMH
  mh = test(eng, keys=[_key_words(keys[c]) for c in range(len(chains))] if keys else None)
  eng.run(iterations)
  torch.cuda.synchronize()
  out = []
  for c, s in enumerate(init_samples):
    _, td = T.flatten(s)
    samples = {"variables": T.unflatten(td, [eng.sample_trace(c)[:kept.size].clone()]),
               "energy": eng.sample_potentials(c)[:kept.size].clone(),
               "acceptance_ratio": mh.acceptance_ratios(c)[:kept.size].clone(),
               "step_size": torch.as_tensor(eps[kept]), "model_state": None}
    if mode == "sggmc":
      samples["kinetic_energy"] = mh.kinetic_energies(c)[:kept.size].clone()
      samples["potential"] = mh.full_potentials(c)[:kept.size].clone()
    if to_numpy:
      samples = T.tree_map(lambda x: x.cpu().numpy() if isinstance(x, torch.Tensor) else x,
                           samples)
    out.append({"sample_count": int(kept.size), "samples": samples})
  return out


_MLP_LEAVES = ("W1", "W2", "W3", "b1", "b2", "b3")


def _mlp_eligible(potential_fn, data_loader, init_samples, batch_size, mass=None):
  """The MLP engine applies to an MLPSoftmax likelihood with a Gaussian (or no) prior
  over a DeviceNumpyDataLoader, identity mass, samples {"W1".."b3"} with D_in and n
  multiples of 64 and H1, H2 multiples of 128."""
  from . import mlp as ML
  lik = getattr(potential_fn, "likelihood", None)
  if not (isinstance(potential_fn, PT.StochasticPotential) and isinstance(lik, ML.MLPSoftmax)
          and isinstance(data_loader, DT.DeviceNumpyDataLoader) and mass is None
          and PT._prior_spec(potential_fn.prior) is not None):
    return False
  X = data_loader.reference_data.get(lik.x)
  if X is None or X.dim() != 2 or X.dtype != torch.float32:
    return False
  for smp in init_samples:
    if not isinstance(smp, dict) or sorted(smp) != list(_MLP_LEAVES):
      return False
    D_in, H1 = tuple(smp["W1"].shape)
    H2 = smp["W2"].shape[1]
    if (D_in != X.shape[1] or D_in % 64 or H1 % 128 or H2 % 128 or batch_size % 64 or
        tuple(smp["W3"].shape) != (H2, ML.NCLS)):
      return False
  return True


def _run_fused_mlp(potential_fn, data_loader, batch_size, init_samples, iterations, sched,
                   integrator, keys, data_keys, to_numpy=True, rms=False, alpha=0.9, lmbd=1e-5,
                   steps=1, friction=1.0, result_key="likelihood", precision="fp32"):
  """alias.sgld / obabo / sggmc on the MLP engine (fp32, bf16 or fp8 GEMM operands,
  float32 state): every chain advances in the same launches."""
  from . import mlp as ML
  init_fn = sched[0]
  st, _ = init_fn(iterations)
  eps, temp, keep = init_fn.tables(st)
  slots = np.full(iterations, -1, np.int32)
  kept = np.nonzero(keep)[0]
  slots[kept] = np.arange(kept.size, dtype=np.int32)
  lik = potential_fn.likelihood
  X = data_loader.reference_data[lik.x]
  labels = data_loader.reference_data[lik.y]
  has_prior, pscale = PT._prior_spec(potential_fn.prior)
  first = init_samples[0]
  dims = (int(first["W1"].shape[0]), int(first["W1"].shape[1]), int(first["W2"].shape[1]))
  th0 = np.stack([np.concatenate([np.asarray(smp[k].detach().cpu() if isinstance(smp[k], torch.Tensor)
                                             else smp[k], np.float32).reshape(-1)
                                  for k in _MLP_LEAVES]) for smp in init_samples])
  C = len(init_samples)
  kw = [_key_words(keys[c] if keys else None) for c in range(C)]
  dk = [_key_words(data_keys[c] if data_keys else None) for c in range(C)]
  eng = ML.FusedMLP(X, labels, dims, batch_size, th0, eps, keys=kw, dkeys=dk, rms=rms,
                    alpha=alpha, lmbd=lmbd, prior_scale=pscale, has_prior=bool(has_prior),
                    temperature=potential_fn.temperature,
                    noise_temperature=float(temp[0]) if temp.size else 1.0,
                    save_slots=slots if kept.size else None, integrator=integrator,
                    friction=friction, steps=steps, precision=precision)
  eng.run(iterations)
  torch.cuda.synchronize()
  shapes = ML.leaf_shapes(*dims)
  out = []
  for c in range(C):
    tr = eng.samples[c, :kept.size] if eng.samples is not None else eng.theta[c:c + 1, :0]
    var, off = {}, 0
    for k in _MLP_LEAVES:
      size = int(np.prod(shapes[k]))
      var[k] = tr[:, off:off + size].reshape((kept.size,) + shapes[k]).clone()
      off += size
    Us = eng.samples_U[c, :kept.size] if eng.samples_U is not None else eng.U[:0]
    val = -Us if result_key == "likelihood" else Us
    samples = {"variables": var, result_key: val.clone(), "model_state": None}
    if to_numpy:
      samples = T.tree_map(lambda x: x.cpu().numpy() if isinstance(x, torch.Tensor) else x,
                           samples)
    out.append({"sample_count": int(kept.size), "samples": samples})
  return out


def sgld(potential_fn, data_loader, cache_size=512, batch_size=32, first_step_size=0.05,
         last_step_size=0.001, burn_in=0, accepted_samples=1000, rms_prop=False, alpha=0.9,
         lmbd=1e-5, save_to_numpy=True, progress_bar=True):
  """alias.py:30-120."""
  random_data = DT.random_reference_data(data_loader, cache_size, batch_size)
  adapt = AD.rms_prop() if rms_prop else None
  integ = IG.langevin_diffusion(potential_fn, random_data, adaption=adapt)
  sched = _schedule(first_step_size, last_step_size, burn_in, accepted_samples)
  saving = IO.save(to_numpy=True) if save_to_numpy else None
  solver = SV.sgmc(integ)
  mcmc = SV.mcmc(solver, sched, strategy="map", saving=saving)

  def run_fn(*init_samples, init_model_state=None, iterations=1000, keys=None,
             data_keys=None, fused=None, precision="fp32"):
    if (fused is not False and init_model_state is None and
        _mlp_eligible(potential_fn, data_loader, init_samples, batch_size)):
      return _run_fused_mlp(potential_fn, data_loader, batch_size, init_samples, iterations,
                            sched, "sgld", keys, data_keys, to_numpy=save_to_numpy,
                            rms=rms_prop, alpha=alpha, lmbd=lmbd, precision=precision)
    if init_model_state is None and _fused_eligible(potential_fn, data_loader, init_samples,
                                                    fused):
      return _run_fused(potential_fn, data_loader, batch_size, init_samples, iterations,
                        sched, "sgld_rms" if rms_prop else "sgld", keys, data_keys,
                        alpha=alpha, lmbd=lmbd, to_numpy=save_to_numpy)
    states = []
    for c, s in enumerate(init_samples):
      kw = dict(adaption_kwargs={"alpha": alpha, "lmbd": lmbd}, init_model_state=init_model_state)
      if keys:
        kw["key"] = keys[c]
      if data_keys:
        kw["batch_kwargs"] = {"key": data_keys[c]}
      states.append(solver[0](s, **kw))
    return mcmc(*states, iterations=iterations)

  return run_fn


def sghmc(potential_fn, data_loader, cache_size=512, batch_size=32, integration_steps=10,
          friction=1.0, mass=None, first_step_size=0.05, last_step_size=0.001, burn_in=0,
          accepted_samples=1000, adapt_noise_model=False, diagonal_noise=True,
          save_to_numpy=True, progress_bar=True):
  """alias.py:451-540. adapt_noise_model=True estimates the gradient noise with the
  empirical Fisher information of each minibatch (adaption.fisher_information)."""
  random_data = DT.random_reference_data(data_loader, cache_size, batch_size)
  noise_model = (AD.fisher_information(minibatch_potential=potential_fn,
                                       diagonal=diagonal_noise)
                 if adapt_noise_model else None)
  integ = IG.friction_leapfrog(potential_fn, random_data, friction=friction, const_mass=mass,
                               steps=integration_steps, noise_model=noise_model)
  sched = _schedule(first_step_size, last_step_size, burn_in, accepted_samples)
  saving = IO.save(to_numpy=True) if save_to_numpy else None
  solver = SV.sgmc(integ)
  mcmc = SV.mcmc(solver, sched, strategy="map", saving=saving)

  def run_fn(*init_samples, init_model_state=None, iterations=1000, keys=None,
             data_keys=None, fused=None):
    scalar_fric = np.ndim(friction) == 0 and not isinstance(friction, (dict, list, tuple))
    if (init_model_state is None and mass is None and scalar_fric and not adapt_noise_model
        and _fused_eligible(potential_fn, data_loader, init_samples, fused)):
      return _run_fused(potential_fn, data_loader, batch_size, init_samples, iterations, sched,
                        "sghmc", keys, data_keys, steps=integration_steps,
                        friction=float(friction), to_numpy=save_to_numpy,
                        result_key="energy")
    states = []
    for c, s in enumerate(init_samples):
      kw = dict(init_model_state=init_model_state)
      if keys:
        kw["key"] = keys[c]
      if data_keys:
        kw["batch_kwargs"] = {"key": data_keys[c]}
      states.append(solver[0](s, **kw))
    return mcmc(*states, iterations=iterations)

  return run_fn


def obabo(potential_fn, data_loader, cache_size=512, batch_size=32, integration_steps=10,
          friction=1.0, mass=None, first_step_size=0.05, last_step_size=0.001, burn_in=0,
          accepted_samples=1000, save_to_numpy=True, progress_bar=True):
  """alias.py:542-619."""
  random_data = DT.random_reference_data(data_loader, cache_size, batch_size)
  integ = IG.obabo(potential_fn=potential_fn, batch_fn=random_data, steps=integration_steps,
                   friction=friction, const_mass=mass)
  sched = _schedule(first_step_size, last_step_size, burn_in, accepted_samples)
  saving = IO.save(to_numpy=True) if save_to_numpy else None
  solver = SV.sgmc(integ)
  mcmc = SV.mcmc(solver, sched, strategy="map", saving=saving)

  def run_fn(*init_samples, init_model_state=None, iterations=1000, keys=None,
             data_keys=None, fused=None, precision="fp32"):
    if (fused is not False and init_model_state is None and np.ndim(friction) == 0 and
        not isinstance(friction, (dict, list, tuple)) and
        _mlp_eligible(potential_fn, data_loader, init_samples, batch_size, mass)):
      return _run_fused_mlp(potential_fn, data_loader, batch_size, init_samples, iterations,
                            sched, "obabo", keys, data_keys, to_numpy=save_to_numpy,
                            steps=integration_steps, friction=float(friction),
                            result_key="energy", precision=precision)
    if (init_model_state is None and mass is None and np.ndim(friction) == 0
        and not isinstance(friction, (dict, list, tuple))
        and _fused_eligible(potential_fn, data_loader, init_samples,
                            None if fused is None else bool(fused))):
      return _run_fused(potential_fn, data_loader, batch_size, init_samples, iterations, sched,
                        "obabo", keys, data_keys, steps=integration_steps,
                        friction=float(friction), to_numpy=save_to_numpy,
                        result_key="energy")
    if fused:
      raise ValueError("fused=True needs a GLM family or MLPSoftmax potential over a "
                       "DeviceNumpyDataLoader with identity mass and scalar friction")
    states = []
    for c, s in enumerate(init_samples):
      kw = dict(init_model_state=init_model_state)
      if keys:
        kw["key"] = keys[c]
      if data_keys:
        kw["batch_kwargs"] = {"key": data_keys[c]}
      states.append(solver[0](s, **kw))
    return mcmc(*states, iterations=iterations)

  return run_fn


def re_sgld(potential_fn, data_loader, cache_size=512, batch_size=32, temperature=1000.0,
            first_step_size=0.05, last_step_size=0.001, burn_in=0, accepted_samples=100,
            save_to_numpy=True, progress_bar=True):
  """alias.py:122-208: a normal and a hot chain exchanging states."""
  random_data = DT.random_reference_data(data_loader, cache_size, batch_size)
  integ = IG.langevin_diffusion(potential_fn, random_data)
  sched = _schedule(first_step_size, last_step_size, burn_in, accepted_samples,
                    temperature=SC.constant_temperature(1.0))
  saving = IO.save(to_numpy=True) if save_to_numpy else None
  solver = SV.parallel_tempering(integ)
  mcmc = SV.mcmc(solver, sched, strategy="map", saving=saving)

  def run_fn(*init_samples, init_model_state=None, iterations=1000, keys=None,
             data_keys=None, fused=None):
    flat = [s for pair in init_samples for s in pair]
    if init_model_state is None and _fused_eligible(potential_fn, data_loader, flat, fused):
      return _run_fused_resgld(potential_fn, data_loader, batch_size, init_samples,
                               iterations, sched, temperature, keys, data_keys,
                               to_numpy=save_to_numpy)
    states = []
    for c, (a, b) in enumerate(init_samples):
      kw = dict(init_model_state=init_model_state)
      if keys:
        kw["key"] = keys[c]
      if data_keys:
        kw["batch_kwargs"] = {"key": data_keys[c]}
      states.append(solver[0](a, b, **kw))
    return mcmc(*states, iterations=iterations,
                schedulers=[{"temperature": {"tau": temperature}}])

  return run_fn


def sggmc(stochastic_potential_fn, full_potential_fn, data_loader, cache_size=512,
          batch_size=32, integration_steps=10, friction_coefficient=1.0, first_step_size=0.001,
          last_step_size=0.001, adaptive_step_size=False, stabilization_constant=10,
          decay_constant=0.75, speed_constant=0.05, target_acceptance_rate=0.25, burn_in=0,
          accepted_samples=None, mass=None, save_to_numpy=True, progress_bar=True):
  """alias.py:330-449: OBABO proposals with a full-data Metropolis-Hastings test."""
  random_data = DT.random_reference_data(data_loader, cache_size, batch_size)
  full_map = DT.full_reference_data(data_loader, cache_size, batch_size)
  integ = IG.obabo(stochastic_potential_fn, random_data, integration_steps,
                   friction_coefficient, mass)
  solver = SV.sggmc(integ, full_potential_fn, full_map)
  sched = _mh_schedule(adaptive_step_size, first_step_size, last_step_size, burn_in,
                       accepted_samples, stabilization_constant, decay_constant,
                       speed_constant, target_acceptance_rate)
  saving = IO.save(to_numpy=True) if save_to_numpy else None
  mcmc = SV.mcmc(solver, sched, strategy="map", saving=saving)

  def run_fn(*init_samples, init_model_state=None, iterations=1000, keys=None, data_keys=None,
             fused=None, precision="fp32"):
    from . import mlp as ML
    if (fused is not False and not adaptive_step_size and init_model_state is None and
        mass is None and np.ndim(friction_coefficient) == 0 and
        isinstance(getattr(full_potential_fn, "likelihood", None), ML.MLPSoftmax) and
        _mlp_eligible(stochastic_potential_fn, data_loader, init_samples, batch_size)):
      return _run_fused_mlp(stochastic_potential_fn, data_loader, batch_size, init_samples,
                            iterations, sched, "sggmc", keys, data_keys, to_numpy=save_to_numpy,
                            steps=integration_steps, friction=float(friction_coefficient),
                            result_key="energy", precision=precision)
    if (init_model_state is None and np.ndim(friction_coefficient) == 0 and _mh_fused_eligible(
        stochastic_potential_fn, full_potential_fn, data_loader, init_samples, mass,
        adaptive_step_size, None if fused is None else bool(fused))):
      return _run_fused_mh(stochastic_potential_fn, data_loader, batch_size, init_samples,
                           iterations, sched, "sggmc", integration_steps,
                           float(friction_coefficient), keys, data_keys,
                           to_numpy=save_to_numpy)
    if fused:
      raise ValueError("fused=True needs matching GLM or MLPSoftmax potentials over a "
                       "DeviceNumpyDataLoader with identity mass and a fixed step-size "
                       "schedule")
    if data_keys:
      raise ValueError("data_keys is supported by the fused path only")
    states = []
    for c, s in enumerate(init_samples):
      kw = dict(init_model_state=init_model_state)
      if keys:
        kw["key"] = keys[c]
      states.append(solver[0](s, **kw))
    return mcmc(*states, iterations=iterations)

  return run_fn


def amagold(stochastic_potential_fn, full_potential_fn, data_loader, cache_size=512,
            batch_size=32, integration_steps=10, friction=0.25, first_step_size=0.001,
            last_step_size=0.001, adaptive_step_size=False, stabilization_constant=10,
            decay_constant=0.75, speed_constant=0.05, target_acceptance_rate=0.25, burn_in=0,
            accepted_samples=None, mass=None, save_to_numpy=True, progress_bar=True):
  """alias.py:211-327: reversible-leapfrog proposals with the amortized full-data
  Metropolis-Hastings test (AMAGOLD). The result's samples hold "variables",
  "energy", "acceptance_ratio" and "step_size"."""
  random_data = DT.random_reference_data(data_loader, cache_size, batch_size)
  full_map = DT.full_reference_data(data_loader, cache_size, batch_size)
  integ = IG.reversible_leapfrog(stochastic_potential_fn, random_data, integration_steps,
                                 friction, mass)
  solver = SV.amagold(integ, full_potential_fn, full_map)
  sched = _mh_schedule(adaptive_step_size, first_step_size, last_step_size, burn_in,
                       accepted_samples, stabilization_constant, decay_constant,
                       speed_constant, target_acceptance_rate)
  saving = IO.save(to_numpy=True) if save_to_numpy else None
  mcmc = SV.mcmc(solver, sched, strategy="map", saving=saving)

  def run_fn(*init_samples, init_model_state=None, iterations=1000, keys=None, data_keys=None,
             fused=None):
    if init_model_state is None and _mh_fused_eligible(
        stochastic_potential_fn, full_potential_fn, data_loader, init_samples, mass,
        adaptive_step_size, fused):
      return _run_fused_mh(stochastic_potential_fn, data_loader, batch_size, init_samples,
                           iterations, sched, "amagold", integration_steps, float(friction),
                           keys, data_keys, to_numpy=save_to_numpy)
    if data_keys:
      raise ValueError("data_keys is supported by the fused path only")
    states = []
    for c, s in enumerate(init_samples):
      kw = dict(init_model_state=init_model_state)
      if keys:
        kw["key"] = keys[c]
      states.append(solver[0](s, **kw))
    return mcmc(*states, iterations=iterations)

  return run_fn
