"""ctypes handle on the oracle's C port (oracle/_build/liboracle_port*.so).

Test infrastructure: the timed CPU baseline of bench.py and a compiled cross-check of
the numpy oracle. Build with ``make -C oracle``. ``fast=False`` (the default) loads the
oracle build (written operation order, float64 accumulators: what tests/ compare);
``fast=True`` the baseline build (vectorized float32 sums with FMA contraction, AVX-512
when the host has it, AVX2 otherwise), which bench.py times.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle_port.so")
_libs = {}


def host_has_avx512():
  try:
    flags = open("/proc/cpuinfo").read()
  except OSError:
    return False
  return all(f in flags for f in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))


def lib_path(fast=False):
  if not fast:
    return LIB
  return os.path.join(_HERE, "_build",
                      f"liboracle_port_fast_{'v4' if host_has_avx512() else 'v3'}.so")


def load(fast=False):
  path = lib_path(fast)
  if path not in _libs:
    # The MLP port alternates BLAS calls (numpy's own thread pool) with OpenMP loops:
    # threads that spin while the other pool works halve its speed, so OpenMP threads
    # sleep between parallel regions (read by libgomp when it first loads).
    os.environ.setdefault("OMP_WAIT_POLICY", "passive")
    if not os.path.exists(path):
      raise ImportError(f"{path} missing: run make -C oracle")
    lib = ctypes.CDLL(path)
    lib.cport_obabo_first.restype = ctypes.c_double
    lib.cport_obabo_second.restype = ctypes.c_double
    lib.cport_full_potential.restype = ctypes.c_double
    _libs[path] = lib
  return _libs[path]


def _p(a):
  return a.ctypes.data_as(ctypes.c_void_p)


def sghmc(family, X, y, n, theta, keys, dkeys, eps, steps, friction, sigma=1.0,
          prior_scale=1.0, threads=1, fast=False):
  """Runs len(eps) SGHMC iterations for theta.shape[0] chains in place."""
  lib = load(fast)
  X = np.ascontiguousarray(X, np.float32)
  y = np.ascontiguousarray(y, np.float32)
  eps = np.ascontiguousarray(eps, np.float32)
  C, D = theta.shape
  U = np.zeros(C, np.float32)
  lib.cport_sghmc(ctypes.c_int(family), _p(X), ctypes.c_int64(X.shape[1]), _p(y),
                  ctypes.c_int64(X.shape[0]), ctypes.c_int(D), ctypes.c_int(n), ctypes.c_int(C),
                  _p(theta), _p(keys), _p(dkeys), _p(eps), ctypes.c_int(eps.size),
                  ctypes.c_int(steps), ctypes.c_float(friction), ctypes.c_float(sigma),
                  ctypes.c_float(prior_scale), _p(U), ctypes.c_int(threads))
  return U


def sgld(family, X, y, n, theta, keys, dkeys, eps, temperature=1.0, sigma=1.0,
         prior_scale=1.0, threads=1, fast=False):
  lib = load(fast)
  X = np.ascontiguousarray(X, np.float32)
  y = np.ascontiguousarray(y, np.float32)
  eps = np.ascontiguousarray(eps, np.float32)
  C, D = theta.shape
  U = np.zeros(C, np.float32)
  lib.cport_sgld(ctypes.c_int(family), _p(X), ctypes.c_int64(X.shape[1]), _p(y),
                 ctypes.c_int64(X.shape[0]), ctypes.c_int(D), ctypes.c_int(n), ctypes.c_int(C),
                 _p(theta), _p(keys), _p(dkeys), _p(eps), ctypes.c_float(temperature),
                 ctypes.c_int(eps.size), ctypes.c_float(sigma), ctypes.c_float(prior_scale),
                 _p(U), ctypes.c_int(threads))
  return U


def amagold_init(theta0, key):
  """solver.amagold init (solver.py:340-352) for one chain with unit mass: returns the
  initial momentum, the integrator key and the MH key (numpy oracle RNG)."""
  from oracle import jaxrandom as jr
  from oracle import steps as ost
  k_int, k_mh = jr.split(np.asarray(key, np.uint32))
  p, k_int = ost.reversible_leapfrog_init(np.asarray(theta0), k_int, [np.size(theta0)])
  return p.astype(np.float32), k_int, k_mh


def amagold(family, X, y, n, theta, mom, keys, mh_keys, dkeys, U_full, eps, steps, friction,
            sigma=1.0, prior_scale=1.0, threads=1, fast=False):
  """Runs len(eps) AMAGOLD iterations for theta.shape[0] chains in place; returns the
  accept decisions [C, iterations]."""
  lib = load(fast)
  X = np.ascontiguousarray(X, np.float32)
  y = np.ascontiguousarray(y, np.float32)
  eps = np.ascontiguousarray(eps, np.float32)
  C, D = theta.shape
  acc = np.zeros((C, eps.size), np.int8)
  lib.cport_amagold(ctypes.c_int(family), _p(X), ctypes.c_int64(X.shape[1]), _p(y),
                    ctypes.c_int64(X.shape[0]), ctypes.c_int(D), ctypes.c_int(n),
                    ctypes.c_int(C), _p(theta), _p(mom), _p(keys), _p(mh_keys), _p(dkeys),
                    _p(U_full), _p(eps), ctypes.c_int(eps.size), ctypes.c_int(steps),
                    ctypes.c_float(friction), ctypes.c_float(sigma),
                    ctypes.c_float(prior_scale), _p(acc), ctypes.c_int(threads))
  return acc.astype(bool)


def sggmc(family, X, y, n, theta, mom, keys, mh_keys, dkeys, U_full, eps, steps, friction,
          temperature=1.0, sigma=1.0, prior_scale=1.0, threads=1, fast=False):
  """Runs len(eps) SGGMC iterations (OBABO proposals, full-data MH test) for
  theta.shape[0] chains in place; returns the accept decisions [C, iterations]."""
  lib = load(fast)
  X = np.ascontiguousarray(X, np.float32)
  y = np.ascontiguousarray(y, np.float32)
  eps = np.ascontiguousarray(eps, np.float32)
  C, D = theta.shape
  acc = np.zeros((C, eps.size), np.int8)
  lib.cport_sggmc(ctypes.c_int(family), _p(X), ctypes.c_int64(X.shape[1]), _p(y),
                  ctypes.c_int64(X.shape[0]), ctypes.c_int(D), ctypes.c_int(n),
                  ctypes.c_int(C), _p(theta), _p(mom), _p(keys), _p(mh_keys), _p(dkeys),
                  _p(U_full), _p(eps), ctypes.c_int(eps.size), ctypes.c_int(steps),
                  ctypes.c_float(friction), ctypes.c_float(temperature), ctypes.c_float(sigma),
                  ctypes.c_float(prior_scale), _p(acc), ctypes.c_int(threads))
  return acc.astype(bool)


def full_potential(family, X, y, theta, sigma=1.0, prior_scale=1.0):
  """FP64 full-data potential of one position (the C port's MH reference value)."""
  lib = load()
  X = np.ascontiguousarray(X, np.float32)
  y = np.ascontiguousarray(y, np.float32)
  w = np.ascontiguousarray(theta, np.float32)
  return lib.cport_full_potential(ctypes.c_int(family), _p(X), ctypes.c_int64(X.shape[1]),
                                  _p(y), ctypes.c_int64(X.shape[0]), ctypes.c_int(X.shape[1]),
                                  _p(w), ctypes.c_float(sigma), ctypes.c_float(prior_scale))


def normal(key, S):
  lib = load()
  out = np.zeros(S, np.float32)
  k = np.ascontiguousarray(key, np.uint32)
  lib.cport_normal(_p(k), ctypes.c_uint32(S), _p(out))
  return out


# ---- helpers of the float32 MLP port (oracle/mlp_port.py) ----

def tree_normals(key, sizes, threads=1, fast=False):
  """random_tree(key) over a flat vector with leaves of the given sizes."""
  lib = load(fast)
  sz = np.ascontiguousarray(sizes, np.int64)
  out = np.empty(int(sz.sum()), np.float32)
  k = np.ascontiguousarray(key, np.uint32)
  lib.cport_tree_normals(_p(k), _p(sz), ctypes.c_int(sz.size), _p(out), ctypes.c_int(threads))
  return out


def draw_indices(dkey, n, N, fast=False):
  """Advances dkey (uint32[2], in place) and returns the n minibatch indices."""
  lib = load(fast)
  idx = np.empty(n, np.int32)
  lib.cport_draw_indices(_p(dkey), ctypes.c_uint32(n), ctypes.c_uint32(N), _p(idx))
  return idx


def sgld_update(theta, v, g, xi, eps, noise_scale, alpha, lmbd, rms, threads=1, fast=False):
  load(fast).cport_sgld_update(
      _p(theta), _p(v) if v is not None else None, _p(g), _p(xi), ctypes.c_int64(theta.size),
      ctypes.c_float(eps), ctypes.c_float(noise_scale), ctypes.c_float(alpha),
      ctypes.c_float(lmbd), ctypes.c_int(1 if rms else 0), ctypes.c_int(threads))


def obabo_first(theta, p, g, xi, eps, decay, noise, threads=1, fast=False):
  return load(fast).cport_obabo_first(_p(theta), _p(p), _p(g), _p(xi),
                                      ctypes.c_int64(theta.size), ctypes.c_float(eps),
                                      ctypes.c_float(decay), ctypes.c_float(noise),
                                      ctypes.c_int(threads))


def obabo_second(p, g, xi, eps, decay, noise, threads=1, fast=False):
  return load(fast).cport_obabo_second(_p(p), _p(g), _p(xi), ctypes.c_int64(p.size),
                                       ctypes.c_float(eps), ctypes.c_float(decay),
                                       ctypes.c_float(noise), ctypes.c_int(threads))
