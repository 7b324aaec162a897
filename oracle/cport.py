"""ctypes handle on the oracle's C port (oracle/_build/liboracle_port.so).

Test infrastructure: the timed CPU baseline of bench.py and a compiled cross-check of
the numpy oracle. Build with ``make -C oracle``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle_port.so")
_lib = None


def load():
  global _lib
  if _lib is None:
    if not os.path.exists(LIB):
      raise ImportError(f"{LIB} missing: run make -C oracle")
    _lib = ctypes.CDLL(LIB)
  return _lib


def _p(a):
  return a.ctypes.data_as(ctypes.c_void_p)


def sghmc(family, X, y, n, theta, keys, dkeys, eps, steps, friction, sigma=1.0,
          prior_scale=1.0, threads=1):
  """Runs len(eps) SGHMC iterations for theta.shape[0] chains in place."""
  lib = load()
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
         prior_scale=1.0, threads=1):
  lib = load()
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
            sigma=1.0, prior_scale=1.0, threads=1):
  """Runs len(eps) AMAGOLD iterations for theta.shape[0] chains in place; returns the
  accept decisions [C, iterations]."""
  lib = load()
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
          temperature=1.0, sigma=1.0, prior_scale=1.0, threads=1):
  """Runs len(eps) SGGMC iterations (OBABO proposals, full-data MH test) for
  theta.shape[0] chains in place; returns the accept decisions [C, iterations]."""
  lib = load()
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
  lib.cport_full_potential.restype = ctypes.c_double
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
