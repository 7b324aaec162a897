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


def normal(key, S):
  lib = load()
  out = np.zeros(S, np.float32)
  k = np.ascontiguousarray(key, np.uint32)
  lib.cport_normal(_p(k), ctypes.c_uint32(S), _p(out))
  return out
