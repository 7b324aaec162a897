"""ctypes binding of libsgmc.so (the HIP library built from csrc/).

Every entry point of include/sgmc.h is declared here with its argument types.
Loading fails loudly: there is no fallback implementation of any kernel.
"""
import atexit
import contextlib
import ctypes
import gc
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsgmc.so")

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float
c_double = ctypes.c_double

ABI_VERSION = 2  # SGMC_ABI_VERSION of the structures mirrored below
OK = 0
E_ARG = 1
E_HIP = 2
PAIRS_PER_BLOCK = 1024
LIK_LINEAR_GAUSSIAN = 0
LIK_LOGISTIC = 1
MODE_SGLD = 0
MODE_SGLD_RMS = 1
MODE_SGHMC = 2
MODE_AMAGOLD = 3
MODE_OBABO = 4
MODE_SGGMC = 5


class GlmChain(ctypes.Structure):
  """Mirror of ``sgmc_glm_chain`` (include/sgmc.h)."""
  _fields_ = [
      ("X", c_void_p), ("y", c_void_p), ("theta", c_void_p), ("mom", c_void_p),
      ("v", c_void_p), ("key", c_void_p), ("dkey", c_void_p), ("U", c_void_p),
      ("var", c_void_p), ("lik", c_void_p), ("partial", c_void_p), ("counter", c_void_p),
      ("samples", c_void_p), ("samples_U", c_void_p), ("N", c_int64), ("ldx", c_int64), ("D", c_int32),
      ("n", c_int32), ("block_begin", c_int32), ("nblocks", c_int32), ("dclass", c_int32),
      ("temperature", c_float), ("energy", c_void_p), ("theta_old", c_void_p),
      ("mom_old", c_void_p),
  ]


class SwapPair(ctypes.Structure):
  """Mirror of ``sgmc_swap_pair`` (include/sgmc.h)."""
  _fields_ = [
      ("lo", c_int32), ("hi", c_int32), ("t_lo", c_float), ("t_hi", c_float),
      ("parity", c_int32), ("reserved", c_int32), ("ssq", c_void_p), ("step", c_void_p),
      ("key", c_void_p), ("swaps", c_void_p), ("samples", c_void_p), ("samples_U", c_void_p),
  ]


class MlpState(ctypes.Structure):
  """Mirror of ``sgmc_mlp_state`` (include/sgmc.h)."""
  _fields_ = [
      ("X", c_void_p), ("labels", c_void_p), ("N", c_int64), ("ldx", c_int64),
      ("D_in", c_int32), ("H1", c_int32), ("H2", c_int32), ("n", c_int32), ("C", c_int32),
      ("theta", c_void_p), ("v", c_void_p), ("shadow", c_void_p), ("P", c_int64),
      ("key", c_void_p), ("dkey", c_void_p), ("work", c_void_p), ("work_bytes", c_int64),
      ("U", c_void_p), ("grad", c_void_p), ("samples", c_void_p), ("samples_U", c_void_p),
      ("n_slots", c_int32), ("eps_table", c_void_p), ("save_slot", c_void_p),
      ("iter", c_void_p), ("counters", c_void_p), ("done", c_void_p), ("table_len", c_int32),
      ("rms", c_int32), ("alpha", c_float), ("lmbd", c_float), ("temperature", c_float),
      ("noise_temperature", c_float), ("has_prior", c_int32), ("prior_scale", c_float),
      ("mode", c_int32), ("mom", c_void_p), ("ke", c_void_p), ("U_half", c_void_p),
      ("decay_table", c_void_p), ("noise_table", c_void_p), ("U_full", c_void_p),
      ("U_prop", c_void_p), ("theta_old", c_void_p), ("mom_old", c_void_p),
      ("U_old", c_void_p), ("mh_key", c_void_p), ("mh_stats", c_void_p),
      ("accepted", c_void_p), ("precision", c_int32), ("shadow8", c_void_p), ("q", c_void_p),
      ("tune", c_int32 * 5),
  ]


_SIGNATURES = {
    "sgmc_last_error": (ctypes.c_char_p, []),
    "sgmc_abi_version": (c_int, []),
    "sgmc_stream_create": (c_int, [c_void_p]),
    "sgmc_stream_destroy": (c_int, [c_void_p]),
    "sgmc_split": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "sgmc_random_bits": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "sgmc_uniform": (c_int, [c_void_p, c_void_p, c_int64, c_float, c_float, c_void_p]),
    "sgmc_normal": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "sgmc_randint": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p]),
    "sgmc_tree_normal": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "sgmc_draw_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int64, c_void_p]),
    "sgmc_gather_rows": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int, c_int64,
                                 c_int64, c_void_p]),
    "sgmc_glm_blocks": (c_int, [c_int, c_int]),
    "sgmc_glm_workspace_floats": (c_int64, [c_int, c_int, c_int]),
    "sgmc_glm_potential": (c_int, [c_int, c_void_p, c_int64, c_void_p, c_void_p, c_int, c_int64,
                                   c_int, c_void_p, c_float, c_int, c_float, c_float, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int,
                                   c_void_p]),
    "sgmc_glm_potential_noalloc": (c_int, [c_int, c_void_p, c_int64, c_void_p, c_void_p, c_int,
                                           c_int64, c_int, c_void_p, c_float, c_int, c_float,
                                           c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                           c_void_p]),
    "sgmc_glm_full_workspace_doubles": (c_int64, [c_int64, c_int]),
    "sgmc_glm_full_potential": (c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64, c_int,
                                        c_void_p, c_float, c_int, c_float, c_double, c_void_p,
                                        c_void_p, c_int64, c_int, c_void_p]),
    "sgmc_glm_full_chains_workspace_doubles": (c_int64, [c_int]),
    "sgmc_glm_full_potential_chains": (c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64,
                                               c_int, c_void_p, c_int64, c_int, c_float, c_int,
                                               c_float, c_double, c_void_p, c_void_p, c_int64,
                                               c_int, c_void_p]),
    "sgmc_glm_full_potential_chains_split": (c_int, [c_int, c_void_p, c_int64, c_void_p, c_int64,
                                                     c_int, c_void_p, c_int64, c_int, c_float,
                                                     c_int, c_float, c_double, c_int, c_void_p,
                                                     c_void_p, c_int64, c_int, c_void_p]),
    "sgmc_glm_check_table": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int]),
    "sgmc_glm_fused_loop": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_int, c_float,
                                    c_float, c_int, c_int, c_float, c_float, c_float, c_void_p,
                                    c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                    c_int, c_void_p]),
    "sgmc_glm_amagold_mh": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "sgmc_glm_sggmc_mh": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                  c_void_p]),
    "sgmc_sgld_step": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int, c_void_p,
                               c_void_p, c_float, c_float, c_float, c_float, c_void_p]),
    "sgmc_sghmc_resample": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                    c_void_p]),
    "sgmc_sghmc_position": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int64,
                                    c_float, c_void_p]),
    "sgmc_sghmc_momentum": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_int,
                                    c_void_p, c_void_p, c_float, c_void_p]),
    "sgmc_obabo_workspace_floats": (c_int64, [c_int]),
    "sgmc_obabo_o": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_float,
                             c_float, c_float, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sgmc_sghmc_momentum_noise_model": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                                c_void_p, c_void_p, c_float, c_void_p]),
    "sgmc_obabo_ba": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_float,
                              c_void_p]),
    "sgmc_obabo_b": (c_int, [c_void_p, c_void_p, c_int64, c_float, c_void_p]),
    "sgmc_amagold_position": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_float,
                                      c_void_p]),
    "sgmc_amagold_momentum": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                      c_void_p, c_void_p, c_float, c_float, c_void_p, c_void_p,
                                      c_void_p, c_float, c_void_p, c_void_p]),
    "sgmc_glm_dclass": (c_int, [c_int]),
    "sgmc_glm_fused_step": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_float,
                                    c_int, c_float, c_float, c_int, c_int, c_int, c_float,
                                    c_float, c_float, c_void_p, c_void_p, c_void_p, c_int,
                                    c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "sgmc_resgld_swap": (c_int, [c_void_p, c_int, c_void_p, c_int, c_float, c_int, c_void_p,
                                 c_int, c_void_p, c_void_p]),
    "sgmc_mlp_full_potential": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sgmc_mlp_init_fp8": (c_int, [c_void_p, c_void_p]),
    "sgmc_mlp_mh": (c_int, [c_void_p, c_void_p]),
    "sgmc_mlp_param_count": (c_int64, [c_int, c_int, c_int]),
    "sgmc_mlp_workspace_bytes": (c_int64, [c_int] * 5),
    "sgmc_mlp_shadow": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "sgmc_mlp_step": (c_int, [ctypes.POINTER(MlpState), c_int, c_void_p]),
    "sgmc_mlp_gemm_plan": (c_int, [c_int] * 7 + [c_void_p, c_void_p]),
}

# Entry points added after ABI version 1 register themselves here.
EXTRA_SIGNATURES = {}

_lib = None


class SgmcError(RuntimeError):
  pass


def load():
  """Loads libsgmc.so once; raises ImportError if it has not been built."""
  global _lib
  if _lib is not None:
    return _lib
  if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: the HIP library must be built first "
        "(python jax-sgmc_amd/build.py, or __graft_entry__.build()). There is no "
        "CPU fallback.")
  lib = ctypes.CDLL(LIB_PATH)
  lib.sgmc_abi_version.restype = c_int
  if lib.sgmc_abi_version() != ABI_VERSION:
    raise ImportError(f"{LIB_PATH} has ABI version {lib.sgmc_abi_version()}, the bindings "
                      f"expect {ABI_VERSION}: rebuild it (python jax-sgmc_amd/build.py)")
  for name, (res, args) in list(_SIGNATURES.items()) + list(EXTRA_SIGNATURES.items()):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
  _lib = lib
  return lib


def exported_names():
  return sorted(list(_SIGNATURES) + list(EXTRA_SIGNATURES))


def call(name, *args):
  """Calls an entry point and raises on a nonzero status."""
  lib = load()
  status = getattr(lib, name)(*args)
  if status != OK:
    msg = lib.sgmc_last_error().decode(errors="replace")
    if status == E_ARG:
      raise ValueError(f"{name}: {msg}")
    raise SgmcError(f"{name} failed ({status}): {msg}")
  return status


def query(name, *args):
  return getattr(load(), name)(*args)


def ptr(t):
  """Device pointer of a tensor (None -> NULL). Tensors must live on the GPU."""
  if t is None:
    return None
  if not t.is_cuda:
    raise ValueError("libsgmc entry points take device tensors; got a CPU tensor")
  return c_void_p(t.data_ptr())


@contextlib.contextmanager
def capture_guard():
  """Around a HIP graph capture: collects garbage first and keeps the cyclic collector
  off until the capture ends. A graph kept alive only by a dead reference cycle would
  otherwise be destroyed whenever the collector runs, and destroying a graph while a
  stream is capturing is an error that aborts the process."""
  gc.collect()
  enabled = gc.isenabled()
  gc.disable()
  try:
    yield
  finally:
    if enabled:
      gc.enable()


_DEDICATED = {}


def _release_dedicated():
  """At interpreter exit: waits for the pool's streams and destroys them, so that the
  runtime's own teardown finds no queue of ours left."""
  for (dev, _), st in list(_DEDICATED.items()):
    try:
      with torch.cuda.device(dev):
        st.synchronize()
        call("sgmc_stream_destroy", c_void_p(st.cuda_stream))
    except Exception:  # exiting: nothing left to report to
      pass
  _DEDICATED.clear()


def dedicated_stream(device, index):
  """Stream ``index`` of a per-device pool of streams with hardware queues of their own
  (sgmc_stream_create), created on first use and destroyed at interpreter exit; wrapped
  as a torch stream."""
  dev = torch.device(device)
  key = (dev.index if dev.index is not None else torch.cuda.current_device(), int(index))
  if key not in _DEDICATED:
    if not _DEDICATED:
      atexit.register(_release_dedicated)
    with torch.cuda.device(key[0]):
      h = c_void_p()
      call("sgmc_stream_create", ctypes.byref(h))
    _DEDICATED[key] = torch.cuda.ExternalStream(h.value, device=torch.device("cuda", key[0]))
  return _DEDICATED[key]


def stream():
  return c_void_p(torch.cuda.current_stream().cuda_stream)


def require_gpu():
  if not torch.cuda.is_available():
    raise RuntimeError("the sgmc kernels need a ROCm GPU (torch.cuda.is_available() is False)")
  load()
