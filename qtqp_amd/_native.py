"""ctypes bindings of the HIP libraries built into qtqp_amd/_lib.

Loading fails loudly (OSError) when a library has not been built or no HIP
device is usable: there is no CPU fallback anywhere in the product path.
"""

from __future__ import annotations

import ctypes
import pathlib
import threading

import numpy as np

LIBDIR = pathlib.Path(__file__).resolve().parent / "_lib"

QTQP_OK = 0
QTQP_ERR_ARG = 1
QTQP_ERR_HIP = 2
QTQP_ERR_FACTOR = 3
QTQP_ERR_NAN = 4
QTQP_ERR_STATE = 5

PREC_FP64 = 0
PREC_MIXED = 1

_c_double_p = ctypes.POINTER(ctypes.c_double)
_lock = threading.Lock()
_libs: dict[str, ctypes.CDLL] = {}


class IrStats(ctypes.Structure):
  _fields_ = [("solves", ctypes.c_int32), ("status", ctypes.c_int32),
              ("final_residual_norm", ctypes.c_double), ("rhs_norm", ctypes.c_double),
              ("tolerance", ctypes.c_double)]


class KktInfo(ctypes.Structure):
  _fields_ = [("precision", ctypes.c_int32), ("factor_precision", ctypes.c_int32),
              ("factorizations", ctypes.c_int64), ("fp64_fallbacks", ctypes.c_int64),
              ("backend_solves", ctypes.c_int64), ("inner_steps", ctypes.c_int64),
              ("factor_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
              ("jitter_retries", ctypes.c_int64), ("graph_redos", ctypes.c_int64)]


class IpmOpts(ctypes.Structure):
  """qtqp_ipm_opts (include/qtqp_ipm.h)."""
  _fields_ = [("atol", ctypes.c_double), ("rtol", ctypes.c_double),
              ("atol_infeas", ctypes.c_double), ("rtol_infeas", ctypes.c_double),
              ("step_size_scale", ctypes.c_double), ("min_static_regularization", ctypes.c_double),
              ("linear_solver_atol", ctypes.c_double), ("linear_solver_rtol", ctypes.c_double),
              ("max_iterative_refinement_steps", ctypes.c_int32),
              ("collect_stats", ctypes.c_int32), ("use_graph", ctypes.c_int32),
              ("graph_steps", ctypes.c_int32)]


IPM_RECORD_DOUBLES = ("ctx", "bty", "pcost", "dcost", "pres", "dres", "gap", "pinfeas", "dinfeas",
                      "dinfeas_a", "dinfeas_p", "mu", "sigma", "alpha", "tau", "norm_x", "norm_y",
                      "prelrhs", "drelrhs", "complementarity", "norm_s", "max_sy", "min_sy",
                      "std_sy", "max_s_over_y", "min_s_over_y", "mean_s_over_y", "std_s_over_y")


class IpmRecord(ctypes.Structure):
  """qtqp_ipm_record (include/qtqp_ipm.h)."""
  _fields_ = ([("status", ctypes.c_int32), ("tau_method", ctypes.c_int32 * 2),
               ("reserved", ctypes.c_int32), ("lin", IrStats * 3)]
              + [(name, ctypes.c_double) for name in IPM_RECORD_DOUBLES])


# Exported symbol -> (restype, argtypes), matching include/qtqp_kkt.h.
KKT_SYMBOLS = {
    "qtqp_kkt_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _c_double_p, _c_double_p,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_kkt_update_diag": (ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    "qtqp_kkt_factorize": (ctypes.c_int, [ctypes.c_void_p]),
    "qtqp_kkt_solve": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    "qtqp_kkt_matvec": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    "qtqp_kkt_set_correction": (ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    "qtqp_kkt_matvec_noalloc": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]),
    "qtqp_kkt_solve_ir": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p, ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_double, _c_double_p,
                                         ctypes.POINTER(IrStats)]),
    "qtqp_kkt_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(KktInfo)]),
    "qtqp_kkt_free": (None, [ctypes.c_void_p]),
    "qtqp_last_error": (ctypes.c_char_p, []),
    "qtqp_device_count": (ctypes.c_int, []),
}


# Exported symbol -> (restype, argtypes), matching include/qtqp_ipm.h (same library).
IPM_SYMBOLS = {
    "qtqp_ipm_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _c_double_p,
                                       _c_double_p, ctypes.c_int32, ctypes.c_int32, _c_double_p,
                                       _c_double_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_ipm_iterate": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(IpmOpts),
                                        ctypes.POINTER(IpmRecord)]),
    "qtqp_ipm_get_iterates": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p, _c_double_p,
                                             _c_double_p]),
    "qtqp_ipm_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(KktInfo)]),
    "qtqp_ipm_free": (None, [ctypes.c_void_p]),
}


class SparseInfo(ctypes.Structure):
  """qtqp_sparse_info (include/qtqp_sparse.h)."""
  _fields_ = [("n", ctypes.c_int64), ("nnz_kkt", ctypes.c_int64), ("nnz_l", ctypes.c_int64),
              ("supernodes", ctypes.c_int64), ("levels", ctypes.c_int64),
              ("max_rows", ctypes.c_int64), ("update_tiles", ctypes.c_int64),
              ("factorizations", ctypes.c_int64), ("solves", ctypes.c_int64),
              ("factor_flops", ctypes.c_double), ("analyze_ms", ctypes.c_double),
              ("factor_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
              ("phases", ctypes.c_int64)]


_c_int64_p = ctypes.POINTER(ctypes.c_int64)

SPARSE_CSC = 0
SPARSE_CSR = 1

# Exported symbol -> (restype, argtypes), matching include/qtqp_sparse.h.
SPARSE_SYMBOLS = {
    "qtqp_sparse_analyze": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _c_int64_p, _c_int64_p,
                                           _c_int64_p, ctypes.POINTER(SparseInfo)]),
    "qtqp_sparse_analyze_ex": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _c_int64_p,
                                              _c_int64_p, ctypes.c_int32, _c_int64_p,
                                              ctypes.POINTER(SparseInfo)]),
    "qtqp_sparse_create": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _c_int64_p, _c_int64_p,
                                          _c_double_p, ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_sparse_update_diag": (ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    "qtqp_sparse_factorize": (ctypes.c_int, [ctypes.c_void_p]),
    "qtqp_sparse_solve": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    "qtqp_sparse_matvec": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    "qtqp_sparse_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SparseInfo)]),
    "qtqp_sparse_free": (None, [ctypes.c_void_p]),
    "qtqp_sparse_last_error": (ctypes.c_char_p, []),
    "qtqp_sparse_device_count": (ctypes.c_int, []),
}


def lib_path(name: str) -> pathlib.Path:
  return LIBDIR / name


def load(name: str, symbols: dict) -> ctypes.CDLL:
  """Loads a built library and declares its C signatures (raises OSError)."""
  with _lock:
    if name in _libs:
      return _libs[name]
    path = lib_path(name)
    if not path.exists():
      raise OSError(f"{path} is not built; run `python -m qtqp_amd.build` "
                    "(or __graft_entry__.build()) first")
    lib = ctypes.CDLL(str(path))
    for sym, (res, args) in symbols.items():
      fn = getattr(lib, sym)
      fn.restype = res
      fn.argtypes = args
    _libs[name] = lib
    return lib


def kkt_lib() -> ctypes.CDLL:
  return load("libqtqp_kkt.so", {**KKT_SYMBOLS, **IPM_SYMBOLS})


def sparse_lib() -> ctypes.CDLL:
  return load("libqtqp_sparse.so", SPARSE_SYMBOLS)


def iptr(a: np.ndarray):
  return a.ctypes.data_as(_c_int64_p)


def as_i64(a) -> np.ndarray:
  return np.ascontiguousarray(a, dtype=np.int64)


def require_device(lib: ctypes.CDLL) -> None:
  count = lib.qtqp_sparse_device_count if hasattr(lib, "qtqp_sparse_device_count") else \
      lib.qtqp_device_count
  if count() < 1:
    raise OSError("no HIP device is visible; the qtqp_amd HIP backends need an MI355X GPU")


def dptr(a: np.ndarray):
  return a.ctypes.data_as(_c_double_p)


def as_f64(a) -> np.ndarray:
  return np.ascontiguousarray(a, dtype=np.float64)


def check(lib: ctypes.CDLL, rc: int) -> None:
  """Maps a C status to the exception the reference backends raise."""
  if rc == QTQP_OK:
    return
  last = lib.qtqp_sparse_last_error if hasattr(lib, "qtqp_sparse_last_error") else \
      lib.qtqp_last_error
  msg = (last() or b"").decode(errors="replace")
  if rc == QTQP_ERR_FACTOR:
    raise np.linalg.LinAlgError(msg)
  if rc in (QTQP_ERR_ARG, QTQP_ERR_NAN):
    raise ValueError(msg)
  raise RuntimeError(f"qtqp_amd native error {rc}: {msg}")
