"""KKT linear-system layer: the LinearSolver plug-in protocol, the refined
KKT solver, and the MI355X HIP backends.

Mirrors qtqp v0.0.5 ``src/qtqp/direct.py``: ``diag_data_indices`` (33-46),
``LinearSolver`` (49-109) and ``DirectKktSolver`` (112-311), so any backend
written for the reference protocol runs here unchanged, and the HIP backends
below run under the reference's own ``DirectKktSolver`` unchanged.
"""

from __future__ import annotations

import logging
from typing import Any, Literal

import numpy as np
import scipy.sparse as sp

from . import _native

_IR_STATUS = {0: "converged", 1: "non-converged", 2: "stalled"}


def diag_data_indices(mat: sp.spmatrix) -> np.ndarray:
  """Index into ``mat.data`` of each diagonal entry of a CSC/CSR matrix
  (direct.py:33-46).  Requires every diagonal entry to be stored."""
  mat.sort_indices()
  k = np.arange(mat.shape[0])
  starts = mat.indptr[:-1]
  out = np.empty(mat.shape[0], dtype=np.intp)
  for i in range(mat.shape[0]):
    seg = mat.indices[starts[i]:mat.indptr[i + 1]]
    out[i] = starts[i] + np.searchsorted(seg, k[i])
  return out


class LinearSolver:
  """Backend protocol for the KKT system (direct.py:49-109).

  A backend receives the dimensions, then the upper-triangular KKT scaffold,
  then per IPM iteration a regularised diagonal (y-block negated) followed by
  ``factorize``; ``solve`` applies K_reg^{-1} and ``@`` applies K_reg.
  """

  def set_dims(self, n: int, m: int, z: int) -> None:
    pass

  def set_kkt(self, kkt: sp.spmatrix) -> None:
    self._kkt = kkt
    self._kkt_diag = kkt.diagonal()
    self._kkt_diag_idxs = diag_data_indices(kkt)

  def update_diag(self, diag: np.ndarray) -> None:
    self._kkt.data[self._kkt_diag_idxs] = diag
    np.copyto(self._kkt_diag, diag)

  def factorize(self) -> None:
    raise NotImplementedError

  def solve(self, rhs: np.ndarray) -> np.ndarray:
    raise NotImplementedError

  def __matmul__(self, x: np.ndarray) -> np.ndarray:
    # Symmetric product from the stored upper triangle: U x + U' x - diag*x.
    out = self._kkt @ x
    out += self._kkt.T @ x
    out -= self._kkt_diag * x
    return out

  def format(self) -> str:
    raise NotImplementedError

  def free(self) -> None:
    pass


class _HipDenseBase(LinearSolver):
  """Dense KKT backend on an MI355X through the qtqp_amd C ABI.

  Data stays on the device for the life of the handle; each protocol call
  copies only N-vectors across PCIe.  ``solve`` is accurate to FP64 in one
  call (the mixed-precision handle refines internally), which is what the
  reference protocol expects of a direct backend.
  """

  _precision = _native.PREC_FP64
  # QTQP.solve runs the whole iteration loop on the device (DeviceIpm) for
  # backends that set this; the protocol methods below stay available.
  device_driver = True
  # Each device-loop iteration is one HIP graph launch with the refinement and
  # precision decisions taken on the device (False: host-driven launches);
  # device_graph_steps refinement steps per solve are in the graph (0: adaptive).
  device_graph = True
  device_graph_steps = 0

  def __init__(self):
    self._lib = _native.kkt_lib()
    _native.require_device(self._lib)
    self._h = None
    self._n = self._m = 0

  # -- protocol ------------------------------------------------------------
  def set_dims(self, n: int, m: int, z: int) -> None:
    self._n, self._m = int(n), int(m)

  def set_kkt(self, kkt: sp.spmatrix) -> None:
    super().set_kkt(kkt)
    n = self._n
    kkt = sp.csr_matrix(kkt)
    a = kkt[:n, n:].T.toarray()
    pu = kkt[:n, :n].toarray()
    p = pu + pu.T - np.diag(np.diag(pu))
    self.set_dense_blocks(a, p)

  def set_dense_blocks(self, a: np.ndarray, p: np.ndarray | None) -> None:
    """Loads the dense A (m x n) and P (n x n) blocks, skipping the scaffold."""
    self._release()
    a = _native.as_f64(a)
    self._m, self._n = a.shape
    pp = None if p is None else _native.as_f64(p)
    h = _native.ctypes.c_void_p()
    rc = self._lib.qtqp_kkt_create(self._n, self._m, _native.dptr(a),
                                   None if pp is None else _native.dptr(pp),
                                   self._precision, _native.ctypes.byref(h))
    _native.check(self._lib, rc)
    self._h = h
    self._buf = np.empty(self._n + self._m)

  def update_diag(self, diag: np.ndarray) -> None:
    d = _native.as_f64(diag)
    _native.check(self._lib, self._lib.qtqp_kkt_update_diag(self._h, _native.dptr(d)))

  def factorize(self) -> None:
    _native.check(self._lib, self._lib.qtqp_kkt_factorize(self._h))

  def solve(self, rhs: np.ndarray) -> np.ndarray:
    r = _native.as_f64(rhs)
    out = np.empty_like(r)
    _native.check(self._lib, self._lib.qtqp_kkt_solve(self._h, _native.dptr(r), _native.dptr(out)))
    return out

  def __matmul__(self, x: np.ndarray) -> np.ndarray:
    v = _native.as_f64(x)
    out = np.empty_like(v)
    _native.check(self._lib, self._lib.qtqp_kkt_matvec(self._h, _native.dptr(v), _native.dptr(out)))
    return out

  def format(self) -> Literal["csr"]:
    return "csr"

  def free(self) -> None:
    self._release()

  def device_ipm(self, a, p, has_p, b, c, z, equilibrate) -> DeviceIpm:
    """The device-resident IPM driver on this backend's engine and precision."""
    return DeviceIpm(self._lib, self._precision, a, p, has_p, b, c, z, equilibrate)

  # -- device-resident refinement (used by DirectKktSolver when present) ---
  def set_correction(self, corr: np.ndarray) -> None:
    c = _native.as_f64(corr)
    _native.check(self._lib, self._lib.qtqp_kkt_set_correction(self._h, _native.dptr(c)))

  def solve_refined(self, rhs, warm_start, max_steps, atol, rtol):
    r = _native.as_f64(rhs)
    w = _native.as_f64(warm_start)
    sol = np.empty_like(r)
    st = _native.IrStats()
    rc = self._lib.qtqp_kkt_solve_ir(self._h, _native.dptr(r), _native.dptr(w), int(max_steps),
                                     float(atol), float(rtol), _native.dptr(sol),
                                     _native.ctypes.byref(st))
    _native.check(self._lib, rc)
    status = _IR_STATUS[st.status]
    return sol, {"solves": int(st.solves), "final_residual_norm": float(st.final_residual_norm),
                 "rhs_norm": float(st.rhs_norm), "tolerance": float(st.tolerance),
                 "converged": status == "converged", "status": status}

  def info(self) -> dict[str, Any]:
    inf = _native.KktInfo()
    _native.check(self._lib, self._lib.qtqp_kkt_get_info(self._h, _native.ctypes.byref(inf)))
    return {k: getattr(inf, k) for k, _ in inf._fields_}

  def _release(self) -> None:
    if getattr(self, "_h", None) is not None:
      self._lib.qtqp_kkt_free(self._h)
      self._h = None

  def __del__(self):
    try:
      self._release()
    except Exception:  # pylint: disable=broad-except
      pass


class DeviceIpm:
  """One QP held on the device for the device-resident IPM driver
  (include/qtqp_ipm.h): ``iterate`` runs one whole IPM iteration on the GPU
  and returns its record; the iterate never leaves HBM until ``iterates``.
  """

  def __init__(self, lib, precision: int, a: np.ndarray, p: np.ndarray | None, has_p: bool,
               b: np.ndarray, c: np.ndarray, z: int, equilibrate: bool):
    self._lib = lib
    self._h = None
    self.m, self.n = a.shape
    # Column-major inputs (what scipy's CSC toarray() returns) go up as they
    # are and are transposed on the device.
    col_major = (a.dtype == np.float64 and a.flags.f_contiguous and not a.flags.c_contiguous and
                 (p is None or (p.dtype == np.float64 and p.flags.f_contiguous)))
    if col_major:
      a, p = a.T, (None if p is None else p.T)
    a = _native.as_f64(a)
    pp = None if p is None else _native.as_f64(p)
    bb, cc = _native.as_f64(b), _native.as_f64(c)
    h = _native.ctypes.c_void_p()
    rc = lib.qtqp_ipm_create(self.n, self.m, int(z), _native.dptr(a),
                             None if pp is None else _native.dptr(pp), int(col_major),
                             int(bool(has_p)),
                             _native.dptr(bb), _native.dptr(cc), precision, int(bool(equilibrate)),
                             _native.ctypes.byref(h))
    _native.check(lib, rc)
    self._h = h
    self._rec = _native.IpmRecord()

  def iterate(self, opts: "_native.IpmOpts") -> "_native.IpmRecord":
    _native.check(self._lib, self._lib.qtqp_ipm_iterate(self._h, _native.ctypes.byref(opts),
                                                        _native.ctypes.byref(self._rec)))
    return self._rec

  def iterates(self):
    """Unequilibrated (x, y, s, tau) of the current iterate."""
    x, y, s = np.empty(self.n), np.empty(self.m), np.empty(self.m)
    tau = _native.ctypes.c_double()
    _native.check(self._lib, self._lib.qtqp_ipm_get_iterates(
        self._h, _native.dptr(x), _native.dptr(y), _native.dptr(s),
        _native.ctypes.cast(_native.ctypes.byref(tau), _native._c_double_p)))
    return x, y, s, float(tau.value)

  def info(self) -> dict[str, Any]:
    inf = _native.KktInfo()
    _native.check(self._lib, self._lib.qtqp_ipm_get_info(self._h, _native.ctypes.byref(inf)))
    return {k: getattr(inf, k) for k, _ in inf._fields_}

  def free(self) -> None:
    if getattr(self, "_h", None) is not None:
      self._lib.qtqp_ipm_free(self._h)
      self._h = None

  def __del__(self):
    try:
      self.free()
    except Exception:  # pylint: disable=broad-except
      pass


class HipDenseSolver(_HipDenseBase):
  """Mixed precision: FP32 MFMA factor, FP64 refinement, FP64 fallback."""
  _precision = _native.PREC_MIXED


class HipDenseFp64Solver(_HipDenseBase):
  """Plain FP64: FP64 MFMA Gram product and factor, FP64 solves."""
  _precision = _native.PREC_FP64


class HipSparseSolver(LinearSolver):
  """Sparse KKT backend on an MI355X: supernodal LDL' through the qtqp_amd
  sparse C ABI (include/qtqp_sparse.h).

  The counterpart of the reference's sparse direct backends, CuDssSolver
  (solvers_gpu.py:27-115) and QdldlSolver (solvers_sparse.py:74-93): it takes
  the upper-triangular KKT scaffold in CSC, analyses its pattern once
  (ordering, elimination tree, supernodes) and keeps pattern, values and
  factor on the device.  ``solve`` is one application of the FP64 factor;
  DirectKktSolver refines around it on the host, as it does for cuDSS.
  """

  def __init__(self):
    self._lib = _native.sparse_lib()
    _native.require_device(self._lib)
    self._h = None

  def set_kkt(self, kkt: sp.spmatrix) -> None:
    super().set_kkt(kkt)
    self._release()
    csc = sp.csc_matrix(kkt)
    csc.sort_indices()
    ptr, ind = _native.as_i64(csc.indptr), _native.as_i64(csc.indices)
    val = _native.as_f64(csc.data)
    h = _native.ctypes.c_void_p()
    rc = self._lib.qtqp_sparse_create(csc.shape[0], _native.SPARSE_CSC, _native.iptr(ptr),
                                      _native.iptr(ind), _native.dptr(val),
                                      _native.ctypes.byref(h))
    _native.check(self._lib, rc)
    self._h = h

  def update_diag(self, diag: np.ndarray) -> None:
    d = _native.as_f64(diag)
    _native.check(self._lib, self._lib.qtqp_sparse_update_diag(self._h, _native.dptr(d)))
    np.copyto(self._kkt_diag, d)

  def factorize(self) -> None:
    _native.check(self._lib, self._lib.qtqp_sparse_factorize(self._h))

  def solve(self, rhs: np.ndarray) -> np.ndarray:
    r = _native.as_f64(rhs)
    out = np.empty_like(r)
    _native.check(self._lib, self._lib.qtqp_sparse_solve(self._h, _native.dptr(r),
                                                         _native.dptr(out)))
    return out

  def __matmul__(self, x: np.ndarray) -> np.ndarray:
    v = _native.as_f64(x)
    out = np.empty_like(v)
    _native.check(self._lib, self._lib.qtqp_sparse_matvec(self._h, _native.dptr(v),
                                                          _native.dptr(out)))
    return out

  def format(self) -> Literal["csc"]:
    return "csc"

  def info(self) -> dict[str, Any]:
    inf = _native.SparseInfo()
    _native.check(self._lib, self._lib.qtqp_sparse_get_info(self._h, _native.ctypes.byref(inf)))
    return {k: getattr(inf, k) for k, _ in inf._fields_}

  def free(self) -> None:
    self._release()

  def _release(self) -> None:
    if getattr(self, "_h", None) is not None:
      self._lib.qtqp_sparse_free(self._h)
      self._h = None

  def __del__(self):
    try:
      self._release()
    except Exception:  # pylint: disable=broad-except
      pass


def sparse_analyze(kkt: sp.spmatrix, ordering: int = 0) -> tuple[np.ndarray, dict[str, Any]]:
  """Symbolic analysis of an upper-triangular KKT pattern on the host only
  (qtqp_sparse_analyze_ex): the elimination order and the factor's statistics.
  ordering: 0 exact minimum degree up to N = 40000 and approximate minimum
  degree above (what the backend uses), 1 exact, 2 approximate."""
  lib = _native.sparse_lib()
  csc = sp.csc_matrix(kkt)
  csc.sort_indices()
  ptr, ind = _native.as_i64(csc.indptr), _native.as_i64(csc.indices)
  perm = np.empty(csc.shape[0], dtype=np.int64)
  inf = _native.SparseInfo()
  rc = lib.qtqp_sparse_analyze_ex(csc.shape[0], _native.SPARSE_CSC, _native.iptr(ptr),
                                  _native.iptr(ind), int(ordering), _native.iptr(perm),
                                  _native.ctypes.byref(inf))
  _native.check(lib, rc)
  return perm, {k: getattr(inf, k) for k, _ in inf._fields_}


class DirectKktSolver:
  """Regularised quasi-definite KKT solves with iterative refinement.

      [ P + mu I      A'         ] [x]   [ rhs_x]
      [   A      -(D + mu I)     ] [y] = [-rhs_y],   D = diag([0]*z, s/y)

  Same constructor, ``update``/``solve``/``free`` semantics and stats as
  direct.py:112-311.  When the backend offers ``solve_refined`` (the HIP
  backends) the refinement loop runs on the device in one call.
  """

  def __init__(self, *, a, p, z, min_static_regularization, max_iterative_refinement_steps,
               atol, rtol, solver: LinearSolver):
    self.m, self.n = a.shape
    self.z = z
    self._p_diags = np.asarray(p.diagonal(), dtype=np.float64)
    self.min_static_regularization = min_static_regularization
    self.max_iterative_refinement_steps = max_iterative_refinement_steps
    self._atol, self._rtol = atol, rtol
    self._solver = solver
    solver.set_dims(n=self.n, m=self.m, z=self.z)
    self._device_ir = hasattr(solver, "solve_refined") and hasattr(solver, "set_dense_blocks")
    if self._device_ir:
      dense = lambda mat: mat.toarray() if sp.issparse(mat) else np.asarray(mat, dtype=np.float64)
      solver.set_dense_blocks(dense(a), dense(p))
      self._kkt = None
    else:
      # Upper-triangular scaffold [[triu(P)+I, A'], [., I]] (direct.py:166-175).
      self._kkt = sp.bmat(
          [[sp.triu(p, format="csc") + sp.eye(self.n, format="csc"), a.T],
           [None, sp.eye(self.m, format="csc")]],
          format=solver.format(), dtype=np.float64)
      solver.set_kkt(self._kkt)
    N = self.n + self.m
    self._true = np.empty(N)
    self._reg = np.empty(N)
    self._corr = np.zeros(N)

  def update(self, mu: float, s: np.ndarray, y: np.ndarray) -> None:
    """Regularised diagonal, factorisation, diagonal correction (direct.py:183-216)."""
    n, z = self.n, self.z
    t = self._true
    t[:n] = self._p_diags + mu
    t[n:n + z] = mu
    t[n + z:] = s[z:] / y[z:] + mu
    np.maximum(t, self.min_static_regularization, out=self._reg)
    t[n:] *= -1.0
    self._reg[n:] *= -1.0
    self._solver.update_diag(self._reg)
    self._solver.factorize()
    np.subtract(self._reg, t, out=self._corr)
    if self._device_ir:
      self._solver.set_correction(self._corr)

  def solve(self, rhs: np.ndarray, warm_start: np.ndarray) -> tuple[np.ndarray, dict[str, Any]]:
    """Refined solve of K_true sol = [rhs_x; -rhs_y] (direct.py:218-307)."""
    if self._device_ir:
      sol, st = self._solver.solve_refined(rhs, warm_start, self.max_iterative_refinement_steps,
                                           self._atol, self._rtol)
      logging.debug("KKT solve: status=%s, solves=%d, res=%e", st["status"], st["solves"],
                    st["final_residual_norm"])
      return sol, st
    n = self.n
    target = np.array(rhs, dtype=np.float64)
    target[n:] = -target[n:]
    rhs_norm = float(np.linalg.norm(target, np.inf))
    tol = self._atol + self._rtol * rhs_norm
    sol = np.array(warm_start, dtype=np.float64)

    def residual(v):
      return target - (self._solver @ v) + self._corr * v

    res = residual(sol)
    res_norm = float(np.linalg.norm(res, np.inf))
    status, count = "non-converged", 0
    for count in range(1, self.max_iterative_refinement_steps + 1):
      prev = res_norm
      sol += self._solver.solve(res)
      res = residual(sol)
      res_norm = float(np.linalg.norm(res, np.inf))
      if res_norm < tol:
        status = "converged"
        break
      if res_norm >= prev:
        status = "stalled"
        break
    if np.any(np.isnan(sol)):
      raise ValueError("Linear solver returned NaNs.")
    return sol, {"solves": count, "final_residual_norm": res_norm, "rhs_norm": rhs_norm,
                 "tolerance": tol, "converged": status == "converged", "status": status}

  def free(self) -> None:
    self._solver.free()
