"""qtqp_amd: qtqp's interior-point QP solver with an MI355X-native KKT engine.

Public API mirrors qtqp v0.0.5 (``src/qtqp/__init__.py``): ``QTQP(a, b, c, z,
p).solve(**kwargs) -> Solution`` with the same keyword arguments, statuses,
certificate scaling and stats keys, and a ``LinearSolver`` enum whose values
are backend classes (``HIP_DENSE`` mixed precision, ``HIP_DENSE_FP64`` plain
FP64).  The per-iteration linear algebra runs in hand-written HIP kernels
through the C ABI in include/qtqp_kkt.h; ``qtqp_amd.batch`` solves thousands
of QPs per launch on the device.

Algorithm (homogeneous self-dual embedding, Mehrotra predictor-corrector):
each iteration normalises (x, y, tau, s), factors the KKT matrix once,
solves K^{-1}[c; b], takes a predictor (mu_target = 0) and a corrector
(mu_target = sigma mu, second-order term) Newton step with a quadratic (or
linearised) equation for tau, and checks termination on the unscaled data.
"""

from __future__ import annotations

import dataclasses
import enum
import logging
import math
import sys
import timeit
from typing import Any, Dict, List

import numpy as np
import scipy.sparse as sp

from . import _native, direct

__version__ = "0.0.5"
_HEADER = ("| iter |      pcost |      dcost |     pres |     dres |      gap |   infeas |"
           "       mu |  q, p, c |     time |")
_SEPARA = ("|------|------------|------------|----------|----------|----------|----------|"
           "----------|----------|----------|")
_EPS = 1e-15
_inf_norm = lambda v: float(np.linalg.norm(v, np.inf)) if np.size(v) else 0.0


class LinearSolver(enum.Enum):
  """KKT backends.  AUTO picks by problem: the sparse backend for large sparse
  problems and for sizes past the dense engine's limits, the dense mixed-precision
  backend otherwise; the first of its candidates that initialises on this machine."""

  AUTO = "auto"
  HIP_DENSE = direct.HipDenseSolver
  HIP_DENSE_FP64 = direct.HipDenseFp64Solver
  HIP_SPARSE = direct.HipSparseSolver


# (platform, backend name) -> why the backend could not initialise here; only
# availability is cached, the choice itself depends on each problem
_AUTO_SOLVER_CACHE: dict[tuple, str] = {}
_AUTO_UNAVAILABLE_ERRORS = (ImportError, OSError)

_DENSE_N_LIMIT = 46340          # largest n of the dense KKT engine (n*n entries of G in int32)
_DENSE_BYTES_LIMIT = 4 << 30    # A, P and G in FP64 past this go to the sparse backend
_SPARSE_MIN_CELLS = 1 << 22     # below this many cells of A and P the dense path is kept
_SPARSE_MAX_FILL = 0.05         # stored fraction of A and P below which a larger problem is sparse


def _instantiate_linear_solver(linear_solver: LinearSolver) -> direct.LinearSolver:
  if linear_solver is LinearSolver.AUTO:
    raise ValueError("AUTO must be resolved before instantiating a backend.")
  return linear_solver.value()


def _auto_linear_solver_order(m: int = 0, n: int = 0, nnz: int = 0) -> list[LinearSolver]:
  """AUTO's candidates, in order, for A of m x n with nnz stored entries in A and P.

  The sparse backend comes first when n is past the dense engine's limit, when the
  dense A, P and G would take more than _DENSE_BYTES_LIMIT, or when a problem of
  more than _SPARSE_MIN_CELLS cells stores less than _SPARSE_MAX_FILL of them (the
  reference's AUTO is sparse-first; the dense MFMA path is faster for the small and
  dense QPs it keeps)."""
  cells = m * n + n * n
  dense_bytes = 8 * (m * n + 2 * n * n)
  sparse_first = (n > _DENSE_N_LIMIT or dense_bytes > _DENSE_BYTES_LIMIT
                  or (cells > _SPARSE_MIN_CELLS and nnz < _SPARSE_MAX_FILL * cells))
  dense = [LinearSolver.HIP_DENSE, LinearSolver.HIP_DENSE_FP64]
  return [LinearSolver.HIP_SPARSE] + dense if sparse_first else dense + [LinearSolver.HIP_SPARSE]


class _ProductMat:
  """A matrix used only in products: a dense copy (BLAS) when at least a
  quarter of it is stored, the scipy.sparse matrix otherwise."""

  def __init__(self, mat):
    self.shape, self.nnz = mat.shape, mat.nnz
    cells = mat.shape[0] * mat.shape[1]
    self._m = mat.toarray() if cells and mat.nnz >= 0.25 * cells and cells <= (1 << 28) else mat

  def __matmul__(self, v):
    return self._m @ v

  def rmatvec(self, v):
    return self._m.T @ v


def _segment_max(vals: np.ndarray, indptr: np.ndarray) -> np.ndarray:
  """Max of vals over each compressed segment [indptr[k], indptr[k+1]); 0 if empty."""
  out = np.zeros(len(indptr) - 1)
  nonempty = np.diff(indptr) > 0
  if vals.size:
    out[nonempty] = np.maximum.reduceat(vals, indptr[:-1][nonempty])
  return out


def _resolve_linear_solver(linear_solver: LinearSolver, shape: tuple = (0, 0, 0)):
  """(enum, backend instance); mirrors __init__.py:107-126.  shape = (m, n, nnz of
  A and P) of the problem AUTO chooses for."""
  if linear_solver is not LinearSolver.AUTO:
    return linear_solver, _instantiate_linear_solver(linear_solver)
  errors = []
  for cand in _auto_linear_solver_order(*shape):
    known = _AUTO_SOLVER_CACHE.get((sys.platform, cand.name))
    if known is not None:
      errors.append(f"{cand.name}: {known}")
      continue
    try:
      backend = _instantiate_linear_solver(cand)
    except _AUTO_UNAVAILABLE_ERRORS as e:
      _AUTO_SOLVER_CACHE[(sys.platform, cand.name)] = str(e)
      errors.append(f"{cand.name}: {e}")
      logging.debug("AUTO skipped %s: %s", cand.name, e)
      continue
    return cand, backend
  raise RuntimeError("AUTO could not initialize any linear solver backend: " + "; ".join(errors))


class SolutionStatus(enum.Enum):
  SOLVED = "solved"
  INFEASIBLE = "infeasible"
  UNBOUNDED = "unbounded"
  HIT_MAX_ITER = "hit_max_iter"
  FAILED = "failed"
  UNFINISHED = "unfinished"


@dataclasses.dataclass(frozen=True)
class Solution:
  """x, y, s (solution or certificate), per-iteration stats and status."""
  x: np.ndarray
  y: np.ndarray
  s: np.ndarray
  stats: List[Dict[str, Any]]
  status: SolutionStatus


@dataclasses.dataclass(frozen=True)
class _PresolveState:
  keep: np.ndarray
  a_dropped: sp.csc_matrix
  b_dropped: np.ndarray


class QTQP:
  """Primal-dual interior point method for convex QPs.

    min 1/2 x'Px + c'x   s.t.  Ax + s = b,  s[:z] = 0,  s[z:] >= 0
    max -1/2 x'Px - b'y  s.t.  Px + A'y = -c,  y[z:] >= 0
  """

  def __init__(self, *, a: sp.csc_matrix, b: np.ndarray, c: np.ndarray, z: int,
               p: sp.csc_matrix | None = None):
    self.m, self.n = a.shape
    self.z = z
    if not sp.isspmatrix_csc(a):
      raise TypeError("Constraint matrix 'a' must be in CSC format.")
    self.a = a
    self.b = np.array(b, dtype=np.float64)
    if self.b.shape != (self.m,):
      raise ValueError(f"b must have shape ({self.m},), got {self.b.shape}")
    self.c = np.array(c, dtype=np.float64)
    if self.c.shape != (self.n,):
      raise ValueError(f"c must have shape ({self.n},), got {self.c.shape}")
    if not 0 <= self.z <= self.m:
      raise ValueError(f"Number of equality constraints z={self.z} must satisfy "
                       f"0 <= z <= m={self.m}")
    self._presolve()
    if self.z == self.m:
      raise ValueError("effective z == m after presolve; some rows may have been dropped")
    if p is None:
      self.p = sp.csc_matrix((self.n, self.n))
    else:
      if not sp.isspmatrix_csc(p):
        raise TypeError("QP matrix 'p' must be in CSC format.")
      if p.shape != (self.n, self.n):
        raise ValueError(f"p must have shape ({self.n}, {self.n}), got {p.shape}")
      if not np.all(np.isfinite(p.data)):
        raise ValueError("QP matrix 'p' must contain only finite values.")
      asym = p - p.T
      if asym.nnz and np.max(np.abs(asym.data), initial=0.0) > 1e-12:
        raise ValueError("QP matrix 'p' must be symmetric.")
      self.p = p

  # -- presolve (__init__.py:248-297) --------------------------------------
  def _presolve(self, inf_bound: float = 1e20):
    self._presolve_state = None
    if not np.all(np.isfinite(self.b[:self.z])):
      raise ValueError("Equality RHS entries in 'b' must be finite.")
    tail = self.b[self.z:]
    if np.any(np.isnan(tail) | np.isneginf(tail)):
      raise ValueError("Inequality RHS entries in 'b' must be finite, +inf, or >= inf_bound.")
    drop = np.zeros(self.m, dtype=bool)
    drop[self.z:] = tail >= inf_bound
    if not drop.any():
      return
    keep = ~drop
    self._presolve_state = _PresolveState(keep=keep, a_dropped=self.a[drop],
                                          b_dropped=self.b[drop])
    self.a = self.a[keep]
    self.b = self.b[keep]
    self.m = int(keep.sum())

  def _postsolve(self, y, s, y_dropped=0.0, s_dropped=np.nan):
    ps = self._presolve_state
    if ps is None:
      return y, s
    y_full = np.empty(ps.keep.shape[0], dtype=y.dtype)
    s_full = np.empty(ps.keep.shape[0], dtype=s.dtype)
    y_full[ps.keep], y_full[~ps.keep] = y, y_dropped
    s_full[ps.keep], s_full[~ps.keep] = s, s_dropped
    return y_full, s_full

  def _dropped_slack(self, x):
    ps = self._presolve_state
    return 0.0 if ps is None else ps.b_dropped - ps.a_dropped @ x

  def _init_variables(self):
    """Default start x=0, y[z:]=s[z:]=1, tau=1 (equilibrated). __init__.py:299-307."""
    x, y, s = np.zeros(self.n), np.zeros(self.m), np.zeros(self.m)
    y[self.z:] = 1.0
    s[self.z:] = 1.0
    if self.equilibrate:
      x, y, s = self._equilibrate_iterates(x, y, s)
    return x, y, s, 1.0, {}

  # -- solve ---------------------------------------------------------------
  def solve(self, *, atol: float = 1e-7, rtol: float = 1e-8, atol_infeas: float = 1e-8,
            rtol_infeas: float = 1e-9, max_iter: int = 100, step_size_scale: float = 0.99,
            min_static_regularization: float = 1e-8, max_iterative_refinement_steps: int = 20,
            linear_solver_atol: float = 1e-12, linear_solver_rtol: float = 1e-12,
            linear_solver: LinearSolver = LinearSolver.AUTO, verbose: bool = True,
            equilibrate: bool = True, collect_stats: bool = False) -> Solution:
    """Solves the QP; the backend is always freed, also on exceptions."""
    self._linear_solver = None
    try:
      return self._solve_impl(
          atol=atol, rtol=rtol, atol_infeas=atol_infeas, rtol_infeas=rtol_infeas,
          max_iter=max_iter, step_size_scale=step_size_scale,
          min_static_regularization=min_static_regularization,
          max_iterative_refinement_steps=max_iterative_refinement_steps,
          linear_solver_atol=linear_solver_atol, linear_solver_rtol=linear_solver_rtol,
          linear_solver=linear_solver, verbose=verbose, equilibrate=equilibrate,
          collect_stats=collect_stats)
    finally:
      if self._linear_solver is not None:
        self._linear_solver.free()
        self._linear_solver = None

  def _solve_impl(self, *, atol, rtol, atol_infeas, rtol_infeas, max_iter, step_size_scale,
                  min_static_regularization, max_iterative_refinement_steps,
                  linear_solver_atol, linear_solver_rtol, linear_solver, verbose,
                  equilibrate, collect_stats) -> Solution:
    assert atol >= 0 and rtol >= 0 and atol_infeas >= 0 and rtol_infeas >= 0
    assert max_iter > 0
    assert 0 < step_size_scale < 1
    assert min_static_regularization >= 0
    assert max_iterative_refinement_steps >= 1
    assert linear_solver_atol >= 0 and linear_solver_rtol >= 0

    resolved, backend = _resolve_linear_solver(
        linear_solver, (self.m, self.n, int(self.a.nnz) + int(self.p.nnz)))
    self.start_time = timeit.default_timer()
    self.atol, self.rtol = atol, rtol
    self.atol_infeas, self.rtol_infeas = atol_infeas, rtol_infeas
    self.verbose = verbose
    self.equilibrate = equilibrate
    if verbose:
      print(f"| QTQP v{__version__}: m={self.m}, n={self.n}, z={self.z}, nnz(A)={self.a.nnz},"
            f" nnz(P)={self.p.nnz}, linear_solver={resolved.name}")
    if getattr(backend, "device_driver", False):
      return self._solve_on_device(
          backend, max_iter=max_iter, collect_stats=collect_stats,
          opts=_native.IpmOpts(
              atol=atol, rtol=rtol, atol_infeas=atol_infeas, rtol_infeas=rtol_infeas,
              step_size_scale=step_size_scale,
              min_static_regularization=min_static_regularization,
              linear_solver_atol=linear_solver_atol, linear_solver_rtol=linear_solver_rtol,
              max_iterative_refinement_steps=max_iterative_refinement_steps,
              collect_stats=int(bool(collect_stats)),
              use_graph=int(bool(getattr(backend, "device_graph", True))),
              graph_steps=int(getattr(backend, "device_graph_steps", 0))))
    if equilibrate:
      a, p, b, c, self.d, self.e = self._equilibrate()
    else:
      a, p, b, c, self.d, self.e = self.a, self.p, self.b, self.c, None, None
    self.q = np.concatenate([c, b])
    self._a_op, self._p_op = _ProductMat(self.a), _ProductMat(self.p)
    p_op = _ProductMat(p)
    self._norm_b = _inf_norm(self.b)
    self._norm_c = _inf_norm(self.c)
    self._linear_solver = direct.DirectKktSolver(
        a=a, p=p, z=self.z, min_static_regularization=min_static_regularization,
        max_iterative_refinement_steps=max_iterative_refinement_steps,
        atol=linear_solver_atol, rtol=linear_solver_rtol, solver=backend)

    n, m, z = self.n, self.m, self.z
    stats: list = []
    self.kinv_q = np.zeros_like(self.q)
    x, y, s, tau, _ = self._init_variables()
    status = SolutionStatus.UNFINISHED
    self._log_header()
    anchor = np.empty(n + m)
    d_s = np.zeros(m)
    alpha = sigma = 0.0

    for self.it in range(max_iter):
      st: dict = {}
      x, y, tau, s = self._normalize(x, y, tau, s)
      mu = (y @ s) / (m - z)
      self._linear_solver.update(mu=mu, s=s, y=y)
      self.kinv_q, st["q_lin_sys_stats"] = self._linear_solver.solve(
          rhs=self.q, warm_start=self.kinv_q)

      # Predictor: mu_target = 0, anchored at the current point.
      anchor[:n], anchor[n:] = x, y
      x_p, y_p, tau_p, st["predictor_lin_sys_stats"] = self._newton_step(
          p=p_op, mu=mu, mu_target=0.0, r_anchor=anchor, tau_anchor=tau, x=x, y=y, s=s,
          tau=tau, correction=None)
      d_x_p, d_y_p, d_tau_p = x_p - x, y_p - y, tau_p - tau
      d_s[z:] = -y_p[z:] * s[z:] / y[z:]
      alpha_p = self._compute_step_size(y, s, d_y_p, d_s)
      sigma = self._compute_sigma(mu, x, y, tau, s, alpha_p, d_x_p, d_y_p, d_tau_p, d_s)

      # Corrector: centring target sigma*mu plus Mehrotra's second-order term.
      correction = -d_s[z:] * d_y_p[z:] / y[z:]
      anchor[:n], anchor[n:] = x_p, y_p
      x_c, y_c, tau_c, st["corrector_lin_sys_stats"] = self._newton_step(
          p=p_op, mu=mu, mu_target=sigma * mu, r_anchor=anchor, tau_anchor=tau_p, x=x, y=y,
          s=s, tau=tau, correction=correction)
      d_x, d_y, d_tau = x_c - x, y_c - y, tau_c - tau
      d_s[z:] = sigma * mu / y[z:] + correction - y_c[z:] * s[z:] / y[z:]
      alpha = self._compute_step_size(y, s, d_y, d_s)
      step = step_size_scale * alpha
      x += step * d_x
      y += step * d_y
      tau += step * d_tau
      s += step * d_s
      y[z:] = np.maximum(y[z:], 1e-30)
      s[z:] = np.maximum(s[z:], 1e-30)
      tau = max(tau, 1e-30)

      status = self._check_termination(x, y, tau, s, alpha, mu, sigma, st, collect_stats)
      self._log_iteration(st)
      if collect_stats:
        stats.append(st)
      if status != SolutionStatus.UNFINISHED:
        break
    else:
      status = SolutionStatus.HIT_MAX_ITER
      if collect_stats:
        stats[-1]["status"] = status

    if self.equilibrate:
      x, y, s = self._unequilibrate_iterates(x, y, s)
    return self._finish(x, y, s, tau, status, stats)

  def _solve_on_device(self, backend, *, max_iter, collect_stats, opts) -> Solution:
    """The iteration loop of _solve_impl (__init__.py:473-578) with the whole
    iteration on the GPU: equilibration, normalisation, KKT update and
    factor, the three refined solves, tau, sigma, step and termination run
    as kernels on data held in HBM; each iteration returns one record."""
    has_p = self.p.nnz > 0
    ipm = backend.device_ipm(self.a.toarray(), self.p.toarray() if has_p else None, has_p,
                             self.b, self.c, self.z, self.equilibrate)
    self._linear_solver = ipm
    stats: list = []
    status = SolutionStatus.UNFINISHED
    self._log_header()
    for self.it in range(max_iter):
      st = self._record_stats(ipm.iterate(opts), collect_stats)
      status = st["status"]
      self._log_iteration(st)
      if collect_stats:
        stats.append(st)
      if status != SolutionStatus.UNFINISHED:
        break
    else:
      status = SolutionStatus.HIT_MAX_ITER
      if collect_stats:
        stats[-1]["status"] = status
    x, y, s, tau = ipm.iterates()
    return self._finish(x, y, s, tau, status, stats)

  _DEVICE_STATUS = {0: SolutionStatus.UNFINISHED, 1: SolutionStatus.SOLVED,
                    2: SolutionStatus.INFEASIBLE, 3: SolutionStatus.UNBOUNDED}

  def _record_stats(self, rec, collect_stats) -> dict:
    """The stats dict of one iteration (__init__.py:479-566, 966-1008) from a
    device record."""
    def lin(i, tau_method=None):
      ls = rec.lin[i]
      status = direct._IR_STATUS[ls.status]
      d = {"solves": int(ls.solves), "final_residual_norm": float(ls.final_residual_norm),
           "rhs_norm": float(ls.rhs_norm), "tolerance": float(ls.tolerance),
           "converged": status == "converged", "status": status}
      if tau_method is not None:
        d["tau_method"] = "quadratic" if tau_method else "linearized"
      return d

    st = {"q_lin_sys_stats": lin(0), "predictor_lin_sys_stats": lin(1, rec.tau_method[0]),
          "corrector_lin_sys_stats": lin(2, rec.tau_method[1])}
    st.update({k: float(getattr(rec, k)) for k in (
        "ctx", "bty", "pcost", "dcost", "pres", "dres", "gap", "pinfeas", "dinfeas", "dinfeas_a",
        "dinfeas_p", "mu", "sigma", "alpha", "tau", "norm_x", "norm_y")})
    st.update({"iter": self.it, "status": self._DEVICE_STATUS[rec.status],
               "time": timeit.default_timer() - self.start_time,
               "prelrhs": float(rec.prelrhs), "drelrhs": float(rec.drelrhs)})
    if collect_stats:
      st.update({k: float(getattr(rec, k)) for k in (
          "complementarity", "norm_s", "max_sy", "min_sy", "std_sy", "max_s_over_y",
          "min_s_over_y", "mean_s_over_y", "std_s_over_y")})
    return st

  def _finish(self, x, y, s, tau, status, stats) -> Solution:
    """Certificate or tau scaling and postsolve of the unequilibrated iterate
    (__init__.py:580-614)."""
    if status is SolutionStatus.INFEASIBLE:
      self._log_footer("Primal infeasible / dual unbounded")
      x.fill(np.nan)
      s.fill(np.nan)
      y_cert, s = self._postsolve(y / abs(self.b @ y), s)
      return Solution(x, y_cert, s, stats, status)
    if status is SolutionStatus.UNBOUNDED:
      self._log_footer("Dual infeasible / primal unbounded")
      y.fill(np.nan)
      scale = abs(self.c @ x)
      x, s = x / scale, s / scale
      y, s = self._postsolve(y, s, y_dropped=np.nan)
      return Solution(x, y, s, stats, status)
    footer = {SolutionStatus.SOLVED: "Solved",
              SolutionStatus.HIT_MAX_ITER: "Hit maximum iterations",
              SolutionStatus.UNFINISHED: "Failed to converge"}
    if status not in footer:
      raise ValueError(f"Unknown convergence status: {status}")
    self._log_footer(footer[status])
    x, y, s = x / tau, y / tau, s / tau
    y, s = self._postsolve(y, s, s_dropped=self._dropped_slack(x))
    if status is SolutionStatus.UNFINISHED:
      status = SolutionStatus.FAILED
    return Solution(x, y, s, stats, status)

  # -- equilibration (__init__.py:616-669) ---------------------------------
  def _equilibrate(self, num_iters=10, min_scale=1e-3, max_scale=1e3):
    """Ruiz scaling: A <- D A E, P <- E P E, b <- D b, c <- E c (__init__.py:616-669).

    Same passes and the same floating-point products as the reference.  The
    inf-norms are segment maxima over CSC (columns) and a CSR copy of A
    (rows) that is scaled alongside, instead of format conversions inside
    every norm call; maxima are exact, so d and e are bitwise the same.
    """
    a, p = self.a.copy().tocsc(), self.p.copy().tocsc()
    ar = a.tocsr()
    ar_rows = np.repeat(np.arange(self.m), np.diff(ar.indptr))
    d, e = np.ones(self.m), np.ones(self.n)
    for _ in range(num_iters):
      row = _segment_max(np.abs(ar.data), ar.indptr)
      d_i = np.clip(1.0 / np.sqrt(np.where(row == 0.0, 1.0, row)), min_scale, max_scale)
      col = np.maximum(_segment_max(np.abs(a.data), a.indptr),
                       _segment_max(np.abs(p.data), p.indptr))
      e_i = np.clip(1.0 / np.sqrt(np.where(col == 0.0, 1.0, col)), min_scale, max_scale)
      a.data *= d_i[a.indices] * np.repeat(e_i, np.diff(a.indptr))
      ar.data *= d_i[ar_rows] * e_i[ar.indices]
      if p.nnz > 0:
        p.data *= e_i[p.indices] * np.repeat(e_i, np.diff(p.indptr))
      d *= d_i
      e *= e_i
    return a, p, self.b * d, self.c * e, d, e

  def _unequilibrate_iterates(self, x, y, s):
    return self.e * x, self.d * y, s / self.d

  def _equilibrate_iterates(self, x, y, s):
    return x / self.e, y / self.d, s * self.d

  # -- scalar pieces -------------------------------------------------------
  def _max_step_size(self, y: np.ndarray, delta_y: np.ndarray) -> float:
    """Largest alpha in [0, 1] with y + alpha*delta_y >= 0 (__init__.py:671-680)."""
    mask = delta_y < -_EPS
    if not mask.any():
      return 1.0
    return min(1.0, float(np.min(-y[mask] / delta_y[mask])))

  def _compute_step_size(self, y, s, d_y, d_s) -> float:
    z = self.z
    return min(self._max_step_size(s[z:], d_s[z:]), self._max_step_size(y[z:], d_y[z:]))

  def _compute_sigma(self, mu_curr, x, y, tau, s, alpha, d_x, d_y, d_tau, d_s) -> float:
    """Mehrotra's sigma = clip((mu_aff/mu)^3, 0, 1) (__init__.py:682-706)."""
    xa, ya = x + alpha * d_x, y + alpha * d_y
    ta, sa = tau + alpha * d_tau, s + alpha * d_s
    nrm2 = xa @ xa + ya @ ya + ta * ta
    k = self.m - self.z
    mu_aff = (k + 1) / max(_EPS * _EPS, nrm2) * (ya @ sa) / k
    base = mu_aff / max(_EPS, mu_curr)
    return np.clip(base * base * base, 0.0, 1.0)

  def _normalize(self, x, y, tau, s):
    """Scales (x, y, tau, s) so ||(x, y, tau)||^2 = m - z + 1 (__init__.py:859-880)."""
    scale = math.sqrt(self.m - self.z + 1) / max(_EPS, math.sqrt(x @ x + y @ y + tau * tau))
    x *= scale
    y *= scale
    s *= scale
    return x, y, tau * scale, s

  def _newton_step(self, *, p, mu, mu_target, r_anchor, tau_anchor, x, y, s, tau, correction):
    """[x+; y+] = K^{-1} r - K^{-1} q tau+, tau+ from the embedding's tau
    equation (__init__.py:708-757)."""
    n, z = self.n, self.z
    r = (mu - mu_target) * r_anchor
    tail = r[n + z:]
    if mu_target != 0.0:
      tail += mu_target / y[z:]
    tail += s[z:]
    if correction is not None:
      tail += correction
    kinv_r, lin = self._linear_solver.solve(rhs=r, warm_start=r_anchor)
    tau_plus = None
    if lin["converged"] or lin["final_residual_norm"] < 1e-7:
      try:
        tau_plus = self._solve_for_tau(p, kinv_r, mu, mu_target, (mu - mu_target) * tau_anchor)
        lin["tau_method"] = "quadratic"
      except ValueError:
        logging.debug("Primary tau solve failed; falling back to linearized.")
    if tau_plus is None:
      lin["tau_method"] = "linearized"
      tau_plus = self._solve_for_tau_linearized_fallback(p, kinv_r, mu, mu_target, x, y, tau,
                                                         tau_anchor)
    kinv_r -= self.kinv_q * tau_plus
    return kinv_r[:n], kinv_r[n:], tau_plus, lin

  def _solve_for_tau(self, p, kinv_r, mu, mu_target, r_tau) -> float:
    """Stable root of t_a tau^2 + t_b tau + t_c = 0 (__init__.py:759-817)."""
    n, q, kq = self.n, self.q, self.kinv_q
    t_a = mu + kq @ q
    t_b = -r_tau - kinv_r @ q
    t_c = -mu_target
    if p.nnz > 0:
      pr, pq = (p @ np.stack([kinv_r[:n], kq[:n]], axis=1)).T
      t_a -= kq[:n] @ pq
      t_b += kinv_r[:n] @ pq + kq[:n] @ pr
      t_c -= kinv_r[:n] @ pr
    if abs(t_a) < _EPS:
      if abs(t_b) < _EPS:
        raise ValueError(f"Degenerate tau equation: t_a={t_a}, t_b={t_b}, t_c={t_c}")
      root = -t_c / t_b
      if not np.isfinite(root) or root < -1e-10:
        raise ValueError(f"Invalid linear tau solution found: {root}")
      return max(0.0, root)
    disc = t_b * t_b - 4 * t_a * t_c
    if disc < -1e-9:
      raise ValueError(f"Negative discriminant: {disc}")
    sq = math.sqrt(max(0.0, disc))
    root = t_c / (-0.5 * (t_b + sq)) if t_b > 0 else (-0.5 * (t_b - sq)) / t_a
    if not np.isfinite(root) or root < -1e-10:
      raise ValueError(f"Invalid tau solution found: {root}")
    return max(0.0, root)

  def _solve_for_tau_linearized_fallback(self, p, kinv_r, mu, mu_target, x, y, tau_curr,
                                         tau_anchor) -> float:
    """First-order tau step with a [0.1, 10] trust region (__init__.py:819-857)."""
    n, q, kq = self.n, self.q, self.kinv_q
    px = p @ x if p.nnz > 0 else np.zeros(n)
    q_z = q[:n] @ x + q[n:] @ y
    x_px = x @ px
    q_kq = q @ kq
    px_kq = px @ kq[:n]
    q_rz = q @ kinv_r - tau_curr * q_kq - q_z
    px_rz = px @ kinv_r[:n] - tau_curr * px_kq - x_px
    g = (mu * tau_curr * tau_curr + (mu_target - mu) * tau_anchor * tau_curr
         - tau_curr * q_z - mu_target - x_px)
    num = g - tau_curr * q_rz - 2.0 * px_rz
    den = (2.0 * mu * tau_curr + (mu_target - mu) * tau_anchor - q_z + tau_curr * q_kq
           + 2.0 * px_kq)
    tau_new = tau_curr + (0.0 if abs(den) < 1e-16 else -num / den)
    if not np.isfinite(tau_new):
      logging.warning("Linearized tau fallback non-finite; using current tau.")
      return tau_curr
    return min(max(tau_new, 0.1 * tau_curr), 10.0 * tau_curr)

  # -- termination (__init__.py:888-1008) ----------------------------------
  def _check_termination(self, x, y, tau, s, alpha, mu, sigma, stats_i, collect_stats):
    if self.equilibrate:
      x, y, s = self._unequilibrate_iterates(x, y, s)
    inv_tau = 1.0 / max(tau, _EPS)
    ax = self._a_op @ x
    aty = self._a_op.rmatvec(y)
    if self._p_op.nnz == 0:
      px, xpx = np.zeros(self.n), 0.0
    else:
      px = self._p_op @ x
      xpx = x @ px
    ctx, bty = self.c @ x, self.b @ y
    pcost = (ctx + 0.5 * xpx * inv_tau) * inv_tau
    dcost = (-bty - 0.5 * xpx * inv_tau) * inv_tau
    pres = _inf_norm((ax + s) * inv_tau - self.b)
    dres = _inf_norm((px + aty) * inv_tau + self.c)
    gap = abs((ctx + bty + xpx * inv_tau) * inv_tau)
    norm_aty, norm_px = _inf_norm(aty), _inf_norm(px)
    dinfeas_a = _inf_norm(ax + s) / (abs(ctx) + _EPS)
    dinfeas_p = norm_px / (abs(ctx) + _EPS)
    dinfeas = max(dinfeas_a, dinfeas_p)
    pinfeas = norm_aty / (abs(bty) + _EPS)
    prelrhs = max(_inf_norm(ax) * inv_tau, _inf_norm(s) * inv_tau, self._norm_b)
    drelrhs = max(norm_px * inv_tau, norm_aty * inv_tau, self._norm_c)
    norm_x, norm_y = _inf_norm(x), _inf_norm(y)
    if (gap < self.atol + self.rtol * min(abs(pcost), abs(dcost))
        and pres < self.atol + self.rtol * prelrhs
        and dres < self.atol + self.rtol * drelrhs):
      status = SolutionStatus.SOLVED
    elif ctx < -_EPS and dinfeas < self.atol_infeas + self.rtol_infeas * norm_x / abs(ctx):
      status = SolutionStatus.UNBOUNDED
    elif bty < -_EPS and pinfeas < self.atol_infeas + self.rtol_infeas * norm_y / abs(bty):
      status = SolutionStatus.INFEASIBLE
    else:
      status = SolutionStatus.UNFINISHED
    stats_i.update({
        "iter": self.it, "ctx": ctx, "bty": bty, "pcost": pcost, "dcost": dcost,
        "pres": pres, "dres": dres, "gap": gap, "pinfeas": pinfeas, "dinfeas": dinfeas,
        "dinfeas_a": dinfeas_a, "dinfeas_p": dinfeas_p, "mu": mu, "sigma": sigma,
        "alpha": alpha, "tau": tau, "norm_x": norm_x, "norm_y": norm_y, "status": status,
        "time": timeit.default_timer() - self.start_time, "prelrhs": prelrhs,
        "drelrhs": drelrhs,
    })
    if collect_stats:
      stats_i["complementarity"] = abs((y @ s) * inv_tau * inv_tau)
      stats_i["norm_s"] = _inf_norm(s)
      if self.z < self.m:
        sy = s[self.z:] * y[self.z:]
        soy = s[self.z:] / np.maximum(_EPS, y[self.z:])
        stats_i.update({"max_sy": np.max(sy), "min_sy": np.min(sy), "std_sy": np.std(sy),
                        "max_s_over_y": np.max(soy), "min_s_over_y": np.min(soy),
                        "mean_s_over_y": np.mean(soy), "std_s_over_y": np.std(soy)})
    return status

  # -- logging (__init__.py:1010-1039) -------------------------------------
  def _log_header(self):
    if self.verbose:
      print(f"{_SEPARA}\n{_HEADER}\n{_SEPARA}")

  def _log_iteration(self, st):
    if not self.verbose:
      return
    fmt = lambda d: " *" if d.get("status") == "stalled" else f"{d.get('solves', 0):2}"
    solves = ",".join(fmt(st[k]) for k in ("q_lin_sys_stats", "predictor_lin_sys_stats",
                                           "corrector_lin_sys_stats"))
    print(f"| {st['iter']:>4} | {st['pcost']:>10.3e} | {st['dcost']:>10.3e} |"
          f" {st['pres']:>8.2e} | {st['dres']:>8.2e} | {st['gap']:>8.2e} |"
          f" {min(st['pinfeas'], st['dinfeas']):>8.2e} | {st['mu']:>8.2e} | {solves:>8} |"
          f" {st['time']:>8.2e} |")

  def _log_footer(self, message: str):
    if self.verbose:
      print(f"{_SEPARA}\n| {message}")
