"""Batched device IPM: many uneven dense QPs solved per launch on one MI355X.

``BatchQP(qps).solve(**kwargs)`` runs qtqp's algorithm (the same iteration as
``qtqp_amd.QTQP``) for every QP with every per-iteration step on the GPU
(include/qtqp_batch.h).  Presolve (qtqp __init__.py:248-270) and the
status-dependent scaling of the result (580-614) run on the host here.
"""

from __future__ import annotations

import ctypes
import dataclasses
from typing import Sequence

import numpy as np

from . import _native
from . import SolutionStatus

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_dp = ctypes.POINTER(ctypes.c_double)


class BatchOpts(ctypes.Structure):
  _fields_ = [("atol", ctypes.c_double), ("rtol", ctypes.c_double),
              ("atol_infeas", ctypes.c_double), ("rtol_infeas", ctypes.c_double),
              ("step_size_scale", ctypes.c_double),
              ("min_static_regularization", ctypes.c_double),
              ("linear_solver_atol", ctypes.c_double), ("linear_solver_rtol", ctypes.c_double),
              ("max_iter", ctypes.c_int32), ("max_iterative_refinement_steps", ctypes.c_int32),
              ("precision", ctypes.c_int32), ("use_graph", ctypes.c_int32),
              ("equilibrate", ctypes.c_int32), ("per_iteration", ctypes.c_int32)]


class BatchCounters(ctypes.Structure):
  _fields_ = [("backend_solves", ctypes.c_int64), ("inner_steps", ctypes.c_int64),
              ("fp64_fallbacks", ctypes.c_int64), ("ir_not_converged", ctypes.c_int64),
              ("iterations", ctypes.c_int32), ("device_ms", ctypes.c_double),
              ("jitter_retries", ctypes.c_int64)]


BATCH_SYMBOLS = {
    "qtqp_batch_create": (ctypes.c_int, [ctypes.c_int32, _i32p, _i32p, _i32p, _i64p, _i64p, _i64p,
                                         _i64p, _dp, _dp, _dp, _dp,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_batch_create_ex": (ctypes.c_int, [ctypes.c_int32, _i32p, _i32p, _i32p, _i64p, _i64p,
                                            _i64p, _i64p, _dp, _dp, _dp, _dp, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_batch_create_family": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                ctypes.c_int32, _dp, _dp, _dp, _dp, ctypes.c_int32,
                                                ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_batch_set_bc": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "qtqp_batch_stored_rows": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
    "qtqp_batch_row_bytes": (ctypes.c_int, [ctypes.c_void_p, _i64p]),
    "qtqp_batch_set_start": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp]),
    "qtqp_batch_set_start_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p]),
    "qtqp_batch_device_buffers": (ctypes.c_int, [ctypes.c_void_p] +
                                  [ctypes.POINTER(ctypes.c_void_p)] * 4),
    "qtqp_batch_solution": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _i32p, _i32p]),
    "qtqp_batch_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchOpts),
                                        ctypes.POINTER(BatchCounters)]),
    "qtqp_batch_results": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp, _dp, _dp, _i32p, _i32p,
                                          _dp, _dp, _dp]),
    "qtqp_batch_launch_times": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int32]),
    "qtqp_batch_instance_ms": (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int32]),
    "qtqp_batch_work": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
    "qtqp_batch_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.c_int32]),
    "qtqp_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchOpts),
                                           ctypes.c_int32, ctypes.c_int32, _i32p, _dp]),
    "qtqp_batch_set_tiles": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
    "qtqp_batch_free": (None, [ctypes.c_void_p]),
    "qtqp_batch_last_error": (ctypes.c_char_p, []),
}

_STATUS = {1: SolutionStatus.SOLVED, 2: SolutionStatus.INFEASIBLE,
           3: SolutionStatus.UNBOUNDED}


def batch_lib() -> ctypes.CDLL:
  lib = _native.load("libqtqp_batch.so", BATCH_SYMBOLS)
  _native.require_device(_native.kkt_lib())
  return lib


def _check(lib, rc):
  if rc != 0:
    msg = (lib.qtqp_batch_last_error() or b"").decode(errors="replace")
    if rc == _native.QTQP_ERR_ARG:
      raise ValueError(msg)
    raise RuntimeError(f"qtqp_amd batch error {rc}: {msg}")


@dataclasses.dataclass
class BatchSolution:
  x: np.ndarray
  y: np.ndarray
  s: np.ndarray
  status: SolutionStatus
  iters: int
  pres: float
  dres: float
  gap: float


@dataclasses.dataclass
class BatchResult:
  solutions: list
  device_ms: float
  counters: dict

  @property
  def statuses(self):
    return [s.status for s in self.solutions]

  @property
  def iterations(self) -> np.ndarray:
    return np.array([s.iters for s in self.solutions])


def _ptr(a, t):
  return a.ctypes.data_as(t)


# -- packed batch-instance files ----------------------------------------------
# One .npz archive (no pickled objects): per-instance sizes n, m, z, element
# offsets of each QP's dense row-major A (m x n) and P (n x n) and of its b, c
# in the concatenated arrays, optional generator seeds (-1 = none), and a
# format tag.  The same packing the C ABI takes (qtqp_batch_create).
_INSTANCE_FORMAT = "qtqp_amd.batch-instances.v1"


def save_instances(path, qps: Sequence, seeds=None) -> None:
  """Writes dense QPs ``(a, b, c, z, p)`` (or objects with those fields)."""
  ns, ms, zs, a_parts, p_parts, b_parts, c_parts = [], [], [], [], [], [], []
  for q in qps:
    a, b, c, z, p = (q.a, q.b, q.c, q.z, q.p) if hasattr(q, "a") else q
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    p = np.zeros((n, n)) if p is None else np.asarray(p, dtype=np.float64)
    ns.append(n)
    ms.append(m)
    zs.append(int(z))
    a_parts.append(a.ravel())
    p_parts.append(p.ravel())
    b_parts.append(np.asarray(b, dtype=np.float64))
    c_parts.append(np.asarray(c, dtype=np.float64))
  off = lambda parts: np.concatenate([[0], np.cumsum([x.size for x in parts])[:-1]]).astype(np.int64)
  seeds = np.full(len(ns), -1, dtype=np.int64) if seeds is None else np.asarray(seeds, np.int64)
  if seeds.shape != (len(ns),):
    raise ValueError("one seed per instance")
  np.savez(path, format=np.array(_INSTANCE_FORMAT), n=np.array(ns, np.int32),
           m=np.array(ms, np.int32), z=np.array(zs, np.int32), a_off=off(a_parts),
           p_off=off(p_parts), b_off=off(b_parts), c_off=off(c_parts),
           a=np.concatenate(a_parts), p=np.concatenate(p_parts), b=np.concatenate(b_parts),
           c=np.concatenate(c_parts), seeds=seeds)


def load_instances(path) -> tuple[list, np.ndarray]:
  """Reads a file written by save_instances: ([(a, b, c, z, p), ...], seeds)."""
  with np.load(path, allow_pickle=False) as f:
    if str(f["format"]) != _INSTANCE_FORMAT:
      raise ValueError(f"not a {_INSTANCE_FORMAT} file")
    n, m, z = f["n"], f["m"], f["z"]
    a, p, b, c = f["a"], f["p"], f["b"], f["c"]
    ao, po, bo, co = f["a_off"], f["p_off"], f["b_off"], f["c_off"]
    qps = []
    for i in range(len(n)):
      ni, mi = int(n[i]), int(m[i])
      qps.append((a[ao[i]:ao[i] + mi * ni].reshape(mi, ni).copy(), b[bo[i]:bo[i] + mi].copy(),
                  c[co[i]:co[i] + ni].copy(), int(z[i]),
                  p[po[i]:po[i] + ni * ni].reshape(ni, ni).copy()))
    return qps, f["seeds"].copy()


class BatchQP:
  """A batch of dense QPs ``(a, b, c, z, p)`` resident on the GPU.

  Each QP follows qtqp's conventions: ``min 1/2 x'Px + c'x  s.t.
  Ax + s = b, s[:z] = 0, s[z:] >= 0``; ``p`` may be None (LP).
  With ``pair_rows`` (default) rows of A that are exact negations of another
  row are stored once on the device (``stored_rows`` gives the count); with
  ``row_extents`` (default) the passes over A read only each row's nonzero
  column extent (same results as full rows).  With ``fused_solves``
  (default) a QP that factors in FP64 (n <= 256) refines each iteration's
  K^{-1} [c; b] and predictor solves together, one pass over A and P
  serving both (same results, bit for bit, as solving them in turn).
  """

  NO_ROW_PAIRS = 1
  NO_ROW_EXTENTS = 2
  NO_FUSED_SOLVES = 4

  @classmethod
  def _flags(cls, pair_rows, row_extents, fused_solves):
    return ((0 if pair_rows else cls.NO_ROW_PAIRS) | (0 if row_extents else cls.NO_ROW_EXTENTS)
            | (0 if fused_solves else cls.NO_FUSED_SOLVES))

  def __init__(self, qps: Sequence, pair_rows: bool = True, row_extents: bool = True,
               fused_solves: bool = True):
    self._lib = batch_lib()
    self._h = None
    ns, ms, zs, a_parts, p_parts, b_parts, c_parts = [], [], [], [], [], [], []
    self._presolve = []
    for q in qps:
      a, b, c, z, p = (q.a, q.b, q.c, q.z, q.p) if hasattr(q, "a") else q
      a = np.asarray(a, dtype=np.float64)
      m, n = a.shape
      b = np.array(b, dtype=np.float64)
      c = np.asarray(c, dtype=np.float64)
      p = np.zeros((n, n)) if p is None else np.asarray(p, dtype=np.float64)
      if b.shape != (m,) or c.shape != (n,) or p.shape != (n, n) or not 0 <= z <= m:
        raise ValueError("inconsistent QP shapes")
      if not np.all(np.isfinite(b[:z])):
        raise ValueError("Equality RHS entries in 'b' must be finite.")
      if np.any(np.isnan(b[z:]) | np.isneginf(b[z:])):
        raise ValueError("Inequality RHS entries in 'b' must be finite, +inf, or >= inf_bound.")
      keep = np.ones(m, dtype=bool)
      keep[z:] = b[z:] < 1e20
      self._presolve.append((keep, a[~keep], b[~keep]) if not keep.all() else None)
      a, b = a[keep], b[keep]
      if z == a.shape[0]:
        raise ValueError("effective z == m after presolve; some rows may have been dropped")
      ns.append(n)
      ms.append(a.shape[0])
      zs.append(z)
      a_parts.append(a.ravel())
      p_parts.append(p.ravel())
      b_parts.append(b)
      c_parts.append(c)
    self.count = len(ns)
    self.n = np.array(ns, dtype=np.int32)
    self.m = np.array(ms, dtype=np.int32)
    self.z = np.array(zs, dtype=np.int32)
    off = lambda sizes: np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    self.a_off = off([x.size for x in a_parts])
    self.p_off = off([x.size for x in p_parts])
    self.n_off = off(ns)
    self.m_off = off(ms)
    self._b = [b for b in b_parts]
    self._c = [c for c in c_parts]
    a_all = np.concatenate(a_parts)
    p_all = np.concatenate(p_parts)
    b_all = np.concatenate(b_parts)
    c_all = np.concatenate(c_parts)
    h = ctypes.c_void_p()
    rc = self._lib.qtqp_batch_create_ex(
        self.count, _ptr(self.n, _i32p), _ptr(self.m, _i32p), _ptr(self.z, _i32p),
        _ptr(self.a_off, _i64p), _ptr(self.p_off, _i64p), _ptr(self.n_off, _i64p),
        _ptr(self.m_off, _i64p), _ptr(a_all, _dp), _ptr(p_all, _dp), _ptr(b_all, _dp),
        _ptr(c_all, _dp), self._flags(pair_rows, row_extents, fused_solves), ctypes.byref(h))
    _check(self._lib, rc)
    self._h = h

  @classmethod
  def family(cls, a, p, z: int, b, c, pair_rows: bool = True, row_extents: bool = True,
             fused_solves: bool = True) -> "BatchQP":
    """count QPs of one parametric family: the same A (m x n), P (n x n or
    None) and z, with b of shape (count, m) and c of shape (count, n).  A and
    P are held once on the device and every QP reads them (results as for
    ``BatchQP`` over ``count`` copies).  Rows with b >= 1e20 are dropped by
    the presolve as in ``BatchQP``; they must be the same rows in every QP."""
    self = cls.__new__(cls)
    self._lib = batch_lib()
    self._h = None
    a = np.asarray(a, dtype=np.float64)
    m, n = a.shape
    p = np.zeros((n, n)) if p is None else np.asarray(p, dtype=np.float64)
    if not 0 <= z <= m or p.shape != (n, n):
      raise ValueError("inconsistent QP shapes")
    self._a_family = a
    self.count = count = int(np.shape(b)[0]) if np.ndim(b) == 2 else 1
    self.z = np.full(count, z, dtype=np.int32)
    self.n = np.full(count, n, dtype=np.int32)
    bk, ck, keep, self._b_drop = self._family_vectors(b, c, m)
    a_k = np.ascontiguousarray(a[keep])
    if z == a_k.shape[0]:
      raise ValueError("effective z == m after presolve; some rows may have been dropped")
    mk = a_k.shape[0]
    self._keep = keep
    self.m = np.full(count, mk, dtype=np.int32)
    self.n_off = np.arange(count, dtype=np.int64) * n
    self.m_off = np.arange(count, dtype=np.int64) * mk
    h = ctypes.c_void_p()
    rc = self._lib.qtqp_batch_create_family(
        count, n, mk, z, _ptr(a_k, _dp), _ptr(p, _dp), _ptr(bk, _dp), _ptr(ck, _dp),
        cls._flags(pair_rows, row_extents, fused_solves), ctypes.byref(h))
    _check(self._lib, rc)
    self._h = h
    return self

  @staticmethod
  def _family_keep(b, z):
    """Rows every QP keeps after the presolve (the same in all, or an error)."""
    if z and not np.all(np.isfinite(b[:, :z])):
      raise ValueError("Equality RHS entries in 'b' must be finite.")
    # per-row minimum and maximum over the QPs: NaN and -inf show in them,
    # and a row is dropped in every QP or in none iff both sides agree
    lo, hi = b[:, z:].min(axis=0), b[:, z:].max(axis=0)
    if np.any(np.isnan(lo) | np.isnan(hi) | np.isneginf(lo)):
      raise ValueError("Inequality RHS entries in 'b' must be finite, +inf, or >= inf_bound.")
    if np.any((lo >= 1e20) != (hi >= 1e20)):
      raise ValueError("a family batch needs the same rows with b >= 1e20 in every QP")
    keep = np.ones(b.shape[1], dtype=bool)
    keep[z:] = ~(lo >= 1e20)
    return keep

  def _family_vectors(self, b, c, m):
    """Checks a family batch's b (count, m) and c (count, n) and returns the
    contiguous presolved b and c to upload, the kept rows and the b of the
    dropped rows (None if none), the only part of b kept on the host (for
    the dropped rows' slacks)."""
    count, n, z = self.count, int(self.n[0]), int(self.z[0])
    b = np.asarray(b, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    if b.shape != (count, m) or c.shape != (count, n):
      raise ValueError("inconsistent QP shapes")
    keep = self._family_keep(b, z)
    if keep.all():
      return np.ascontiguousarray(b), np.ascontiguousarray(c), keep, None
    return np.ascontiguousarray(b[:, keep]), np.ascontiguousarray(c), keep, b[:, ~keep].copy()

  def set_bc(self, b, c) -> None:
    """New right-hand sides for every QP (the next solve re-solves with them):
    per-QP lists, or (count, m) / (count, n) arrays for a family batch.  The
    presolve must drop the same rows as at creation."""
    if getattr(self, "_a_family", None) is not None:
      bk, ck, keep, b_drop = self._family_vectors(b, c, self._keep.size)
      if not np.array_equal(keep, self._keep):
        raise ValueError("the presolve would drop other rows than at creation")
      self._b_drop = b_drop
    else:
      if len(b) != self.count or len(c) != self.count:
        raise ValueError("one b and one c per QP")
      bk, ck = [], []
      for q in range(self.count):
        bq = np.array(b[q], dtype=np.float64)
        cq = np.asarray(c[q], dtype=np.float64)
        ps = self._presolve[q]
        keep = np.ones(int(self.m[q]), dtype=bool) if ps is None else ps[0]
        zq = int(self.z[q])
        if (bq.shape != keep.shape or cq.shape != (int(self.n[q]),)
            or not np.all(np.isfinite(bq[:zq]))):
          raise ValueError(f"QP {q}: inconsistent b or c")
        newkeep = np.ones(bq.size, dtype=bool)
        newkeep[zq:] = bq[zq:] < 1e20
        if not np.array_equal(newkeep, keep):
          raise ValueError(f"QP {q}: the presolve would drop other rows than at creation")
        if ps is not None:
          self._presolve[q] = (keep, ps[1], bq[~keep])
        self._b[q], self._c[q] = bq[keep], cq
        bk.append(bq[keep])
        ck.append(cq)
      bk, ck = np.concatenate(bk), np.concatenate(ck)
    _check(self._lib, self._lib.qtqp_batch_set_bc(self._h, _ptr(bk, _dp), _ptr(ck, _dp)))

  def row_bytes(self) -> np.ndarray:
    """FP64 bytes one pass over each QP's stored row extents reads."""
    out = np.zeros(self.count, dtype=np.int64)
    _check(self._lib, self._lib.qtqp_batch_row_bytes(self._h, _ptr(out, _i64p)))
    return out

  def stored_rows(self) -> np.ndarray:
    """Rows of A each QP keeps on the device (m less its row pairs)."""
    out = np.zeros(self.count, dtype=np.int32)
    _check(self._lib, self._lib.qtqp_batch_stored_rows(self._h, _ptr(out, _i32p)))
    return out

  def set_start(self, x, y, s, tau) -> None:
    """Starting point of the next solve (unscaled coordinates), packed or lists."""
    pack = lambda v: np.ascontiguousarray(np.concatenate(v) if isinstance(v, (list, tuple)) else v,
                                          dtype=np.float64)
    x, y, s, t = pack(x), pack(y), pack(s), np.ascontiguousarray(tau, dtype=np.float64)
    if x.size != self.n.sum() or y.size != self.m.sum() or s.size != self.m.sum() or t.size != self.count:
      raise ValueError("start has the wrong packed sizes")
    _check(self._lib, self._lib.qtqp_batch_set_start(self._h, _ptr(x, _dp), _ptr(y, _dp),
                                                     _ptr(s, _dp), _ptr(t, _dp)))

  def device_buffers(self) -> tuple:
    """Device addresses of the handle's packed x, y, s and tau buffers (a
    producer on the device may fill them for ``set_start_device``)."""
    ptrs = [ctypes.c_void_p() for _ in range(4)]
    _check(self._lib, self._lib.qtqp_batch_device_buffers(self._h, *map(ctypes.byref, ptrs)))
    return tuple(p.value for p in ptrs)

  def set_start_device(self, x, y, s, tau) -> None:
    """``set_start`` from device addresses (packed as ``set_start`` takes them)."""
    _check(self._lib, self._lib.qtqp_batch_set_start_device(self._h, x, y, s, tau))

  @staticmethod
  def _opts(*, atol=1e-7, rtol=1e-8, atol_infeas=1e-8, rtol_infeas=1e-9, max_iter=100,
            step_size_scale=0.99, min_static_regularization=1e-8,
            max_iterative_refinement_steps=20, linear_solver_atol=1e-12,
            linear_solver_rtol=1e-12, equilibrate=True, precision="mixed",
            use_graph=True, per_iteration=False) -> BatchOpts:
    assert atol >= 0 and rtol >= 0 and atol_infeas >= 0 and rtol_infeas >= 0
    assert max_iter > 0 and 0 < step_size_scale < 1 and max_iterative_refinement_steps >= 1
    prec = {"mixed": _native.PREC_MIXED, "fp64": _native.PREC_FP64}[precision]
    return BatchOpts(atol, rtol, atol_infeas, rtol_infeas, step_size_scale,
                     min_static_regularization, linear_solver_atol, linear_solver_rtol, max_iter,
                     max_iterative_refinement_steps, prec, int(bool(use_graph)),
                     int(bool(equilibrate)), int(bool(per_iteration)))

  # Gram tile shapes and autotuner size buckets (include/qtqp_batch.h).
  TILE_SHAPES = ("64x64 x2", "128x128", "64x64 x4", "64x64 x1")
  BUCKET_EDGES = (64, 128, 192, 256, 384, 512, 768, 1024)

  def autotune(self, *, sample: int = 32, iters: int = 3, **opts) -> dict:
    """Times each Gram tile shape, in FP32 and in FP64, on up to `sample`
    QPs of every size bucket (`iters` Gram products each) and applies the
    fastest per bucket and precision.  Returns {"fp32": {"table": shape
    index per bucket (-1: empty), "ms": bucket x shape timings}, "fp64":
    {...}, "buckets": upper n edges, "shapes": names}.  Discards a start set
    with set_start."""
    o = self._opts(**opts)
    nb, nt = len(self.BUCKET_EDGES), len(self.TILE_SHAPES)
    table = np.zeros(2 * nb, dtype=np.int32)
    ms = np.zeros(2 * nb * nt)
    _check(self._lib, self._lib.qtqp_batch_autotune(self._h, ctypes.byref(o), int(sample),
                                                    int(iters), _ptr(table, _i32p),
                                                    _ptr(ms, _dp)))
    ms = ms.reshape(2, nb, nt)
    rows = lambda k: [[float(v) if np.isfinite(v) else None for v in row] for row in ms[k]]
    return {"fp32": {"table": table[:nb].tolist(), "ms": rows(0)},
            "fp64": {"table": table[nb:].tolist(), "ms": rows(1)},
            "buckets": list(self.BUCKET_EDGES), "shapes": list(self.TILE_SHAPES)}

  def set_tiles(self, table32, table64=None) -> None:
    """Applies tile tables (one shape index per size bucket) for the FP32 and
    the FP64 Gram product, e.g. saved autotune results (FP64 default: two
    64 x 64 tiles per pass)."""
    nb = len(self.BUCKET_EDGES)
    t64 = np.zeros(nb, dtype=np.int32) if table64 is None else table64
    t = np.ascontiguousarray(np.concatenate([table32, t64]), dtype=np.int32)
    assert t.shape == (2 * nb,)
    _check(self._lib, self._lib.qtqp_batch_set_tiles(self._h, _ptr(t, _i32p)))

  def solve_device(self, **opts) -> dict:
    """Runs the device IPM only (results stay on the GPU); returns counters.
    Keyword options as QTQP.solve, plus precision ("mixed" | "fp64"),
    use_graph, and per_iteration (one launch per IPM iteration instead of one
    persistent launch for the whole solve)."""
    o = self._opts(**opts)
    max_iter = o.max_iter
    cnt = BatchCounters()
    _check(self._lib, self._lib.qtqp_batch_solve(self._h, ctypes.byref(o), ctypes.byref(cnt)))
    self._max_iter = max_iter
    return {k: getattr(cnt, k) for k, _ in cnt._fields_}

  def raw_iterates(self) -> dict:
    """Final iterates of the last solve in unscaled coordinates, not divided by
    tau (x = e x_eq, y = d y_eq, s = s_eq / d), with tau, status code and
    iteration count per QP; x/y/s packed by n_off / m_off."""
    tn, tm = int(self.n.sum()), int(self.m.sum())
    x, y, s = np.empty(tn), np.empty(tm), np.empty(tm)
    tau = np.empty(self.count)
    st = np.empty(self.count, dtype=np.int32)
    its = np.empty(self.count, dtype=np.int32)
    _check(self._lib, self._lib.qtqp_batch_results(
        self._h, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(tau, _dp), _ptr(st, _i32p),
        _ptr(its, _i32p), None, None, None))
    return {"x": x, "y": y, "s": s, "tau": tau, "status": st, "iters": its}

  def results(self) -> list:
    if getattr(self, "_a_family", None) is not None:
      ra = self.results_arrays()
      k = self.count
      tau, pres, dres, gap = (np.empty(k) for _ in range(4))
      _check(self._lib, self._lib.qtqp_batch_results(
          self._h, None, None, None, _ptr(tau, _dp), None, None, _ptr(pres, _dp),
          _ptr(dres, _dp), _ptr(gap, _dp)))
      return [BatchSolution(ra["x"][q], ra["y"][q], ra["s"][q],
                            _STATUS.get(int(ra["status"][q]), SolutionStatus.HIT_MAX_ITER),
                            int(ra["iters"][q]), float(pres[q]), float(dres[q]), float(gap[q]))
              for q in range(k)]
    tn, tm = int(self.n.sum()), int(self.m.sum())
    x, y, s = np.empty(tn), np.empty(tm), np.empty(tm)
    tau, pres, dres, gap = (np.empty(self.count) for _ in range(4))
    st = np.empty(self.count, dtype=np.int32)
    its = np.empty(self.count, dtype=np.int32)
    _check(self._lib, self._lib.qtqp_batch_results(
        self._h, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(tau, _dp), _ptr(st, _i32p),
        _ptr(its, _i32p), _ptr(pres, _dp), _ptr(dres, _dp), _ptr(gap, _dp)))
    out = []
    for q in range(self.count):
      n, m = int(self.n[q]), int(self.m[q])
      xq = x[self.n_off[q]:self.n_off[q] + n].copy()
      yq = y[self.m_off[q]:self.m_off[q] + m].copy()
      sq = s[self.m_off[q]:self.m_off[q] + m].copy()
      code = int(st[q])
      if code == 5:
        raise np.linalg.LinAlgError(f"QP {q}: FP64 factorisation broke down")
      status = _STATUS.get(code, SolutionStatus.HIT_MAX_ITER)
      xq, yq, sq = self._finish(q, status, xq, yq, sq, float(tau[q]))
      out.append(BatchSolution(xq, yq, sq, status, int(its[q]), float(pres[q]), float(dres[q]),
                               float(gap[q])))
    return out

  def results_arrays(self) -> dict:
    """The solutions of ``results()`` for a batch whose QPs all have the same
    n and m (a family batch), as stacked arrays: x (count, n), y and s
    (count, m), status codes (QTQP_BATCH_*, 0 for max_iter) and iterations.
    The status-dependent scaling runs on the device; the presolve's dropped
    rows are restored here for all QPs at once."""
    n, m = int(self.n[0]), int(self.m[0])
    if np.any(self.n != n) or np.any(self.m != m):
      raise ValueError("results_arrays needs QPs of equal n and m")
    if getattr(self, "_a_family", None) is None and any(ps is not None for ps in self._presolve):
      # per-QP presolves may drop different rows: finish QP by QP
      sols = self.results()
      out = {k: np.stack([getattr(q, k) for q in sols]) for k in ("x", "y", "s")}
      inv = {v: k for k, v in _STATUS.items()}
      out["status"] = np.array([inv.get(q.status, 0) for q in sols], dtype=np.int32)
      out["iters"] = np.array([q.iters for q in sols], dtype=np.int32)
      return out
    k = self.count
    x, y, s = np.empty((k, n)), np.empty((k, m)), np.empty((k, m))
    st = np.empty(k, dtype=np.int32)
    its = np.empty(k, dtype=np.int32)
    # the status-dependent scaling runs on the device (qtqp_batch_solution)
    _check(self._lib, self._lib.qtqp_batch_solution(
        self._h, _ptr(x, _dp), _ptr(y, _dp), _ptr(s, _dp), _ptr(st, _i32p), _ptr(its, _i32p)))
    if np.any(st == 5):
      raise np.linalg.LinAlgError(f"QP {int(np.argmax(st == 5))}: FP64 factorisation broke down")
    inf = st == 2
    unb = st == 3
    fin = ~(inf | unb)
    if getattr(self, "_b_drop", None) is not None:
      keep = self._keep
      a_drop = self._a_family[~keep]
      b_drop = self._b_drop
      yf = np.empty((k, keep.size))
      sf = np.empty((k, keep.size))
      yf[:, keep], sf[:, keep] = y, s
      yf[:, ~keep] = np.where(unb[:, None], np.nan, 0.0)
      sf[:, ~keep] = np.where(fin[:, None], b_drop - x @ a_drop.T, np.nan)
      y, s = yf, sf
    return {"x": x, "y": y, "s": s, "status": st, "iters": its}

  def _finish(self, q, status, x, y, s, tau):
    """Status-dependent scaling and presolve restore (__init__.py:580-614, 272-297)."""
    ps = self._presolve[q]
    def restore(yv, sv, y_drop, s_drop):
      if ps is None:
        return yv, sv
      keep = ps[0]
      yf = np.empty(keep.size)
      sf = np.empty(keep.size)
      yf[keep], yf[~keep] = yv, y_drop
      sf[keep], sf[~keep] = sv, s_drop
      return yf, sf
    if status is SolutionStatus.INFEASIBLE:
      x[:] = np.nan
      s[:] = np.nan
      y = y / abs(self._b[q] @ y)
      y, s = restore(y, s, 0.0, np.nan)
      return x, y, s
    if status is SolutionStatus.UNBOUNDED:
      y[:] = np.nan
      k = abs(self._c[q] @ x)
      x, s = x / k, s / k
      y, s = restore(y, s, np.nan, np.nan)
      return x, y, s
    x, y, s = x / tau, y / tau, s / tau
    drop_s = 0.0 if ps is None else ps[2] - ps[1] @ x
    y, s = restore(y, s, 0.0, drop_s)
    return x, y, s

  def instance_ms(self) -> np.ndarray:
    """Completion time of each QP in the last solve: device wall-clock ms from
    the start of the first IPM launch to the end of its last iteration."""
    out = np.full(self.count, np.nan)
    k = self._lib.qtqp_batch_instance_ms(self._h, _ptr(out, _dp), self.count)
    if k < 0:
      raise RuntimeError("qtqp_batch_instance_ms failed")
    return out

  def launch_times_ms(self) -> np.ndarray:
    """HIP-event durations of the IPM launches of the last solve (one for the
    persistent schedule, one per iteration with per_iteration)."""
    k = self._lib.qtqp_batch_launch_times(self._h, None, 0)
    out = np.zeros(max(k, 0))
    if k > 0:
      self._lib.qtqp_batch_launch_times(self._h, _ptr(out, _dp), k)
    return out

  TRACE_FIELDS = ("mu", "sigma", "alpha", "tau", "tau_p", "tau_c", "alpha_p", "q_solves",
                  "q_res", "p_solves", "p_res", "c_solves", "c_res", "pres", "dres", "gap",
                  "p_tau_a", "p_tau_b", "p_tau_c", "p_kr_max", "c_tau_a", "c_tau_b",
                  "c_tau_c", "c_kr_max")

  def trace(self, q: int, iters: int) -> list:
    """Per-iteration scalars of QP q in the last solve (list of dicts)."""
    rows = min(int(iters), 128)
    out = np.zeros((max(rows, 1), len(self.TRACE_FIELDS)))
    _check(self._lib, self._lib.qtqp_batch_trace(self._h, int(q), _ptr(out, _dp), rows))
    return [dict(zip(self.TRACE_FIELDS, row)) for row in out[:rows]]

  def work_counts(self) -> np.ndarray:
    """(count, 5): FP32/FP64 factorisations, K products, FP32/FP64 factor solves."""
    out = np.zeros((self.count, 5), dtype=np.int32)
    _check(self._lib, self._lib.qtqp_batch_work(self._h, _ptr(out, _i32p)))
    return out

  def solve(self, **kwargs) -> BatchResult:
    cnt = self.solve_device(**kwargs)
    return BatchResult(self.results(), cnt["device_ms"], cnt)

  def free(self) -> None:
    if getattr(self, "_h", None) is not None:
      self._lib.qtqp_batch_free(self._h)
      self._h = None

  def __del__(self):
    try:
      self.free()
    except Exception:  # pylint: disable=broad-except
      pass
