"""Learned warm start for the batched IPM on a parametric MPC family.

The family fixes (A_d, B_d, Q, R, boxes, horizon) and varies the initial state
x0, so every QP shares A and P and only b, c depend on x0.  The network maps x0
to a starting point of the solver: its own normalised iterate (x, y, s, tau)
after ``k_star`` IPM iterations of a default-start solve, plus kappa = mu/tau.
It predicts x, log y, log s, log tau, log kappa, so a proposal is strictly
interior by construction.

Fallback rule (default start of qtqp __init__.py:299-307): a QP keeps the
default start when its proposal is not finite; a warm-started QP still
unfinished after ``cap`` iterations is re-solved from the default start and
charged both runs' iterations.

Training and inference run on the GPU through include/qtqp_mlp.h (BF16
MFMA products with FP32 accumulation, Adam).
"""

from __future__ import annotations

import ctypes
import dataclasses
import time

import numpy as np

from . import _native, batch, problems
from . import SolutionStatus

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)

MLP_SYMBOLS = {
    "qtqp_mlp_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "qtqp_mlp_num_params": (ctypes.c_int64, [ctypes.c_void_p]),
    "qtqp_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "qtqp_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, _fp]),
    "qtqp_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int64, _fp]),
    "qtqp_mlp_loss_grad": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, ctypes.c_int32, _fp, _dp]),
    "qtqp_mlp_propose": (ctypes.c_int, [ctypes.c_void_p, _fp, ctypes.c_int64, _dp, _dp,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "qtqp_mlp_train": (ctypes.c_int, [ctypes.c_void_p, _fp, _fp, ctypes.c_int64, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _fp]),
    "qtqp_mlp_free": (None, [ctypes.c_void_p]),
    "qtqp_mlp_last_error": (ctypes.c_char_p, []),
}


def mlp_lib() -> ctypes.CDLL:
  lib = _native.load("libqtqp_mlp.so", MLP_SYMBOLS)
  _native.require_device(_native.kkt_lib())
  return lib


def _f32(a):
  return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
  """n_in -> hidden -> hidden -> n_out perceptron with GELU, on the GPU."""

  def __init__(self, n_in: int, hidden: int, n_out: int, max_batch: int = 512, seed: int = 0):
    self._lib = mlp_lib()
    self.n_in, self.hidden, self.n_out, self.max_batch = n_in, hidden, n_out, max_batch
    h = ctypes.c_void_p()
    self._check(self._lib.qtqp_mlp_create(n_in, hidden, n_out, max_batch, ctypes.byref(h)))
    self._h = h
    self.num_params = int(self._lib.qtqp_mlp_num_params(h))
    self.set_params(init_params(n_in, hidden, n_out, seed))

  def _check(self, rc):
    if rc != 0:
      msg = (self._lib.qtqp_mlp_last_error() or b"").decode(errors="replace")
      raise (ValueError if rc == 1 else RuntimeError)(f"qtqp_mlp error {rc}: {msg}")

  def set_params(self, p):
    p = _f32(p)
    assert p.size == self.num_params
    self._check(self._lib.qtqp_mlp_set_params(self._h, p.ctypes.data_as(_fp)))

  def get_params(self) -> np.ndarray:
    p = np.empty(self.num_params, dtype=np.float32)
    self._check(self._lib.qtqp_mlp_get_params(self._h, p.ctypes.data_as(_fp)))
    return p

  def forward(self, x) -> np.ndarray:
    x = _f32(x)
    y = np.empty((x.shape[0], self.n_out), dtype=np.float32)
    self._check(self._lib.qtqp_mlp_forward(self._h, x.ctypes.data_as(_fp), x.shape[0],
                                           y.ctypes.data_as(_fp)))
    return y

  def propose(self, x, tm, ts, n, m, z, out) -> np.ndarray:
    """Decoded starting points of the IPM for the standardised inputs x,
    written to the device addresses out = (x, y, s, tau) (packed; see
    include/qtqp_mlp.h).  Returns the per-row validity flags."""
    x = _f32(x)
    tm = np.ascontiguousarray(tm, dtype=np.float64)
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    valid = np.empty(x.shape[0], dtype=np.int32)
    self._check(self._lib.qtqp_mlp_propose(
        self._h, x.ctypes.data_as(_fp), x.shape[0], tm.ctypes.data_as(_dp),
        ts.ctypes.data_as(_dp), n, m, z, *out, valid.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return valid.astype(bool)

  def loss_grad(self, x, t):
    x, t = _f32(x), _f32(t)
    g = np.empty(self.num_params, dtype=np.float32)
    loss = ctypes.c_double()
    self._check(self._lib.qtqp_mlp_loss_grad(self._h, x.ctypes.data_as(_fp),
                                             t.ctypes.data_as(_fp), x.shape[0],
                                             g.ctypes.data_as(_fp), ctypes.byref(loss)))
    return float(loss.value), g

  def train(self, x, t, *, batch=512, steps=2000, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
            seed=0) -> np.ndarray:
    x, t = _f32(x), _f32(t)
    losses = np.empty(steps, dtype=np.float32)
    self._check(self._lib.qtqp_mlp_train(self._h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp),
                                         x.shape[0], batch, steps, lr, beta1, beta2, eps, seed,
                                         losses.ctypes.data_as(_fp)))
    return losses

  def free(self):
    if getattr(self, "_h", None) is not None:
      self._lib.qtqp_mlp_free(self._h)
      self._h = None

  def __del__(self):
    try:
      self.free()
    except Exception:  # pylint: disable=broad-except
      pass


def init_params(n_in, hidden, n_out, seed=0) -> np.ndarray:
  """He-normal weights, zero biases, packed as W1, b1, W2, b2, W3, b3."""
  rng = np.random.default_rng(seed)
  parts = []
  for fan_in, fan_out in ((n_in, hidden), (hidden, hidden), (hidden, n_out)):
    parts.append(rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in))
    parts.append(np.zeros(fan_out))
  return np.concatenate(parts).astype(np.float32)


@dataclasses.dataclass
class Family:
  """Parametric condensed MPC: shared A, P; b(x0), c(x0) affine in x0."""
  horizon: int = 32
  nx: int = 8
  nu: int = 3
  family_seed: int = 7

  def __post_init__(self):
    self.sys = problems.parametric_mpc_family(self.family_seed, self.horizon, self.nx, self.nu)
    base = problems.condensed_mpc(self.sys, self.horizon, np.zeros(self.nx))
    self.a, self.p, self.z = base.a, base.p, base.z
    self.n, self.m = base.n, base.m
    # b = b0 + Bx x0, c = Cx x0 (read off by linearity)
    eye = np.eye(self.nx)
    cols = [problems.condensed_mpc(self.sys, self.horizon, eye[k]) for k in range(self.nx)]
    self.b0, self.c0 = base.b, base.c
    self.bx = np.stack([q.b - base.b for q in cols], axis=1)
    self.cx = np.stack([q.c - base.c for q in cols], axis=1)
    # [Bx'; b0] and [Cx'; c0]: b and c of many x0 as one product each
    self._bxa = np.vstack([self.bx.T, self.b0[None, :]])
    self._cxa = np.vstack([self.cx.T, self.c0[None, :]])

  def bc(self, x0s):
    """b (count, m) and c (count, n) of the family's QPs at the x0s:
    [x0, 1] [Bx'; b0] and [x0, 1] [Cx'; c0]."""
    x1 = np.empty((x0s.shape[0], self.nx + 1))
    x1[:, :self.nx] = x0s
    x1[:, self.nx] = 1.0
    return x1 @ self._bxa, x1 @ self._cxa

  def qps(self, x0s):
    b, c = self.bc(x0s)
    return [(self.a, b[i], c[i], self.z, self.p) for i in range(x0s.shape[0])]


class FamilySolver:
  """Batched solves of one Family at many x0: a family batch (A and P held
  once on the device) of `chunk` QPs stays resident, and each solve uploads
  only the QPs' b and c (and a start).  A chunk of another size gets a batch
  of its own."""

  def __init__(self, fam: Family, chunk: int = 16384):
    self.fam, self.chunk = fam, chunk
    self._bqs = {}  # resident family batches by QP count

  def _batch(self, x0s):
    b, c = self.fam.bc(x0s)
    bq = self._bqs.get(b.shape[0])
    if bq is not None:
      bq.set_bc(b, c)
    else:
      bq = batch.BatchQP.family(self.fam.a, self.fam.p, self.fam.z, b, c)
      self._bqs[b.shape[0]] = bq
    return bq

  def solve(self, x0s, *, start=None, proposer=None, max_iter=100, precision="mixed",
            raw=False, **opts) -> dict:
    """Solves the family's QPs at the x0s.  start: host arrays (x, y, s, tau)
    of every QP, or proposer: a WarmStart whose network writes the start on
    the device.  Returns per-QP arrays status, iters, sol_x, sol_y, sol_s
    (the solutions as QTQP.solve returns them), valid (proposals that were
    used) and, with raw, the iterates x, y, s, tau before the final scaling;
    plus the summed device_ms."""
    fam = self.fam
    keys = ["status", "iters", "sol_x", "sol_y", "sol_s", "valid"]
    if raw:
      keys += ["x", "y", "s", "tau"]
    out = {k: [] for k in keys}
    t_dev = 0.0
    for c0 in range(0, x0s.shape[0], self.chunk):
      xs = x0s[c0:c0 + self.chunk]
      k = xs.shape[0]
      bq = self._batch(xs)
      valid = np.ones(k, dtype=bool)
      if proposer is not None:
        valid = proposer.propose_device(xs, bq)
      elif start is not None:
        sx, sy, ss, st = (v[c0:c0 + self.chunk] for v in start)
        bq.set_start(sx.ravel(), sy.ravel(), ss.ravel(), st)
      cnt = bq.solve_device(max_iter=max_iter, precision=precision, **opts)
      t_dev += cnt["device_ms"]
      if raw:
        r = bq.raw_iterates()
        out["x"].append(r["x"].reshape(k, fam.n))
        out["y"].append(r["y"].reshape(k, fam.m))
        out["s"].append(r["s"].reshape(k, fam.m))
        out["tau"].append(r["tau"])
      res = bq.results_arrays()
      out["status"].append(res["status"])
      out["iters"].append(res["iters"])
      out["sol_x"].append(res["x"])
      out["sol_y"].append(res["y"])
      out["sol_s"].append(res["s"])
      out["valid"].append(valid)
    # (one chunk: its arrays as they are, without another copy of every solution)
    res = {k: v[0] if len(v) == 1 else np.concatenate(v) for k, v in out.items()}
    res["device_ms"] = t_dev
    return res

  def free(self):
    for bq in self._bqs.values():
      bq.free()
    self._bqs = {}


def _solve(fam: Family, x0s, *, start=None, max_iter=100, chunk=16384, precision="mixed",
           **opts):
  """One-off batched solves of the family at the x0s (a FamilySolver for this
  call); returns the raw iterates too."""
  solver = FamilySolver(fam, chunk)
  try:
    return solver.solve(x0s, start=start, max_iter=max_iter, precision=precision, raw=True,
                        **opts)
  finally:
    solver.free()


class WarmStart:
  """x0 -> starting point of the batched IPM for one Family."""

  def __init__(self, fam: Family, hidden=256, k_star=6, seed=0):
    self.fam, self.k_star = fam, k_star
    self.n_out = fam.n + 2 * fam.m + 2
    self.mlp = MLP(fam.nx, hidden, self.n_out, max_batch=512, seed=seed)
    self.xm = self.xs = self.tm = self.ts = None

  def targets(self, x0s, chunk=16384):
    """Solver iterates after k_star iterations, as (x, log y, log s, log tau, log kappa)."""
    r = _solve(self.fam, x0s, max_iter=self.k_star, chunk=chunk)
    z = self.fam.z
    y, s, tau = r["y"], r["s"], r["tau"]
    mu = np.sum(y[:, z:] * s[:, z:], axis=1) / (self.fam.m - z)
    t = np.concatenate([r["x"], np.log(np.maximum(y, 1e-300)), np.log(np.maximum(s, 1e-300)),
                        np.log(tau)[:, None], np.log(mu / tau)[:, None]], axis=1)
    return t, r

  def fit(self, x0s, t, *, steps=3000, batch_size=512, lr=1e-3, seed=0):
    self.xm, self.xs = x0s.mean(0), x0s.std(0) + 1e-12
    self.tm, self.ts = t.mean(0), t.std(0) + 1e-6
    return self.mlp.train((x0s - self.xm) / self.xs, (t - self.tm) / self.ts, batch=batch_size,
                          steps=steps, lr=lr, seed=seed)

  # -- checkpoint: one .npz (no pickled objects) -----------------------------
  _FORMAT = "qtqp_amd.warmstart.v1"

  def save(self, path) -> None:
    """Writes the family, k*, the network's packed FP32 parameters and the
    input/target standardisation."""
    if self.xm is None:
      raise ValueError("fit() before save()")
    f = self.fam
    np.savez(path, format=np.array(self._FORMAT),
             family=np.array([f.horizon, f.nx, f.nu, f.family_seed], np.int64),
             shape=np.array([self.mlp.n_in, self.mlp.hidden, self.n_out, self.k_star], np.int64),
             params=self.mlp.get_params(), xm=self.xm, xs=self.xs, tm=self.tm, ts=self.ts)

  @classmethod
  def load(cls, path) -> "WarmStart":
    with np.load(path, allow_pickle=False) as d:
      if str(d["format"]) != cls._FORMAT:
        raise ValueError(f"not a {cls._FORMAT} file")
      horizon, nx, nu, fseed = (int(v) for v in d["family"])
      n_in, hidden, n_out, k_star = (int(v) for v in d["shape"])
      ws = cls(Family(horizon=horizon, nx=nx, nu=nu, family_seed=fseed), hidden=hidden,
               k_star=k_star)
      if ws.n_out != n_out or ws.mlp.n_in != n_in:
        raise ValueError("checkpoint does not match its family")
      ws.mlp.set_params(d["params"])
      ws.xm, ws.xs, ws.tm, ws.ts = (d[k].copy() for k in ("xm", "xs", "tm", "ts"))
    return ws

  def propose_device(self, x0s, bq) -> np.ndarray:
    """Writes the proposed start of each x0 into the family batch bq on the
    device (qtqp_mlp_propose, then set_start_device): the decoding of
    ``propose``, with the default start for rows that are not finite.
    Returns the validity flags."""
    fam = self.fam
    bufs = bq.device_buffers()
    valid = self.mlp.propose((x0s - self.xm) / self.xs, self.tm, self.ts, fam.n, fam.m, fam.z,
                             bufs)
    bq.set_start_device(*bufs)
    return valid

  def propose(self, x0s):
    """(x, y, s, tau, kappa, valid) for each x0; unscaled solver coordinates."""
    raw = self.mlp.forward((x0s - self.xm) / self.xs).astype(np.float64) * self.ts + self.tm
    n, m = self.fam.n, self.fam.m
    x = raw[:, :n]
    y = np.exp(np.clip(raw[:, n:n + m], -700, 700))
    s = np.exp(np.clip(raw[:, n + m:n + 2 * m], -700, 700))
    tau = np.exp(np.clip(raw[:, -2], -700, 700))
    kappa = np.exp(np.clip(raw[:, -1], -700, 700))
    z = self.fam.z
    y[:, :z] = raw[:, n:n + z]  # free duals of equality rows (none in MPC)
    s[:, :z] = 0.0
    valid = (np.all(np.isfinite(raw), axis=1) & np.all(y[:, z:] > 0, axis=1)
             & np.all(s[:, z:] > 0, axis=1) & (tau > 0) & (kappa > 0))
    return x, y, s, tau, kappa, valid


def default_start(fam: Family, count: int):
  x = np.zeros((count, fam.n))
  y = np.zeros((count, fam.m))
  s = np.zeros((count, fam.m))
  y[:, fam.z:] = 1.0
  s[:, fam.z:] = 1.0
  return x, y, s, np.ones(count)


def solve_with_warm_start(ws: WarmStart, x0s, *, cap: int, chunk=16384, solver=None, **opts):
  """Warm-started batched solve with the default-start fallback rule: the
  network's proposals are decoded into the batch on the device (invalid ones
  replaced by the default start), and QPs still unfinished after `cap`
  iterations are re-solved from the default start.  `solver`: a resident
  FamilySolver of ws.fam to use (one is made for the call otherwise)."""
  fam = ws.fam
  own = solver is None
  solver = FamilySolver(fam, chunk) if own else solver
  t0 = time.perf_counter()
  try:
    warm = solver.solve(x0s, proposer=ws, max_iter=cap, **opts)
    valid = warm["valid"]
    unfinished = warm["status"] == 0
    iters = warm["iters"].astype(np.int64).copy()
    status = warm["status"].copy()
    sol_x, sol_y, sol_s = warm["sol_x"], warm["sol_y"], warm["sol_s"]
    device_ms = warm["device_ms"]
    if unfinished.any():
      redo = solver.solve(x0s[unfinished], **opts)
      iters[unfinished] += redo["iters"]
      status[unfinished] = redo["status"]
      sol_x[unfinished], sol_y[unfinished], sol_s[unfinished] = (redo["sol_x"], redo["sol_y"],
                                                                 redo["sol_s"])
      device_ms += redo["device_ms"]
  finally:
    if own:
      solver.free()
  return {"iters": iters, "status": status, "fallback": unfinished | ~valid,
          "invalid": ~valid, "x": sol_x, "y": sol_y, "s": sol_s, "device_ms": device_ms,
          "wall_s": time.perf_counter() - t0}
