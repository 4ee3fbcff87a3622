"""CPU FP64 oracle for qtqp's interior-point path -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it.  The shipped solver (``qtqp_amd``) never routes through it.

It restates, in plain dense numpy + scipy LAPACK/BLAS, the algorithm of the
reference qtqp v0.0.5 (paths relative to the reference checkout):

* ``src/qtqp/__init__.py``  -- Mehrotra predictor-corrector IPM on the
  homogeneous embedding (``QTQP._solve_impl`` 351-614), presolve 248-297,
  Ruiz equilibration 616-669, step length 671-686/882-886, sigma 682-706,
  Newton step 708-757, tau quadratic 759-817, linearized tau 819-857,
  normalisation 859-880, termination 888-1008.
* ``src/qtqp/direct.py``    -- ``DirectKktSolver`` (regularised diagonal,
  iterative refinement with the diagonal correction) 183-307.
* ``src/qtqp/solvers_dense.py`` -- ``ScipyDenseSolver`` (Gram/Schur reduction
  G = P_off + diag(R_x) + A' diag(1/R_y) A, dsyrk, jitter 1e-14 max diag,
  dpotrf/dpotrs, block matvec) 25-145.

Pinning (each function has at least one check in tests/test_oracle.py that a
plausible mistake would fail):
  * whole driver: the README example trajectory (README.md:99-112, golden
    fixture), known answers x=[0.3,-0.7] and x=[1], closed-form optima of the
    dense generator, infeasible/unbounded certificates;
  * KKT solve: K_true @ sol = rhs to 1e-10 and numpy.linalg.solve on tiny N;
  * matvec: against the assembled full symmetric K;
  * tau: two independent closed forms (reference tests 971-995), linear case,
    Newton convergence of the linearised form, non-finite guard;
  * sigma: normalise-then-compute form; step size: closed-form cases;
  * normalise: ||(x,y,tau)||^2 = m-z+1; equilibration: diagonal-product form
    and the round trip.
No function here is "parity unpinned".
"""

from __future__ import annotations

import math
import time
from typing import Any

import numpy as np
import scipy.sparse as sp
from scipy.linalg import blas, lapack

EPS = 1e-15  # qtqp's _EPS, src/qtqp/__init__.py:51

SOLVED = "solved"
INFEASIBLE = "infeasible"
UNBOUNDED = "unbounded"
HIT_MAX_ITER = "hit_max_iter"
FAILED = "failed"
UNFINISHED = "unfinished"


# ---------------------------------------------------------------------------
# Dense Gram/Cholesky backend: restates solvers_dense.py:25-145.
# ---------------------------------------------------------------------------
class DenseGramBackend:
  """K_reg = [P_off + diag(d_x), A'; A, diag(d_y)] with d_y < 0 (solvers_dense.py)."""

  def __init__(self, a_dense: np.ndarray, p_dense: np.ndarray):
    self.m, self.n = a_dense.shape
    self.A = np.ascontiguousarray(a_dense, dtype=np.float64)
    p_off = np.array(p_dense, dtype=np.float64, order="F")
    np.fill_diagonal(p_off, 0.0)
    self.P_off = p_off
    self.R_x = np.zeros(self.n)
    self.R_y = np.ones(self.m)
    self.chol = None

  def update_diag(self, diag: np.ndarray) -> None:
    # diag[n:] arrives negated (direct.py:206-214); R_y is its magnitude.
    self.R_x = np.array(diag[: self.n], dtype=np.float64)
    self.R_y = -np.array(diag[self.n :], dtype=np.float64)

  def factorize(self) -> None:
    # solvers_dense.py:99-116
    g = np.array(self.P_off, order="F")
    g[np.diag_indices(self.n)] += self.R_x
    a_s = np.asfortranarray(self.A / np.sqrt(self.R_y)[:, None])
    g = blas.dsyrk(1.0, a_s, beta=1.0, c=g, trans=1, lower=1, overwrite_c=True)
    g[np.diag_indices(self.n)] += 1e-14 * np.max(np.diag(g))
    chol, info = lapack.dpotrf(g, lower=True, overwrite_a=True)
    if info != 0:
      raise np.linalg.LinAlgError(f"Cholesky failed (dpotrf info={info})")
    self.chol = chol

  def matvec(self, v: np.ndarray) -> np.ndarray:
    # solvers_dense.py:118-128 (exact K_reg product, P_off from both triangles)
    n = self.n
    vx, vy = v[:n], v[n:]
    out = np.empty(n + self.m)
    out[:n] = self.P_off @ vx + self.R_x * vx + self.A.T @ vy
    out[n:] = self.A @ vx - self.R_y * vy
    return out

  def solve(self, rhs: np.ndarray) -> np.ndarray:
    # solvers_dense.py:130-142: Gram RHS, dpotrs, back-substitute y.
    n = self.n
    w = 1.0 / self.R_y
    g = rhs[:n] + self.A.T @ (w * rhs[n:])
    x, info = lapack.dpotrs(self.chol, g, lower=True)
    if info != 0:
      raise np.linalg.LinAlgError(f"dpotrs info={info}")
    out = np.empty(n + self.m)
    out[:n] = x
    out[n:] = w * (self.A @ x - rhs[n:])
    return out


# ---------------------------------------------------------------------------
# DirectKktSolver: restates direct.py:112-311.
# ---------------------------------------------------------------------------
class KktIR:
  """Regularised factorisation + iterative refinement on K_true (direct.py)."""

  def __init__(self, a_dense, p_dense, z, min_reg, max_ir, atol, rtol,
               backend_factory=DenseGramBackend):
    self.m, self.n = a_dense.shape
    self.z = z
    self.p_diag = np.diag(p_dense).copy()
    self.min_reg = min_reg
    self.max_ir = max_ir
    self.atol, self.rtol = atol, rtol
    self.backend = backend_factory(a_dense, p_dense)
    self.diag_corr = np.zeros(self.n + self.m)

  def update(self, mu: float, s: np.ndarray, y: np.ndarray) -> None:
    # direct.py:196-216
    n, z = self.n, self.z
    true = np.concatenate([self.p_diag + mu, np.full(z, mu), s[z:] / y[z:] + mu])
    reg = np.maximum(true, self.min_reg)
    true[n:] = -true[n:]
    reg[n:] = -reg[n:]
    self.backend.update_diag(reg)
    self.backend.factorize()
    self.diag_corr = reg - true

  def solve(self, rhs: np.ndarray, warm_start: np.ndarray):
    # direct.py:243-307
    n = self.n
    kkt_rhs = np.array(rhs, dtype=np.float64)
    kkt_rhs[n:] = -kkt_rhs[n:]
    rhs_norm = np.linalg.norm(kkt_rhs, np.inf)
    tol = self.atol + self.rtol * rhs_norm
    sol = np.array(warm_start, dtype=np.float64)
    res = kkt_rhs - self.backend.matvec(sol) + self.diag_corr * sol
    res_norm = np.linalg.norm(res, np.inf)
    status, solves = "non-converged", 0
    for solves in range(1, self.max_ir + 1):
      old = res_norm
      sol = sol + self.backend.solve(res)
      res = kkt_rhs - self.backend.matvec(sol) + self.diag_corr * sol
      res_norm = np.linalg.norm(res, np.inf)
      if res_norm < tol:
        status = "converged"
        break
      if res_norm >= old:
        status = "stalled"
        break
    if np.any(np.isnan(sol)):
      raise ValueError("Linear solver returned NaNs.")
    return sol, {"solves": solves, "final_residual_norm": res_norm,
                 "rhs_norm": rhs_norm, "tolerance": tol,
                 "converged": status == "converged", "status": status}


# ---------------------------------------------------------------------------
# Scalar pieces of the IPM (kept as free functions so tests can pin them).
# ---------------------------------------------------------------------------
def max_step(v: np.ndarray, dv: np.ndarray) -> float:
  """Largest alpha in [0,1] with v + alpha*dv >= 0 (__init__.py:671-680)."""
  neg = dv < -EPS
  if not np.any(neg):
    return 1.0
  return min(1.0, float(np.min(-v[neg] / dv[neg])))


def step_size(y, s, dy, ds, z) -> float:
  """__init__.py:882-886."""
  return min(max_step(s[z:], ds[z:]), max_step(y[z:], dy[z:]))


def sigma_mehrotra(mu, x, y, tau, s, alpha, dx, dy, dtau, ds, m, z) -> float:
  """sigma = clip((mu_aff/mu)^3, 0, 1) (__init__.py:682-706)."""
  xa, ya, ta, sa = x + alpha * dx, y + alpha * dy, tau + alpha * dtau, s + alpha * ds
  nrm2 = xa @ xa + ya @ ya + ta * ta
  mu_aff = (m - z + 1) / max(EPS * EPS, nrm2) * (ya @ sa) / (m - z)
  base = mu_aff / max(EPS, mu)
  return float(np.clip(base * base * base, 0.0, 1.0))


def normalize(x, y, tau, s, m, z):
  """Scale so that ||(x,y,tau)||^2 = m - z + 1 (__init__.py:859-880)."""
  scale = math.sqrt(m - z + 1) / max(EPS, math.sqrt(x @ x + y @ y + tau * tau))
  return x * scale, y * scale, tau * scale, s * scale


def tau_quadratic(n, q, kinv_q, kinv_r, p_dense, mu, mu_target, r_tau) -> float:
  """Root of t_a tau^2 + t_b tau + t_c = 0 (__init__.py:759-817). Raises ValueError."""
  t_a = mu + kinv_q @ q
  t_b = -r_tau - kinv_r @ q
  t_c = -mu_target
  if p_dense is not None and np.any(p_dense):
    pr = p_dense @ kinv_r[:n]
    pq = p_dense @ kinv_q[:n]
    t_a -= kinv_q[:n] @ pq
    t_b += kinv_r[:n] @ pq + kinv_q[:n] @ pr
    t_c -= kinv_r[:n] @ pr
  if abs(t_a) < EPS:
    if abs(t_b) < EPS:
      raise ValueError("degenerate tau equation")
    t = -t_c / t_b
    if not np.isfinite(t) or t < -1e-10:
      raise ValueError("invalid linear tau")
    return max(0.0, t)
  disc = t_b * t_b - 4 * t_a * t_c
  if disc < -1e-9:
    raise ValueError("negative discriminant")
  disc = max(0.0, disc)
  if t_b > 0:
    t = t_c / (-0.5 * (t_b + math.sqrt(disc)))
  else:
    t = (-0.5 * (t_b - math.sqrt(disc))) / t_a
  if not np.isfinite(t) or t < -1e-10:
    raise ValueError("invalid tau")
  return max(0.0, t)


def tau_linearized(n, q, kinv_q, kinv_r, p_dense, mu, mu_target, x, y, tau_c,
                   tau_anchor) -> float:
  """First-order tau update with a [0.1, 10] trust region (__init__.py:819-857)."""
  if p_dense is not None and np.any(p_dense):
    px = p_dense @ x
  else:
    px = np.zeros(n)
  q_z = q[:n] @ x + q[n:] @ y
  x_px = x @ px
  q_kq = q @ kinv_q
  px_kq = px @ kinv_q[:n]
  q_rz = q @ kinv_r - tau_c * q_kq - q_z
  px_rz = px @ kinv_r[:n] - tau_c * px_kq - x_px
  g = mu * tau_c * tau_c + (mu_target - mu) * tau_anchor * tau_c - tau_c * q_z - mu_target - x_px
  num = g - tau_c * q_rz - 2.0 * px_rz
  den = 2.0 * mu * tau_c + (mu_target - mu) * tau_anchor - q_z + tau_c * q_kq + 2.0 * px_kq
  t = tau_c + (0.0 if abs(den) < 1e-16 else -num / den)
  if not np.isfinite(t):
    return tau_c
  return min(max(t, 0.1 * tau_c), 10.0 * tau_c)


def ruiz(a_dense: np.ndarray, p_dense: np.ndarray, iters=10, lo=1e-3, hi=1e3):
  """Ruiz scaling D A E, E P E (__init__.py:616-663), dense form."""
  a = np.array(a_dense, dtype=np.float64)
  p = np.array(p_dense, dtype=np.float64)
  d = np.ones(a.shape[0])
  e = np.ones(a.shape[1])
  for _ in range(iters):
    di = np.abs(a).max(axis=1) if a.shape[1] else np.zeros(a.shape[0])
    di = np.clip(1.0 / np.sqrt(np.where(di == 0.0, 1.0, di)), lo, hi)
    ei = np.maximum(np.abs(a).max(axis=0) if a.shape[0] else 0.0, np.abs(p).max(axis=0))
    ei = np.clip(1.0 / np.sqrt(np.where(ei == 0.0, 1.0, ei)), lo, hi)
    a = di[:, None] * a * ei[None, :]
    p = ei[:, None] * p * ei[None, :]
    d *= di
    e *= ei
  return a, p, d, e


# ---------------------------------------------------------------------------
# The IPM driver.
# ---------------------------------------------------------------------------
def solve_qp(a, b, c, z, p=None, *, atol=1e-7, rtol=1e-8, atol_infeas=1e-8,
             rtol_infeas=1e-9, max_iter=100, step_size_scale=0.99,
             min_static_regularization=1e-8, max_iterative_refinement_steps=20,
             linear_solver_atol=1e-12, linear_solver_rtol=1e-12,
             equilibrate=True, backend_factory=DenseGramBackend,
             force_linearized_tau=False, start=None) -> dict[str, Any]:
  """Solve one QP exactly as qtqp's IPM does, with the dense Gram backend.

  Returns a dict with x, y, s, status (string), stats (per-iteration dicts,
  always collected) and iters.  Accepts dense arrays or scipy sparse matrices.
  start: optional (x, y, s, tau) in unscaled coordinates of the presolved QP,
  replacing the default start of _init_variables (__init__.py:299-307); it is
  equilibrated as that hook equilibrates the default one (668-669).  This is
  the batched engine's warm-start hook; the reference itself always starts
  from the default point.
  """
  a_full = a.toarray() if sp.issparse(a) else np.array(a, dtype=np.float64)
  m_full, n = a_full.shape
  b_full = np.array(b, dtype=np.float64)
  c = np.array(c, dtype=np.float64)
  if p is None:
    p_dense = np.zeros((n, n))
  else:
    p_dense = p.toarray() if sp.issparse(p) else np.array(p, dtype=np.float64)
  if z < 0 or z > m_full:
    raise ValueError("z out of range")
  # Presolve (__init__.py:248-270): drop inequality rows with b >= 1e20.
  if not np.all(np.isfinite(b_full[:z])):
    raise ValueError("Equality RHS entries in 'b' must be finite.")
  if np.any(np.isnan(b_full[z:]) | np.isneginf(b_full[z:])):
    raise ValueError("Inequality RHS entries invalid.")
  keep = np.ones(m_full, dtype=bool)
  keep[z:] = b_full[z:] < 1e20
  a0, b0 = a_full[keep], b_full[keep]
  m = int(keep.sum())
  if z == m:
    raise ValueError("effective z == m after presolve; some rows may have been dropped")

  if equilibrate:
    a_eq, p_eq, d, e = ruiz(a0, p_dense)
    b_eq, c_eq = b0 * d, c * e
  else:
    a_eq, p_eq, d, e = a0, p_dense, np.ones(m), np.ones(n)
    b_eq, c_eq = b0, c
  q = np.concatenate([c_eq, b_eq])
  norm_b, norm_c = np.linalg.norm(b0, np.inf), np.linalg.norm(c, np.inf)
  has_p = bool(np.any(p_dense))
  p_eq_or_none = p_eq if has_p else None

  kkt = KktIR(a_eq, p_eq, z, min_static_regularization,
              max_iterative_refinement_steps, linear_solver_atol,
              linear_solver_rtol, backend_factory)
  if start is None:
    x = np.zeros(n)
    y = np.zeros(m)
    s = np.zeros(m)
    y[z:] = 1.0
    s[z:] = 1.0
    tau = 1.0
  else:
    x, y, s = (np.array(v, dtype=np.float64) for v in start[:3])
    tau = float(start[3])
    if x.shape != (n,) or y.shape != (m,) or s.shape != (m,):
      raise ValueError("start has the wrong sizes")
  x, y, s = x / e, y / d, s * d
  kinv_q = np.zeros(n + m)
  stats = []
  status = UNFINISHED
  t0 = time.perf_counter()

  def newton(mu, mu_t, r_anchor, tau_anchor, x, y, s, tau, corr, kinv_q):
    r = (mu - mu_t) * r_anchor
    if mu_t != 0.0:
      r[n + z:] += mu_t / y[z:]
    r[n + z:] += s[z:]
    if corr is not None:
      r[n + z:] += corr
    kinv_r, st = kkt.solve(r, r_anchor)
    tau_new = None
    if (st["converged"] or st["final_residual_norm"] < 1e-7) and not force_linearized_tau:
      try:
        tau_new = tau_quadratic(n, q, kinv_q, kinv_r, p_eq_or_none, mu, mu_t,
                                (mu - mu_t) * tau_anchor)
        st["tau_method"] = "quadratic"
      except ValueError:
        pass
    if tau_new is None:
      st["tau_method"] = "linearized"
      tau_new = tau_linearized(n, q, kinv_q, kinv_r, p_eq_or_none, mu, mu_t, x,
                               y, tau, tau_anchor)
    xy = kinv_r - kinv_q * tau_new
    return xy[:n], xy[n:], tau_new, st

  for it in range(max_iter):
    st_i: dict[str, Any] = {}
    x, y, tau, s = normalize(x, y, tau, s, m, z)
    mu = (y @ s) / (m - z)
    kkt.update(mu, s, y)
    kinv_q, st_i["q_lin_sys_stats"] = kkt.solve(q, kinv_q)
    xp, yp, taup, st_i["predictor_lin_sys_stats"] = newton(
        mu, 0.0, np.concatenate([x, y]), tau, x, y, s, tau, None, kinv_q)
    dxp, dyp, dtaup = xp - x, yp - y, taup - tau
    ds = np.zeros(m)
    ds[z:] = -yp[z:] * s[z:] / y[z:]
    alpha_p = step_size(y, s, dyp, ds, z)
    sigma = sigma_mehrotra(mu, x, y, tau, s, alpha_p, dxp, dyp, dtaup, ds, m, z)
    corr = -ds[z:] * dyp[z:] / y[z:]
    xc, yc, tauc, st_i["corrector_lin_sys_stats"] = newton(
        mu, sigma * mu, np.concatenate([xp, yp]), taup, x, y, s, tau, corr, kinv_q)
    dx, dy, dtau = xc - x, yc - y, tauc - tau
    ds = np.zeros(m)
    ds[z:] = sigma * mu / y[z:] + corr - yc[z:] * s[z:] / y[z:]
    alpha = step_size(y, s, dy, ds, z)
    stp = step_size_scale * alpha
    x, y, tau, s = x + stp * dx, y + stp * dy, tau + stp * dtau, s + stp * ds
    y[z:] = np.maximum(y[z:], 1e-30)
    s[z:] = np.maximum(s[z:], 1e-30)
    tau = max(tau, 1e-30)

    # Termination on unequilibrated iterates and the post-presolve data
    # (__init__.py:888-965).
    xu, yu, su = e * x, d * y, s / d
    it_tau = 1.0 / max(tau, EPS)
    ax, aty = a0 @ xu, a0.T @ yu
    px = p_dense @ xu if has_p else np.zeros(n)
    xpx = xu @ px if has_p else 0.0
    ctx, bty = c @ xu, b0 @ yu
    pcost = (ctx + 0.5 * xpx * it_tau) * it_tau
    dcost = (-bty - 0.5 * xpx * it_tau) * it_tau
    pres = np.linalg.norm((ax + su) * it_tau - b0, np.inf)
    dres = np.linalg.norm((px + aty) * it_tau + c, np.inf)
    gap = abs((ctx + bty + xpx * it_tau) * it_tau)
    n_aty, n_px = np.linalg.norm(aty, np.inf), np.linalg.norm(px, np.inf)
    dinf_a = np.linalg.norm(ax + su, np.inf) / (abs(ctx) + EPS)
    dinf_p = n_px / (abs(ctx) + EPS)
    dinf = max(dinf_a, dinf_p)
    pinf = n_aty / (abs(bty) + EPS)
    prel = max(np.linalg.norm(ax, np.inf) * it_tau, np.linalg.norm(su, np.inf) * it_tau, norm_b)
    drel = max(n_px * it_tau, n_aty * it_tau, norm_c)
    nx, ny = np.linalg.norm(xu, np.inf), np.linalg.norm(yu, np.inf)
    if (gap < atol + rtol * min(abs(pcost), abs(dcost)) and pres < atol + rtol * prel
        and dres < atol + rtol * drel):
      status = SOLVED
    elif ctx < -EPS and dinf < atol_infeas + rtol_infeas * nx / abs(ctx):
      status = UNBOUNDED
    elif bty < -EPS and pinf < atol_infeas + rtol_infeas * ny / abs(bty):
      status = INFEASIBLE
    else:
      status = UNFINISHED
    st_i.update(iter=it, ctx=ctx, bty=bty, pcost=pcost, dcost=dcost, pres=pres,
                dres=dres, gap=gap, pinfeas=pinf, dinfeas=dinf, dinfeas_a=dinf_a,
                dinfeas_p=dinf_p, mu=mu, sigma=sigma, alpha=alpha, tau=tau,
                norm_x=nx, norm_y=ny, status=status,
                time=time.perf_counter() - t0, prelrhs=prel, drelrhs=drel)
    stats.append(st_i)
    if status != UNFINISHED:
      break
  else:
    status = HIT_MAX_ITER
    stats[-1]["status"] = status

  # Status-dependent scaling and presolve restore (__init__.py:580-614, 272-297).
  xu, yu, su = e * x, d * y, s / d
  drop = ~keep
  y_full = np.zeros(m_full)
  s_full = np.full(m_full, np.nan)
  if status == INFEASIBLE:
    xo = np.full(n, np.nan)
    yk = yu / abs(b0 @ yu)
    y_full[keep] = yk
    s_full[keep] = np.nan
  elif status == UNBOUNDED:
    actx = abs(c @ xu)
    xo = xu / actx
    y_full[:] = np.nan
    s_full[keep] = su / actx
  else:
    xo = xu / tau
    y_full[keep] = yu / tau
    s_full[keep] = su / tau
    if np.any(drop):
      s_full[drop] = b_full[drop] - a_full[drop] @ xo
  if status == UNFINISHED:
    status = FAILED
  return {"x": xo, "y": y_full, "s": s_full, "status": status, "stats": stats,
          "iters": len(stats), "tau": tau}
