"""Restatements of the reference test-suite generators and KKT assertion helpers.

Follows qtqp tests/test_qtqp.py: _gen_feasible 211-235, _gen_infeasible
238-255, _gen_unbounded 258-280, _append_dropped_inequalities 490-496,
_assert_solution 283-309, _assert_infeasible 312-325, _assert_unbounded
328-346.  Dense outputs (a, b, c, p).  `seed` is an int or a
numpy Generator: as with the reference's random_state=rng, a Generator is used
as it is, so one stream can feed several generators in turn.
"""
import numpy as np
import scipy.sparse as sp


def gen_feasible(m, n, z, seed, density=0.1):
  rng = np.random.default_rng(seed)
  w = rng.normal(size=m)
  x = rng.normal(size=n)
  y = w.copy()
  y[z:] = 0.5 * (w[z:] + np.abs(w[z:]))
  s = y - w
  a = sp.random(m, n, density=density, format="csc", random_state=rng,
                data_rvs=lambda k: rng.normal(size=k)).toarray()
  r = sp.random(n, n, density=0.01, format="csc", random_state=rng,
                data_rvs=lambda k: rng.normal(size=k)).toarray()
  p = r.T @ r * 0.01
  return a, a @ x + s, -a.T @ y, p


def gen_infeasible(m, n, z, seed):
  rng = np.random.default_rng(seed)
  w = rng.random(size=m)
  b = rng.normal(size=m)
  y = w.copy()
  y[z:] = 0.5 * (w[z:] + np.abs(w[z:]))
  a = rng.normal(size=(m, n))
  p = rng.normal(size=(n, n))
  a = a - np.outer(y, a.T @ y) / np.linalg.norm(y) ** 2
  b = -b / (b @ y)
  p = p.T @ p * 0.01
  c = rng.normal(size=n)
  return a, b, c, p


def gen_unbounded(m, n, z, seed):
  rng = np.random.default_rng(seed)
  w = rng.random(size=m)
  c = rng.normal(size=n)
  s = np.zeros(m)
  s[z:] = 0.5 * (w[z:] + np.abs(w[z:]))
  a = rng.normal(size=(m, n))
  p = rng.normal(size=(n, n))
  p = p.T @ p * 0.01
  # general eigensolver and the last pair it returns, as the reference does;
  # P = V diag(e) V' is then symmetric only to rounding (within the 1e-12
  # asymmetry that QTQP accepts)
  e, v = np.linalg.eig(p)
  e, v = np.real(e), np.real(v)
  e[-1] = 0.0
  x = v[:, -1]
  p = v @ np.diag(e) @ v.T
  a = a - np.outer(s + a @ x, x) / np.linalg.norm(x) ** 2
  c = -c / (c @ x)
  b = rng.normal(size=m)
  return a, b, c, p


def append_dropped_inequalities(a, b, n_extra, seed, rhs_value):
  """Rows that presolve drops (b >= 1e20): dense N(0,1) rows with rhs_value."""
  rng = np.random.default_rng(seed)
  a_extra = rng.normal(size=(n_extra, a.shape[1]))
  return np.vstack([a, a_extra]), np.concatenate([b, np.full(n_extra, rhs_value)])


def check_solution(x, y, s, a, b, c, p, z, atol=1e-7, rtol=1e-8):
  """KKT optimality at the solve tolerances (tests/test_qtqp.py:283-309)."""
  pcost = c @ x + 0.5 * x @ p @ x
  dcost = -b @ y - 0.5 * x @ p @ x
  pres = np.linalg.norm(a @ x + s - b, np.inf)
  dres = np.linalg.norm(p @ x + a.T @ y + c, np.inf)
  gap = abs(c @ x + b @ y + x @ p @ x)
  prel = max(np.linalg.norm(a @ x, np.inf), np.linalg.norm(s, np.inf), np.linalg.norm(b, np.inf))
  drel = max(np.linalg.norm(p @ x, np.inf), np.linalg.norm(a.T @ y, np.inf), np.linalg.norm(c, np.inf))
  assert gap < atol + rtol * min(abs(pcost), abs(dcost)), gap
  assert pres < atol + rtol * prel, pres
  assert dres < atol + rtol * drel, dres
  assert np.min(y[z:], initial=0.0) > -1e-9
  assert np.min(s[z:], initial=0.0) > -1e-9


def check_infeasible(x, y, s, a, b, z, atol=1e-8, rtol=1e-9):
  """Primal infeasibility certificate (tests/test_qtqp.py:312-325)."""
  assert np.all(np.isnan(x)) and np.all(np.isnan(s))
  np.testing.assert_allclose(b @ y, -1.0, atol=atol, rtol=rtol)
  assert np.min(y[z:], initial=0.0) > -1e-9
  assert np.linalg.norm(a.T @ y, np.inf) < atol + rtol * np.linalg.norm(y, np.inf)


def check_unbounded(x, y, s, a, c, p, z, atol=1e-8, rtol=1e-9):
  """Dual infeasibility certificate (tests/test_qtqp.py:328-346)."""
  assert np.all(np.isnan(y))
  np.testing.assert_allclose(c @ x, -1.0, atol=atol, rtol=rtol)
  assert np.min(s[z:], initial=0.0) > -1e-9
  tol = atol + rtol * np.linalg.norm(x, np.inf)
  assert np.linalg.norm(a @ x + s, np.inf) < tol
  assert np.linalg.norm(p @ x, np.inf) < tol
