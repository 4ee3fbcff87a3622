"""Pins the CPU oracle (oracle/qtqp_oracle.py) before it is trusted as the checker.

Each oracle function is checked against something that does not restate its
own formula: reference trajectories (tests/golden), known answers, closed-form
optima planted by the generators, library routines (numpy.linalg.solve),
independent formulas from the reference test-suite (tests/test_qtqp.py) and
invariants.  A few tests plant a mistake in the oracle and show that the
golden fixtures catch it.
"""
import json
import pathlib

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import qtqp_oracle as O
from qtqp_amd import problems
import golden_cases
import qp_generators as G

TRAJ = golden_cases.load_trajectories()


def _nanvec(v):
  return np.array([np.nan if t is None else t for t in v], dtype=float)


def _relerr(u, v):
  u, v = np.asarray(u, float), np.asarray(v, float)
  if np.all(np.isnan(u)) and np.all(np.isnan(v)):
    return 0.0
  assert np.array_equal(np.isnan(u), np.isnan(v))
  ok = ~np.isnan(v)
  return float(np.max(np.abs(u[ok] - v[ok])) / max(1.0, np.max(np.abs(v[ok]))))


@pytest.mark.parametrize("key", sorted(TRAJ))
def test_oracle_matches_reference_trajectory(key):
  case = TRAJ[key]
  a, b, c, z, p = golden_cases.inputs(case["spec"])
  r = O.solve_qp(a, b, c, z, p, equilibrate=case["equilibrate"])
  assert r["status"] == case["status"]
  assert r["iters"] == case["iters"]
  mus = np.array([st["mu"] for st in r["stats"]])
  ref = np.array([t["mu"] for t in case["trajectory"]])
  # Early iterations agree to rounding; later ones drift with LAPACK order.
  np.testing.assert_allclose(mus[:3], ref[:3], rtol=1e-8)
  taus = [[st[k]["tau_method"] for k in ("predictor_lin_sys_stats", "corrector_lin_sys_stats")]
          for st in r["stats"]]
  assert taus[:3] == [t["tau_method"] for t in case["trajectory"]][:3]
  if case["status"] == "solved" and case["spec"]["kind"] in ("readme", "dense", "mpc"):
    for k in "xys":
      assert _relerr(r[k], _nanvec(case[k])) < 1e-6, k
  else:
    for k in "xys":
      assert _relerr(r[k], _nanvec(case[k])) < 1e-4, k


def test_readme_known_answer():
  fx = json.loads((golden_cases.GOLDEN / "readme_example.json").read_text())
  pr = fx["problem"]
  r = O.solve_qp(np.array(pr["a"]), np.array(pr["b"]), np.array(pr["c"]), pr["z"],
                 np.array(pr["p"]))
  assert r["status"] == O.SOLVED
  np.testing.assert_allclose(r["x"], [0.3, -0.7], atol=1e-7)       # tests/test_qtqp.py:1645
  np.testing.assert_allclose(r["x"], fx["x"], atol=1e-6)           # README.md:110 (v0.0.3)
  np.testing.assert_allclose(r["y"][:2], fx["y"][:2], atol=1e-6)


def test_minimal_known_answer():
  # min 1/2 x^2 - x  s.t. 0 <= x <= 2  ->  x = 1  (tests/test_qtqp.py:2155-2166)
  r = O.solve_qp(np.array([[-1.0], [1.0]]), np.array([0.0, 2.0]), np.array([-1.0]), 0,
                 np.array([[1.0]]))
  assert r["status"] == O.SOLVED
  np.testing.assert_allclose(r["x"], [1.0], atol=1e-7)


@pytest.mark.parametrize("seed", [0, 1])
def test_dense_generator_planted_optimum(seed):
  # At the default tolerances (1e-7/1e-8) the iterate is ~1e-4 from the
  # optimum; tightening them exposes the planted closed-form answer.
  qp = problems.random_dense_qp(200, 300, 50, seed=seed)
  r = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p, atol=1e-10, rtol=1e-11)
  assert r["status"] == O.SOLVED
  for got, want, bound in ((r["x"], qp.x_star, 1e-6), (r["y"], qp.y_star, 1e-6),
                           (r["s"], qp.s_star, 1e-5)):
    assert np.max(np.abs(got - want)) / max(1.0, np.max(np.abs(want))) < bound


def test_mpc_generator_kkt():
  qp = problems.mpc_qp(16, 6, 2, seed=5)
  r = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
  assert r["status"] == O.SOLVED
  G.check_solution(r["x"], r["y"], r["s"], qp.a, qp.b, qp.c, qp.p, qp.z)


def test_mpc_condensing_matches_simulation():
  # X = Phi x0 + Gamma U must equal a forward simulation of x+ = A x + B u.
  rng = np.random.default_rng(3)
  sys_ = problems.mpc_system(5, 2, rng)
  x0 = rng.standard_normal(5)
  qp = problems.condensed_mpc(sys_, 7, x0)
  u = rng.standard_normal(14)
  x, states = x0, []
  for t in range(7):
    x = sys_.ad @ x + sys_.bd @ u[2 * t:2 * t + 2]
    states.append(x)
  states = np.concatenate(states)
  n = 14
  gam = qp.a[2 * n:2 * n + 35]
  fx = 5.0 - qp.b[2 * n:2 * n + 35]
  np.testing.assert_allclose(gam @ u + fx, states, atol=1e-12)
  # objective 1/2 ||X||^2 + 0.05 ||U||^2 = 1/2 u'Pu + c'u + const
  obj = 0.5 * states @ states + 0.05 * u @ u
  quad = 0.5 * u @ qp.p @ u + qp.c @ u + 0.5 * fx @ fx
  np.testing.assert_allclose(obj, quad, rtol=1e-12)


@pytest.mark.parametrize("kind", ["infeasible", "unbounded"])
@pytest.mark.parametrize("seed", [142, 143])
def test_certificates(kind, seed):
  m, n, z = 40, 25, 5
  if kind == "infeasible":
    a, b, c, p = G.gen_infeasible(m, n, z, seed)
    r = O.solve_qp(a, b, c, z, p)
    assert r["status"] == O.INFEASIBLE
    G.check_infeasible(r["x"], r["y"], r["s"], a, b, z)
  else:
    a, b, c, p = G.gen_unbounded(m, n, z, seed)
    r = O.solve_qp(a, b, c, z, p)
    assert r["status"] == O.UNBOUNDED
    G.check_unbounded(r["x"], r["y"], r["s"], a, c, p, z)


def _kkt_case(seed, m=60, n=40, z=6):
  a, b, c, p = G.gen_feasible(m, n, z, seed, density=0.3)
  rng = np.random.default_rng(seed + 1000)
  mu = rng.uniform()
  s = rng.uniform(size=m)
  y = rng.uniform(size=m)
  s[:z] = 0.0
  return a, b, c, p, mu, s, y


@pytest.mark.parametrize("seed", range(3))
def test_kkt_solve_satisfies_block_equations(seed):
  # tests/test_qtqp.py:740-771
  a, b, c, p, mu, s, y = _kkt_case(seed)
  m, n = a.shape
  z = 6
  kkt = O.KktIR(a, p, z, 1e-8, 10, 1e-12, 1e-12)
  kkt.update(mu, s, y)
  sol, st = kkt.solve(np.concatenate([c, b]), np.zeros(n + m))
  d = np.concatenate([np.zeros(z), s[z:] / y[z:]])
  np.testing.assert_allclose(p @ sol[:n] + mu * sol[:n] + a.T @ sol[n:], c, atol=1e-10, rtol=1e-10)
  np.testing.assert_allclose(-a @ sol[:n] + (d + mu) * sol[n:], b, atol=1e-10, rtol=1e-10)
  assert st["converged"]


def test_kkt_solve_matches_numpy_on_tiny_system():
  a, b, c, p, mu, s, y = _kkt_case(7, m=6, n=4, z=1)
  m, n = a.shape
  kkt = O.KktIR(a, p, 1, 1e-8, 10, 1e-12, 1e-12)
  kkt.update(mu, s, y)
  rhs = np.random.default_rng(0).standard_normal(n + m)
  sol, _ = kkt.solve(rhs, np.zeros(n + m))
  hy = np.concatenate([np.full(1, mu), s[1:] / y[1:] + mu])
  k_true = np.block([[p + mu * np.eye(n), a.T], [a, -np.diag(hy)]])
  ref = np.linalg.solve(k_true, np.concatenate([rhs[:n], -rhs[n:]]))
  np.testing.assert_allclose(sol, ref, rtol=1e-10, atol=1e-12)


def test_backend_matvec_matches_full_kkt():
  # tests/test_qtqp.py:774-819
  a, b, c, p, mu, s, y = _kkt_case(11, m=20, n=12, z=4)
  m, n = a.shape
  be = O.DenseGramBackend(a, p)
  hy = np.concatenate([np.full(4, mu), s[4:] / y[4:] + mu])
  be.update_diag(np.concatenate([np.diag(p) + mu, -hy]))
  v = np.random.default_rng(1).standard_normal(n + m)
  full = np.block([[p + mu * np.eye(n), a.T], [a, -np.diag(hy)]])
  np.testing.assert_allclose(be.matvec(v), full @ v, rtol=1e-12, atol=1e-12)


def _tau_alternative(n, kinv_q, kinv_r, mu, mu_target, r, r_tau, s, y):
  # Independent form from tests/test_qtqp.py:982-995 (uses D = s/y, not P).
  kr_d_kr = kinv_r[n:] @ (kinv_r[n:] * s / y)
  kq_d_kq = kinv_q[n:] @ (kinv_q[n:] * s / y)
  kq_d_kr = kinv_q[n:] @ (kinv_r[n:] * s / y)
  t_a = mu + mu * kinv_q @ kinv_q + kq_d_kq
  t_b = -r_tau + kinv_q @ r - 2 * (mu * kinv_q @ kinv_r + kq_d_kr)
  t_c = -kinv_r @ r + mu * kinv_r @ kinv_r + kr_d_kr - mu_target
  return (-t_b + np.sqrt(max(0.0, t_b**2 - 4 * t_a * t_c))) / (2 * t_a)


@pytest.mark.parametrize("seed", range(4))
def test_tau_quadratic_matches_independent_form(seed):
  # tests/test_qtqp.py:998-1044 (z = 0 so that D = s/y on every y-row)
  m, n, z = 50, 30, 0
  a, b, c, p = G.gen_feasible(m, n, z, seed + 20, density=0.3)
  rng = np.random.default_rng(seed)
  mu = rng.uniform()
  mu_t = rng.uniform() * mu
  s = rng.uniform(size=m)
  y = rng.uniform(size=m)
  kkt = O.KktIR(a, p, z, 1e-8, 10, 1e-12, 1e-12)
  kkt.update(mu, s, y)
  q = np.concatenate([c, b])
  kinv_q, _ = kkt.solve(q, np.zeros(n + m))
  r = (mu - mu_t) * rng.uniform(size=n + m)
  r[n:] += mu_t / y + s
  kinv_r, _ = kkt.solve(r, np.zeros(n + m))
  r_tau = (mu - mu_t) * rng.uniform()
  t1 = O.tau_quadratic(n, q, kinv_q, kinv_r, p, mu, mu_t, r_tau)
  t2 = _tau_alternative(n, kinv_q, kinv_r, mu, mu_t, r, r_tau, s, y)
  np.testing.assert_allclose(t1, t2, rtol=1e-10, atol=1e-11)


def test_tau_linear_case():
  # tests/test_qtqp.py:1047-1061: t_a = 0 -> tau = -t_c/t_b = 2
  t = O.tau_quadratic(1, np.array([1.0, 0.0]), np.array([-1.0, 0.0]), np.array([-2.0, 0.0]),
                      None, 1.0, 4.0, 0.0)
  assert t == pytest.approx(2.0)


@pytest.mark.parametrize("seed", range(3))
def test_linearized_tau_is_newton_on_the_quadratic(seed):
  # tests/test_qtqp.py:1064-1127
  rng = np.random.default_rng(seed + 42)
  n, m = 20, 30
  raw = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.2)
  p = raw.T @ raw
  q = rng.standard_normal(n + m)
  kr = rng.standard_normal(n + m)
  kq = rng.standard_normal(n + m)
  tau_anchor = rng.uniform(0.5, 2.0)
  x1, x2 = kr[:n], kq[:n]
  ta0 = q @ kq - x2 @ p @ x2
  mu = max(2.0 * abs(ta0), 1.0)
  mu_t = rng.uniform(0.0, 0.5 * mu)
  t_a = mu + ta0
  t_b = (mu_t - mu) * tau_anchor - q @ kr + 2.0 * x1 @ p @ x2
  t_c = -mu_t - x1 @ p @ x1
  exact = (-t_b + np.sqrt(t_b**2 - 4 * t_a * t_c)) / (2 * t_a)
  tau = 0.9 * exact
  for _ in range(15):
    zk = kr - tau * kq
    tau = O.tau_linearized(n, q, kq, kr, p, mu, mu_t, zk[:n], zk[n:], tau, tau_anchor)
  np.testing.assert_allclose(tau, exact, rtol=1e-12, atol=1e-12)


def test_linearized_tau_nonfinite_guard():
  # tests/test_qtqp.py:1130-1152
  p = np.array([[1e150, 0.0], [0.0, -1e150]])
  tau = O.tau_linearized(2, np.array([1e45, 1e-47, 1e-82, 1e-21]),
                         np.array([-1e79, -1e-73, -1e-68, -1e-78]),
                         np.array([1e112, 1e140, -1e-108, 1e19]), p, 1e-113, 0.0,
                         np.array([-1e-87, 1e117]), np.array([1e-85, 1e-83]), 1e-113, 1e84)
  assert tau == 1e-113


@pytest.mark.parametrize("seed", range(3))
def test_sigma_matches_normalize_then_compute(seed):
  # tests/test_qtqp.py:1299-1337
  rng = np.random.default_rng(seed)
  m, n, z = 30, 20, 4
  mu = rng.uniform(1e-6, 1.0)
  alpha = rng.uniform()
  x, y, tau, s = rng.normal(size=n), rng.uniform(size=m), rng.uniform(0.1, 2), rng.uniform(size=m)
  dx, dy, dtau, ds = rng.normal(size=n), rng.normal(size=m), rng.normal(), rng.normal(size=m)
  sig = O.sigma_mehrotra(mu, x, y, tau, s, alpha, dx, dy, dtau, ds, m, z)
  _, yn, _, sn = O.normalize(x + alpha * dx, y + alpha * dy, tau + alpha * dtau, s + alpha * ds, m, z)
  ref = float(np.clip(((yn @ sn) / (m - z) / mu) ** 3, 0, 1))
  np.testing.assert_allclose(sig, ref, rtol=1e-13, atol=1e-14)


def test_max_step_cases():
  # tests/test_qtqp.py:1723-1746
  y = np.array([1.0, 2.0, 3.0])
  assert O.max_step(y, np.array([1.0, 1.0, 1.0])) == 1.0
  assert O.max_step(y, np.array([-2.0, 1.0, -1.0])) == pytest.approx(0.5)
  assert O.max_step(y, np.array([-0.1, 0.0, 0.0])) == 1.0


def test_normalize_invariant():
  rng = np.random.default_rng(5)
  m, z = 17, 3
  x, y, tau, s = O.normalize(rng.normal(size=9), rng.normal(size=m), 0.7, rng.normal(size=m), m, z)
  np.testing.assert_allclose(x @ x + y @ y + tau * tau, m - z + 1, rtol=1e-12)


def test_ruiz_matches_diagonal_products_and_round_trips():
  # tests/test_qtqp.py:1261-1296 and 1753-1771
  a, b, c, p = G.gen_feasible(40, 25, 5, 9, density=0.3)
  a_eq, p_eq, d, e = O.ruiz(a, p)
  np.testing.assert_allclose(a_eq, np.diag(d) @ a @ np.diag(e), rtol=1e-13, atol=1e-15)
  np.testing.assert_allclose(p_eq, np.diag(e) @ p @ np.diag(e), rtol=1e-13, atol=1e-15)
  rows = np.abs(a_eq).max(axis=1)
  assert np.all((rows > 0.5) & (rows < 2.0))  # Ruiz drives row inf-norms toward 1
  x = np.random.default_rng(0).normal(size=25)
  np.testing.assert_allclose((x / e) * e, x, rtol=1e-14)


# --- planted mistakes: the golden trajectories must catch them --------------
def _fixture_mismatch(monkeypatch_target, attr, bad, key="dense_n40_m60_z10_s1_eq1"):
  case = TRAJ[key]
  a, b, c, z, p = golden_cases.inputs(case["spec"])
  good = getattr(monkeypatch_target, attr)
  setattr(monkeypatch_target, attr, bad(good))
  try:
    r = O.solve_qp(a, b, c, z, p, equilibrate=case["equilibrate"])
  finally:
    setattr(monkeypatch_target, attr, good)
  mus = np.array([st["mu"] for st in r["stats"]])
  ref = np.array([t["mu"] for t in case["trajectory"]])
  k = min(len(mus), len(ref))
  return (r["iters"] != case["iters"] or r["status"] != case["status"]
          or not np.allclose(mus[:k], ref[:k], rtol=1e-6))


def test_planted_sigma_exponent_is_caught():
  def bad(good):
    def f(*args):
      s = good(*args)
      return s ** (2.0 / 3.0)  # (mu_aff/mu)^2 instead of ^3
    return f
  assert _fixture_mismatch(O, "sigma_mehrotra", bad)


def test_planted_tau_sign_is_caught():
  def bad(good):
    return lambda n, q, kq, kr, p, mu, mt, rt: good(n, q, kq, kr, p, mu, mt, -rt)
  assert _fixture_mismatch(O, "tau_quadratic", bad)


def test_planted_step_scale_is_caught():
  def bad(good):
    return lambda y, s, dy, ds, z: 0.9 * good(y, s, dy, ds, z)
  assert _fixture_mismatch(O, "step_size", bad)


def test_oracle_start_hook():
  """The oracle's start argument (the warm-start hook of the batched engine):
  the default point passed explicitly reproduces the default-start solve bit
  for bit, and a start at the solution (y, s moved off zero by 1e-6, tau = 1)
  ends, at tight tolerances, in at most half the iterations at the same
  solution."""
  qp = problems.random_dense_qp(30, 45, 8, seed=4)
  ref = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
  m = qp.a.shape[0]
  y0 = np.zeros(m)
  s0 = np.zeros(m)
  y0[qp.z:] = 1.0
  s0[qp.z:] = 1.0
  same = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p, start=(np.zeros(qp.a.shape[1]), y0, s0, 1.0))
  assert same["iters"] == ref["iters"]
  for k in "xys":
    np.testing.assert_array_equal(same[k], ref[k])
  ys = ref["y"].copy()
  ss = ref["s"].copy()
  ys[qp.z:] = np.maximum(ys[qp.z:], 1e-6)
  ss[qp.z:] = np.maximum(ss[qp.z:], 1e-6)
  tight = dict(atol=1e-10, rtol=1e-10)
  ref_t = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p, **tight)
  warm = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p, start=(ref["x"], ys, ss, 1.0), **tight)
  assert warm["status"] == "solved"
  assert warm["iters"] <= ref_t["iters"] // 2, (warm["iters"], ref_t["iters"])
  for k in "xys":
    assert _relerr(warm[k], ref_t[k]) < 1e-6, k
  with pytest.raises(ValueError, match="wrong sizes"):
    O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p, start=(np.zeros(3), y0, s0, 1.0))
