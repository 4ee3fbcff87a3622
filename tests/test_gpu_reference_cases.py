"""The reference test suite's backend-parametrised cases, run on every HIP
backend: LPs with p=None (feasible, infeasible, unbounded; tests/test_qtqp.py:
1343-1403), the single-inequality boundary z = m-1 (1653-1663), the tiny 5x3
problem (1880-1899), the full solve/infeasible/unbounded grid (348-458,
1428-1486), and the driver's pieces over each backend: the refined KKT solve
(741-771), the resolvent (822-889), Newton steps onto the central path
(892-968) and the tau quadratic against an independent form (998-1044).  Same
seeds, generators, tolerances and assertions as there (generators restated in
tests/qp_generators.py)."""
import numpy as np
import pytest
import scipy.sparse as sp

import qtqp_amd
import qp_generators as G

pytestmark = pytest.mark.gpu

BACKENDS = [qtqp_amd.LinearSolver.HIP_DENSE, qtqp_amd.LinearSolver.HIP_DENSE_FP64,
            qtqp_amd.LinearSolver.HIP_SPARSE]


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("kind,seed0", [("feasible", 4042), ("infeasible", 4142),
                                        ("unbounded", 4242)])
@pytest.mark.parametrize("k", range(5))
@pytest.mark.parametrize("equilibrate", [True, False])
def test_lp_families(require_gpu, linear_solver, kind, seed0, k, equilibrate):
  m, n, z = 50, 30, 5
  seed = seed0 + k
  if kind == "feasible":
    a, b, c, _ = G.gen_feasible(m, n, z, seed)
  elif kind == "infeasible":
    a, b, c, _ = G.gen_infeasible(m, n, z, seed)
  else:
    a, b, c, _ = G.gen_unbounded(m, n, z, seed)
  p0 = np.zeros((n, n))
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z).solve(
      equilibrate=equilibrate, linear_solver=linear_solver, verbose=False)
  if kind == "feasible":
    assert sol.status == qtqp_amd.SolutionStatus.SOLVED
    G.check_solution(sol.x, sol.y, sol.s, a, b, c, p0, z)
  elif kind == "infeasible":
    assert sol.status == qtqp_amd.SolutionStatus.INFEASIBLE
    G.check_infeasible(sol.x, sol.y, sol.s, a, b, z)
  else:
    assert sol.status == qtqp_amd.SolutionStatus.UNBOUNDED
    G.check_unbounded(sol.x, sol.y, sol.s, a, c, p0, z)


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("seed", 5042 + np.arange(5))
def test_single_inequality_constraint(require_gpu, linear_solver, seed):
  m, n = 20, 10
  z = m - 1
  a, b, c, p = G.gen_feasible(m, n, z, int(seed))
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      linear_solver=linear_solver, verbose=False)
  assert sol.status == qtqp_amd.SolutionStatus.SOLVED
  G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, z)


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
def test_solve_tiny(require_gpu, linear_solver):
  rng = np.random.default_rng(42)
  m, n, z = 5, 3, 1
  a = rng.normal(size=(m, n))
  pd = rng.normal(size=(n, n))
  p = pd.T @ pd * 0.1
  w = rng.normal(size=m)
  y = w.copy()
  y[z:] = 0.5 * (w[z:] + np.abs(w[z:]))
  s = y - w
  x = rng.normal(size=n)
  b, c = a @ x + s, -a.T @ y
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      linear_solver=linear_solver, verbose=False)
  assert sol.status == qtqp_amd.SolutionStatus.SOLVED
  G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, z)


def _check(kind, sol, a, b, c, p, z):
  if kind == "feasible":
    assert sol.status == qtqp_amd.SolutionStatus.SOLVED
    G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, z)
  elif kind == "infeasible":
    assert sol.status == qtqp_amd.SolutionStatus.INFEASIBLE
    G.check_infeasible(sol.x, sol.y, sol.s, a, b, z)
  else:
    assert sol.status == qtqp_amd.SolutionStatus.UNBOUNDED
    G.check_unbounded(sol.x, sol.y, sol.s, a, c, p, z)


# (kind, first seed, seeds, sizes): tests/test_qtqp.py test_solve 348-364,
# test_infeasible 368-384, test_unbounded 388-404, *_large 408-458,
# test_solve_all_inequalities 1428-1446, *_small 1450-1486.
_GRID = [("feasible", 42, 10, [(150, 100, 10), (10, 5, 3), (500, 300, 30)]),
         ("infeasible", 142, 10, [(150, 100, 10), (500, 300, 30)]),
         ("unbounded", 242, 10, [(150, 100, 10), (500, 300, 30)]),
         ("feasible", 6042, 3, [(1000, 600, 60)]),
         ("infeasible", 6142, 3, [(1000, 600, 60)]),
         ("unbounded", 6242, 3, [(1000, 600, 60)]),
         ("feasible", 4342, 5, [(50, 30, 0)]),
         ("infeasible", 4442, 5, [(20, 10, 3)]),
         ("unbounded", 4542, 5, [(20, 10, 3)])]
_CASES = [(kind, s0 + k, mnz) for kind, s0, ns, sizes in _GRID for mnz in sizes for k in range(ns)]


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("equilibrate", [True, False])
@pytest.mark.parametrize("kind,seed,mnz", _CASES,
                         ids=[f"{k}-{m}x{n}x{z}-s{s}" for k, s, (m, n, z) in _CASES])
def test_reference_families_full_grid(require_gpu, linear_solver, equilibrate, kind, seed, mnz):
  m, n, z = mnz
  gen = {"feasible": G.gen_feasible, "infeasible": G.gen_infeasible,
         "unbounded": G.gen_unbounded}[kind]
  a, b, c, p = gen(m, n, z, seed)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      equilibrate=equilibrate, linear_solver=linear_solver, verbose=False)
  _check(kind, sol, a, b, c, p, z)


# -- the driver's pieces over each HIP backend (tests/test_qtqp.py:741-1044) ---

def _kkt(a, p, z, linear_solver):
  from qtqp_amd import direct
  return direct.DirectKktSolver(a=a, p=p, z=z, min_static_regularization=1e-8,
                                max_iterative_refinement_steps=10, atol=1e-12, rtol=1e-12,
                                solver=linear_solver.value())


def _iterate_case(seed):
  rng = np.random.default_rng(seed)
  m, n, z = 150, 100, 10
  a, b, c, p = G.gen_feasible(m, n, z, rng)
  return rng, m, n, z, sp.csc_matrix(a), b, c, sp.csc_matrix(p)


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("seed", 842 + np.arange(10))
def test_direct_linear_solver(require_gpu, linear_solver, seed):
  rng, m, n, z, a, b, c, p = _iterate_case(int(seed))
  mu = rng.uniform()
  s, y = rng.uniform(size=m), rng.uniform(size=m)
  s[:z] = 0.0
  d = np.concatenate([np.zeros(z), s[z:] / y[z:]])
  kkt = _kkt(a, p, z, linear_solver)
  kkt.update(mu=mu, s=s, y=y)
  sol, _ = kkt.solve(rhs=np.concatenate([c, b]), warm_start=np.zeros(n + m))
  kkt.free()
  np.testing.assert_allclose(p @ sol[:n] + mu * sol[:n] + a.T @ sol[n:], c, atol=1e-10, rtol=1e-10)
  np.testing.assert_allclose(-a @ sol[:n] + (d + mu) * sol[n:], b, atol=1e-10, rtol=1e-10)


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("seed", 942 + np.arange(20))
def test_resolvent_operator(require_gpu, linear_solver, seed):
  rng, m, n, z, a, b, c, p = _iterate_case(int(seed))
  mu, sigma = rng.uniform(), rng.uniform()
  s, y = rng.uniform(size=m), rng.uniform(size=m)
  s[:z] = 0.0
  solver = qtqp_amd.QTQP(a=a, b=b, c=c, z=z, p=p)
  solver.q = np.concatenate([c, b])
  solver._linear_solver = _kkt(a, p, z, linear_solver)  # pylint: disable=protected-access
  solver._linear_solver.update(mu=mu, s=s, y=y)  # pylint: disable=protected-access
  solver.kinv_q, _ = solver._linear_solver.solve(rhs=solver.q, warm_start=np.zeros(n + m))  # pylint: disable=protected-access
  r_anchor, tau_anchor = rng.uniform(size=n + m), rng.uniform()
  x_new, y_new, tau_new, _ = solver._newton_step(  # pylint: disable=protected-access
      p=p, mu=mu, mu_target=sigma * mu, r_anchor=r_anchor, tau_anchor=tau_anchor,
      x=r_anchor[:n], y=y, s=s, tau=1.0, correction=None)
  solver._linear_solver.free()  # pylint: disable=protected-access
  d = np.concatenate([np.zeros(z), s[z:] / y[z:]])
  np.testing.assert_allclose(p @ x_new + mu * x_new + a.T @ y_new + c * tau_new,
                             (mu - sigma * mu) * r_anchor[:n], atol=1e-10, rtol=1e-10)
  np.testing.assert_allclose(-a @ x_new + (d + mu) * y_new + b * tau_new,
                             np.concatenate([np.zeros(z), sigma * mu / y[z:] + s[z:]])
                             + (mu - sigma * mu) * r_anchor[n:], atol=1e-10, rtol=1e-10)
  np.testing.assert_allclose(-c @ x_new - b @ y_new + mu * tau_new - x_new @ (p @ x_new) / tau_new
                             - sigma * mu / tau_new, (mu - sigma * mu) * tau_anchor,
                             atol=1e-10, rtol=1e-10)


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("seed", 1042 + np.arange(10))
def test_newton_step_converges_to_central_path(require_gpu, linear_solver, seed):
  rng, m, n, z, a, b, c, p = _iterate_case(int(seed))
  mu = rng.uniform()
  s, y, x, tau = np.ones(m), np.ones(m), np.zeros(n), 1.0
  s[:z] = 0.0
  solver = qtqp_amd.QTQP(a=a, b=b, c=c, z=z, p=p)
  solver.q = np.concatenate([c, b])
  solver._linear_solver = _kkt(a, p, z, linear_solver)  # pylint: disable=protected-access
  for _ in range(20):
    solver._linear_solver.update(mu=mu, s=s, y=y)  # pylint: disable=protected-access
    solver.kinv_q, _ = solver._linear_solver.solve(rhs=solver.q, warm_start=np.zeros(n + m))  # pylint: disable=protected-access
    x_t, y_t, tau_t, _ = solver._newton_step(  # pylint: disable=protected-access
        p=p, mu=mu, mu_target=mu, r_anchor=np.zeros(n + m), tau_anchor=0.0, x=x, y=y, s=s,
        tau=tau, correction=None)
    d_x, d_y, d_tau = x_t - x, y_t - y, tau_t - tau
    d_s = np.zeros(m)
    d_s[z:] = mu / y[z:] - y_t[z:] * s[z:] / y[z:]
    step = 0.99 * solver._compute_step_size(y, s, d_y, d_s)  # pylint: disable=protected-access
    x, y, tau, s = x + step * d_x, y + step * d_y, tau + step * d_tau, s + step * d_s
    y[z:] = np.maximum(y[z:], 1e-30)
    s[z:] = np.maximum(s[z:], 1e-30)
    tau = max(tau, 1e-30)
  solver._linear_solver.free()  # pylint: disable=protected-access
  np.testing.assert_allclose(p @ x + mu * x + a.T @ y + c * tau, np.zeros(n), atol=1e-9, rtol=1e-9)
  np.testing.assert_allclose(-a @ x + mu * y + b * tau, np.concatenate([np.zeros(z), s[z:]]),
                             atol=1e-9, rtol=1e-9)
  np.testing.assert_allclose(-c @ x - b @ y + mu * tau - x @ (p @ x) / tau - mu / tau, 0.0,
                             atol=1e-9, rtol=1e-9)
  np.testing.assert_allclose(s[z:] * y[z:], mu * np.ones(m - z), atol=1e-9, rtol=1e-9)


@pytest.mark.parametrize("linear_solver", BACKENDS, ids=lambda e: e.name)
@pytest.mark.parametrize("seed", 1142 + np.arange(10))
def test_equivalent_tau_solution(require_gpu, linear_solver, seed):
  from test_oracle import _tau_alternative
  rng, m, n, z, a, b, c, p = _iterate_case(int(seed))
  mu, sigma = rng.uniform(), rng.uniform()
  mu_t = sigma * mu
  s, y = rng.uniform(size=m), rng.uniform(size=m)
  s[:z] = 0.0
  solver = qtqp_amd.QTQP(a=a, b=b, c=c, z=z, p=p)
  solver.q = np.concatenate([c, b])
  solver._linear_solver = _kkt(a, p, z, linear_solver)  # pylint: disable=protected-access
  solver._linear_solver.update(mu=mu, s=s, y=y)  # pylint: disable=protected-access
  solver.kinv_q, _ = solver._linear_solver.solve(rhs=solver.q, warm_start=np.zeros(n + m))  # pylint: disable=protected-access
  r = (mu - mu_t) * rng.uniform(size=n + m)
  r[n + z:] += mu_t / y[z:] + s[z:]
  kinv_r, _ = solver._linear_solver.solve(rhs=r, warm_start=np.zeros(n + m))  # pylint: disable=protected-access
  r_tau = (mu - mu_t) * rng.uniform()
  tau = solver._solve_for_tau(p, kinv_r, mu, mu_t, r_tau)  # pylint: disable=protected-access
  solver._linear_solver.free()  # pylint: disable=protected-access
  tau_alt = _tau_alternative(n, solver.kinv_q, kinv_r, mu, mu_t, r, r_tau, s, y)
  np.testing.assert_allclose(tau, tau_alt, atol=1e-11, rtol=1e-11)
