"""The reference test suite's backend-parametrised cases, run on every HIP
backend: LPs with p=None (feasible, infeasible, unbounded; tests/test_qtqp.py:
1343-1403), the single-inequality boundary z = m-1 (1653-1663) and the tiny
5x3 problem (1880-1899).  Same seeds, generators and assertions as there
(restated in tests/qp_generators.py)."""
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
