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
