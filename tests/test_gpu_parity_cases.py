"""More parity cases on the GPU.

- The condensed MPC QP with T=64, nx=12, nu=4 (n=256, m=2048, z=0) through both
  dense HIP backends, on the device loop and on the host loop: the FP64
  oracle's status and x/y/s within 1e-6 relative.
- The reference's keyword cases (tests/test_qtqp.py): min_static_regularization
  (2003-2025), step_size_scale (2118-2132), infeasibility tolerances
  (2137-2150), linear-solver tolerances (2197-2210), re-solve with the
  equilibrate flag flipped (1935-1952), stats monotonicity (2046-2079) and
  complementarity at convergence (1960-1971), with the same seeds, shapes and
  assertions, on the device loop and the host loop.  The same keywords go
  through the batched engine too.
- Presolve certificates in the batched engine (tests/test_qtqp.py:561-596):
  rows with b >= 1e20 dropped before the device IPM and restored in the
  infeasible and unbounded certificates and in the solved case's slack.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import qtqp_amd
from qtqp_amd import batch, direct, problems
from oracle import qtqp_oracle as O
import qp_generators as G

pytestmark = pytest.mark.gpu

ENUMS = [qtqp_amd.LinearSolver.HIP_DENSE, qtqp_amd.LinearSolver.HIP_DENSE_FP64]
SOLVED = qtqp_amd.SolutionStatus.SOLVED
INFEASIBLE = qtqp_amd.SolutionStatus.INFEASIBLE
UNBOUNDED = qtqp_amd.SolutionStatus.UNBOUNDED


@pytest.fixture(params=["device", "host"])
def driver(request, monkeypatch):
  """QTQP.solve with a dense HIP backend runs the whole iteration on the device
  (the default) or the reference's host loop over DirectKktSolver."""
  monkeypatch.setattr(direct._HipDenseBase, "device_driver", request.param == "device")
  return request.param


def _rel(u, v):
  ok = ~np.isnan(v)
  if not ok.any():
    return 0.0
  return float(np.max(np.abs(u[ok] - v[ok])) / max(1.0, np.max(np.abs(v[ok]))))


def _qtqp(a, b, c, z, p):
  return qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z,
                       p=None if p is None else sp.csc_matrix(p))


# -- condensed MPC, T=64, nx=12, nu=4 ----------------------------------------

@pytest.fixture(scope="module")
def mpc64():
  qp = problems.mpc_qp(64, 12, 4, seed=0)
  assert (qp.n, qp.m, qp.z) == (256, 2048, 0)
  return qp, O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_mpc_t64_matches_oracle(require_gpu, driver, linear_solver, mpc64):
  qp, ref = mpc64
  print(f"[case] mpc T=64 nx=12 nu=4 {linear_solver.name} {driver}", flush=True)
  sol = _qtqp(qp.a, qp.b, qp.c, qp.z, qp.p).solve(linear_solver=linear_solver, verbose=False,
                                                  collect_stats=True)
  assert sol.status.value == ref["status"] == "solved"
  assert abs(len(sol.stats) - ref["iters"]) <= 1
  for k in "xys":
    assert _rel(getattr(sol, k), ref[k]) < 1e-6, k
  G.check_solution(sol.x, sol.y, sol.s, qp.a, qp.b, qp.c, qp.p, qp.z)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_mpc_t64_in_batch_matches_oracle(require_gpu, precision, mpc64):
  qp, ref = mpc64
  bq = batch.BatchQP([qp, problems.mpc_qp(8, 4, 2, seed=1)])
  sol = bq.solve(precision=precision).solutions[0]
  bq.free()
  assert sol.status.value == ref["status"]
  assert abs(sol.iters - ref["iters"]) <= 1
  for k in "xys":
    assert _rel(getattr(sol, k), ref[k]) < 1e-6, k


# -- the reference's keyword cases ---------------------------------------------

@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("reg", [0.0, 1e-4, 1e-8])
def test_min_static_regularization(require_gpu, driver, linear_solver, reg):
  a, b, c, p = G.gen_feasible(30, 20, 5, 42)
  sol = _qtqp(a, b, c, 5, p).solve(linear_solver=linear_solver, min_static_regularization=reg,
                                   max_iterative_refinement_steps=50, verbose=False)
  assert sol.status == SOLVED
  G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, 5)
  ref = O.solve_qp(a, b, c, 5, p, min_static_regularization=reg,
                   max_iterative_refinement_steps=50)
  assert ref["status"] == "solved"


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("scale", [0.5, 0.9, 0.99])
def test_step_size_scale(require_gpu, driver, linear_solver, scale):
  a, b, c, p = G.gen_feasible(30, 20, 5, 42)
  sol = _qtqp(a, b, c, 5, p).solve(linear_solver=linear_solver, step_size_scale=scale,
                                   verbose=False, collect_stats=True)
  assert sol.status == SOLVED
  G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, 5)
  ref = O.solve_qp(a, b, c, 5, p, step_size_scale=scale)
  assert abs(len(sol.stats) - ref["iters"]) <= 1


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("atol_infeas,rtol_infeas", [(1e-4, 1e-5), (1e-10, 1e-11)])
def test_infeasibility_tolerances(require_gpu, driver, linear_solver, atol_infeas, rtol_infeas):
  a, b, c, p = G.gen_infeasible(150, 100, 10, 142)
  sol = _qtqp(a, b, c, 10, p).solve(linear_solver=linear_solver, atol_infeas=atol_infeas,
                                    rtol_infeas=rtol_infeas, verbose=False)
  assert sol.status == INFEASIBLE
  G.check_infeasible(sol.x, sol.y, sol.s, a, b, 10, atol=atol_infeas, rtol=rtol_infeas)


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_linear_solver_tolerances(require_gpu, driver, linear_solver):
  a, b, c, p = G.gen_feasible(30, 20, 5, 42)
  sol = _qtqp(a, b, c, 5, p).solve(linear_solver=linear_solver, linear_solver_atol=1e-6,
                                   linear_solver_rtol=1e-6, verbose=False)
  assert sol.status == SOLVED
  G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, 5)


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_resolve_flip_equilibrate(require_gpu, driver, linear_solver):
  a, b, c, p = G.gen_feasible(50, 30, 5, 42)
  solver = _qtqp(a, b, c, 5, p)
  sol1 = solver.solve(linear_solver=linear_solver, equilibrate=True, verbose=False)
  sol2 = solver.solve(linear_solver=linear_solver, equilibrate=False, verbose=False)
  for sol in (sol1, sol2):
    assert sol.status == SOLVED
    G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, 5)
  obj = [c @ s.x + 0.5 * s.x @ p @ s.x for s in (sol1, sol2)]
  np.testing.assert_allclose(obj[0], obj[1], atol=1e-5, rtol=1e-5)


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_stats_monotonicity(require_gpu, driver, linear_solver):
  a, b, c, p = G.gen_feasible(50, 30, 5, 42)
  sol = _qtqp(a, b, c, 5, p).solve(linear_solver=linear_solver, verbose=False,
                                   collect_stats=True)
  assert sol.status == SOLVED and len(sol.stats) >= 2
  mus = [s["mu"] for s in sol.stats]
  times = [s["time"] for s in sol.stats]
  assert mus[0] > 0.0
  assert all(mus[i] < mus[i - 1] for i in range(1, len(mus)))
  assert all(times[i] >= times[i - 1] for i in range(1, len(times)))
  assert all(0 < s["alpha"] <= 1.0 for s in sol.stats)
  first, last = sol.stats[0], sol.stats[-1]
  assert last["pres"] < first["pres"] and last["dres"] < first["dres"]
  assert last["gap"] < first["gap"]


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_complementarity_at_convergence(require_gpu, driver, linear_solver):
  a, b, c, p = G.gen_feasible(30, 20, 5, 42)
  sol = _qtqp(a, b, c, 5, p).solve(linear_solver=linear_solver, verbose=False,
                                   collect_stats=True)
  assert sol.status == SOLVED
  assert sol.stats[-1]["complementarity"] < 1e-6


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_keyword_options(require_gpu, precision):
  """The keyword cases above through the batched engine: each option set on a
  batch of the reference's instance, statuses and KKT checks as there."""
  a, b, c, p = G.gen_feasible(30, 20, 5, 42)
  cases = [dict(min_static_regularization=r, max_iterative_refinement_steps=50)
           for r in (0.0, 1e-4, 1e-8)]
  cases += [dict(step_size_scale=s) for s in (0.5, 0.9, 0.99)]
  cases += [dict(linear_solver_atol=1e-6, linear_solver_rtol=1e-6), dict(equilibrate=False)]
  bq = batch.BatchQP([(a, b, c, 5, p)] * 3)
  for kw in cases:
    for sol in bq.solve(precision=precision, **kw).solutions:
      assert sol.status == SOLVED, kw
      G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, 5)
  bq.free()
  ai, bi, ci, pi = G.gen_infeasible(150, 100, 10, 142)
  bq = batch.BatchQP([(ai, bi, ci, 10, pi)])
  for at, rt in ((1e-4, 1e-5), (1e-10, 1e-11)):
    sol = bq.solve(precision=precision, atol_infeas=at, rtol_infeas=rt).solutions[0]
    assert sol.status == INFEASIBLE
    G.check_infeasible(sol.x, sol.y, sol.s, ai, bi, 10, atol=at, rtol=rt)
  bq.free()


# -- presolve certificates in the batched engine --------------------------------

@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_presolve_certificates(require_gpu, precision):
  """tests/test_qtqp.py:561-596 through BatchQP: dropped rows do not corrupt
  the infeasible certificate or the kept part of the unbounded one, and a
  solved instance gets back y = 0 and its slack on the dropped rows."""
  rng = np.random.default_rng(1842)
  ai, bi, ci, pi = G.gen_infeasible(30, 20, 4, rng)
  ai, bi = G.append_dropped_inequalities(ai, bi, 3, rng, 1e21)
  rng = np.random.default_rng(1942)
  au, bu, cu, pu = G.gen_unbounded(30, 20, 4, rng)
  au_base = au
  au, bu = G.append_dropped_inequalities(au, bu, 3, rng, 1e21)
  rng = np.random.default_rng(5)
  af, bf, cf, pf = G.gen_feasible(30, 20, 4, rng)
  extra = rng.normal(size=(2, 20))
  af = np.vstack([af, extra])
  bf = np.concatenate([bf, [1e20, np.inf]])
  bq = batch.BatchQP([(ai, bi, ci, 4, pi), (au, bu, cu, 4, pu), (af, bf, cf, 4, pf)])
  inf, unb, sol = bq.solve(precision=precision).solutions
  bq.free()
  assert inf.status == INFEASIBLE
  G.check_infeasible(inf.x, inf.y, inf.s, ai, bi, 4)
  np.testing.assert_array_equal(inf.y[-3:], 0.0)
  assert unb.status == UNBOUNDED
  assert np.all(np.isnan(unb.y)) and np.all(np.isnan(unb.s[-3:]))
  np.testing.assert_allclose(cu @ unb.x, -1.0, atol=1e-8, rtol=1e-9)
  assert np.min(unb.s[4:30], initial=0.0) > -1e-9
  tol = 1e-8 + 1e-9 * np.linalg.norm(unb.x, np.inf)
  assert np.linalg.norm(au_base @ unb.x + unb.s[:30], np.inf) < tol
  assert np.linalg.norm(pu @ unb.x, np.inf) < tol
  assert sol.status == SOLVED
  G.check_solution(sol.x, sol.y[:30], sol.s[:30], af[:30], bf[:30], cf, pf, 4)
  np.testing.assert_array_equal(sol.y[-2:], 0.0)
  np.testing.assert_allclose(sol.s[-2], 1e20 - extra[0] @ sol.x)
  assert sol.s[-1] == np.inf
  # the single-QP drop-in on the device loop gives the same certificates
  for a, b, c, p, want in ((ai, bi, ci, pi, INFEASIBLE), (au, bu, cu, pu, UNBOUNDED)):
    one = _qtqp(a, b, c, 4, p).solve(linear_solver=qtqp_amd.LinearSolver.HIP_DENSE, verbose=False)
    assert one.status == want
