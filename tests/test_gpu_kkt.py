"""GPU parity of the single-instance HIP KKT engine against the CPU oracle.

Bars: KKT block equations to 1e-10 (tests/test_qtqp.py:740-771), matvec to
1e-12 of the assembled K, end-to-end QTQP solves with the same status and
x/y/s within 1e-6 relative of the oracle (FP64 solves on identical inputs).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import qtqp_amd
from qtqp_amd import direct, problems
from oracle import qtqp_oracle as O
import golden_cases
import qp_generators as G

pytestmark = pytest.mark.gpu

BACKENDS = [direct.HipDenseSolver, direct.HipDenseFp64Solver]
ENUMS = [qtqp_amd.LinearSolver.HIP_DENSE, qtqp_amd.LinearSolver.HIP_DENSE_FP64]
TRAJ = golden_cases.load_trajectories()


@pytest.fixture(params=["device", "host"])
def driver(request, monkeypatch):
  """QTQP.solve with the HIP backends either runs the whole iteration on the
  device (the default) or the reference's host loop over DirectKktSolver."""
  monkeypatch.setattr(direct._HipDenseBase, "device_driver", request.param == "device")
  return request.param


def _rel(u, v):
  return float(np.max(np.abs(u - v)) / max(1.0, np.max(np.abs(v))))


def _kkt_inputs(m, n, z, seed, density=0.3):
  a, b, c, p = G.gen_feasible(m, n, z, seed, density=density)
  rng = np.random.default_rng(seed + 7)
  mu = rng.uniform(0.05, 1.0)
  s = rng.uniform(size=m)
  y = rng.uniform(size=m)
  s[:z] = 0.0
  return a, b, c, p, mu, s, y


class _HostLoop(direct.LinearSolver):
  """Hides solve_refined so DirectKktSolver runs its host refinement loop
  over the HIP backend's protocol methods (the reference's call pattern)."""

  def __init__(self, inner):
    self.inner = inner

  def set_dims(self, n, m, z):
    self.inner.set_dims(n, m, z)

  def set_kkt(self, kkt):
    self.inner.set_kkt(kkt)

  def update_diag(self, diag):
    self.inner.update_diag(diag)

  def factorize(self):
    self.inner.factorize()

  def solve(self, rhs):
    return self.inner.solve(rhs)

  def __matmul__(self, x):
    return self.inner @ x

  def format(self):
    return "csr"

  def free(self):
    self.inner.free()


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("mnz", [(6, 4, 1), (150, 100, 10), (700, 330, 40)])
@pytest.mark.parametrize("host_loop", [False, True])
def test_kkt_block_equations(require_gpu, backend, mnz, host_loop):
  m, n, z = mnz
  a, b, c, p, mu, s, y = _kkt_inputs(m, n, z, seed=842 + n)
  be = backend()
  kkt = direct.DirectKktSolver(a=sp.csc_matrix(a), p=sp.csc_matrix(p), z=z,
                               min_static_regularization=1e-8,
                               max_iterative_refinement_steps=10, atol=1e-12, rtol=1e-12,
                               solver=_HostLoop(be) if host_loop else be)
  kkt.update(mu=mu, s=s, y=y)
  sol, st = kkt.solve(rhs=np.concatenate([c, b]), warm_start=np.zeros(n + m))
  kkt.free()
  d = np.concatenate([np.zeros(z), s[z:] / y[z:]])
  np.testing.assert_allclose(p @ sol[:n] + mu * sol[:n] + a.T @ sol[n:], c, atol=1e-10, rtol=1e-10)
  np.testing.assert_allclose(-a @ sol[:n] + (d + mu) * sol[n:], b, atol=1e-10, rtol=1e-10)
  assert st["status"] == "converged"


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("mn", [(5, 3), (90, 70), (300, 129)])
def test_backend_solve_and_matvec_match_numpy(require_gpu, backend, mn):
  m, n = mn
  rng = np.random.default_rng(m)
  a = rng.standard_normal((m, n))
  r = rng.standard_normal((n, n))
  p = r @ r.T / n
  dx = np.diag(p) + rng.uniform(0.1, 1.0, n)
  dy = -rng.uniform(0.05, 2.0, m)
  k = np.block([[p - np.diag(np.diag(p)) + np.diag(dx), a.T], [a, np.diag(dy)]])
  be = backend()
  be.set_dims(n, m, 0)
  be.set_dense_blocks(a, p)
  be.update_diag(np.concatenate([dx, dy]))
  be.factorize()
  v = rng.standard_normal(n + m)
  np.testing.assert_allclose(be @ v, k @ v, rtol=1e-12, atol=1e-12)
  sol = be.solve(v)
  ref = np.linalg.solve(k, v)
  assert _rel(sol, ref) < 1e-10
  assert np.linalg.norm(k @ sol - v, np.inf) < 1e-10 * (1 + np.linalg.norm(v, np.inf))
  info = be.info()
  assert info["factorizations"] == 1 and info["backend_solves"] >= 1
  be.free()
  be.free()  # idempotent


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("key", sorted(TRAJ))
def test_end_to_end_matches_reference_fixtures(require_gpu, driver, linear_solver, key):
  case = TRAJ[key]
  a, b, c, z, p = golden_cases.inputs(case["spec"])
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      linear_solver=linear_solver, equilibrate=case["equilibrate"], verbose=False,
      collect_stats=True)
  assert sol.status.value == case["status"]
  assert abs(len(sol.stats) - case["iters"]) <= 1
  ref = O.solve_qp(a, b, c, z, p, equilibrate=case["equilibrate"])
  unique = case["spec"]["kind"] in ("readme", "dense", "mpc") or case["status"] != "solved"
  for k in "xys":
    got, want = getattr(sol, k), ref[k]
    assert np.array_equal(np.isnan(got), np.isnan(want)), k
    ok = ~np.isnan(want)
    if ok.any():
      assert _rel(got[ok], want[ok]) < (1e-6 if unique else 1e-4), k


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("kind", ["feasible", "infeasible", "unbounded"])
@pytest.mark.parametrize("seed", [42, 43])
def test_reference_problem_families(require_gpu, driver, linear_solver, kind, seed):
  m, n, z = 150, 100, 10
  gen = {"feasible": G.gen_feasible, "infeasible": G.gen_infeasible,
         "unbounded": G.gen_unbounded}[kind]
  a, b, c, p = gen(m, n, z, seed)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      linear_solver=linear_solver, verbose=False)
  if kind == "feasible":
    G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, z)
  elif kind == "infeasible":
    assert sol.status == qtqp_amd.SolutionStatus.INFEASIBLE
    G.check_infeasible(sol.x, sol.y, sol.s, a, b, z)
  else:
    assert sol.status == qtqp_amd.SolutionStatus.UNBOUNDED
    G.check_unbounded(sol.x, sol.y, sol.s, a, c, p, z)


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_dense_generator_planted_optimum(require_gpu, driver, linear_solver):
  qp = problems.random_dense_qp(200, 300, 50, seed=0)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(qp.a), b=qp.b, c=qp.c, z=qp.z, p=sp.csc_matrix(qp.p)).solve(
      linear_solver=linear_solver, verbose=False, atol=1e-10, rtol=1e-11)
  assert sol.status == qtqp_amd.SolutionStatus.SOLVED
  assert _rel(sol.x, qp.x_star) < 1e-6 and _rel(sol.y, qp.y_star) < 1e-6
  assert _rel(sol.s, qp.s_star) < 1e-5


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_condensed_mpc_matches_oracle(require_gpu, driver, linear_solver):
  qp = problems.mpc_qp(32, 8, 3, seed=11)
  ref = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(qp.a), b=qp.b, c=qp.c, z=qp.z, p=sp.csc_matrix(qp.p)).solve(
      linear_solver=linear_solver, verbose=False)
  assert sol.status.value == ref["status"]
  for k in "xys":
    assert _rel(getattr(sol, k), ref[k]) < 1e-6, k


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("mnzs", [(30, 20, 5, 42), (150, 100, 10, 7)])
def test_min_iterative_refinement_steps(require_gpu, driver, linear_solver, mnzs):
  # tests/test_qtqp.py:1862-1873: one backend solve per KKT solve.  The
  # (150, 100, 10, 7) instance hits max_iter with the FP64 oracle too; the
  # bar is the oracle's status.
  m, n, z, seed = mnzs
  a, b, c, p = G.gen_feasible(m, n, z, seed)
  ref = O.solve_qp(a, b, c, z, p, max_iterative_refinement_steps=1)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p)).solve(
      linear_solver=linear_solver, max_iterative_refinement_steps=1, verbose=False)
  assert sol.status.value == ref["status"]
  if ref["status"] == "solved":
    G.check_solution(sol.x, sol.y, sol.s, a, b, c, p, z)


def test_protocol_errors(require_gpu):
  be = direct.HipDenseFp64Solver()
  be.set_dims(3, 4, 0)
  be.set_dense_blocks(np.ones((4, 3)), np.eye(3))
  with pytest.raises(RuntimeError):
    be.factorize()  # before update_diag
  with pytest.raises(ValueError):
    be.update_diag(np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0]))  # y-block not negative
  be.free()


def test_full_size_dense_qp_matches_oracle(require_gpu):
  """The n=4000, m=6000, z=1000 dense QP (a 10000 x 10000 KKT system) through the
  drop-in driver with the mixed-precision backend: oracle status and x/y/s."""
  qp = problems.random_dense_qp(4000, 6000, 1000, seed=0)
  ref = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(qp.a), b=qp.b, c=qp.c, z=qp.z, p=sp.csc_matrix(qp.p)).solve(
      linear_solver=qtqp_amd.LinearSolver.HIP_DENSE, verbose=False)
  assert sol.status.value == ref["status"] == "solved"
  for k in "xys":
    assert _rel(getattr(sol, k), ref[k]) < 1e-6, k


STAT_KEYS = {"iter", "ctx", "bty", "pcost", "dcost", "pres", "dres", "gap", "pinfeas", "dinfeas",
             "dinfeas_a", "dinfeas_p", "mu", "sigma", "alpha", "tau", "norm_x", "norm_y", "status",
             "time", "prelrhs", "drelrhs", "complementarity", "norm_s", "max_sy", "min_sy",
             "std_sy", "max_s_over_y", "min_s_over_y", "mean_s_over_y", "std_s_over_y",
             "q_lin_sys_stats", "predictor_lin_sys_stats", "corrector_lin_sys_stats"}


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
@pytest.mark.parametrize("equilibrate", [True, False])
def test_device_driver_tracks_host_driver(require_gpu, monkeypatch, linear_solver, equilibrate):
  """The device-resident loop and the host loop over the same kernels take the
  same path: same statuses and iteration counts, per-iteration mu, tau,
  sigma and alpha to 1e-6 while mu is above 1e-6, the same stats keys, and
  IR solve counts that differ only where a residual sits at its tolerance."""
  a, b, c, p = G.gen_feasible(150, 100, 10, 42)
  runs = {}
  for dev in (True, False):
    monkeypatch.setattr(direct._HipDenseBase, "device_driver", dev)
    runs[dev] = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=10, p=sp.csc_matrix(p)).solve(
        linear_solver=linear_solver, equilibrate=equilibrate, verbose=False, collect_stats=True)
  d, h = runs[True], runs[False]
  assert d.status == h.status == qtqp_amd.SolutionStatus.SOLVED
  assert len(d.stats) == len(h.stats)
  for sd, sh in zip(d.stats, h.stats):
    assert set(sd) == set(sh) == STAT_KEYS
    assert set(sd["predictor_lin_sys_stats"]) == set(sh["predictor_lin_sys_stats"])
    if sh["mu"] > 1e-6:
      for k in ("mu", "tau", "sigma", "alpha", "pcost", "dcost"):
        assert abs(sd[k] - sh[k]) <= 1e-6 * max(1.0, abs(sh[k])), (sh["iter"], k)
  for k in "xys":
    assert _rel(getattr(d, k), getattr(h, k)) < 1e-6, k
  G.check_solution(d.x, d.y, d.s, a, b, c, p, 10)


def test_device_driver_verbose_log(require_gpu, capsys):
  """Banner, header, one line per iteration and the footer, as the host loop prints."""
  a, b, c, p = G.gen_feasible(30, 20, 5, 3)
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=5, p=sp.csc_matrix(p)).solve(
      linear_solver=qtqp_amd.LinearSolver.HIP_DENSE, verbose=True, collect_stats=True)
  out = capsys.readouterr().out.splitlines()
  assert out[0].startswith("| QTQP v") and "linear_solver=HIP_DENSE" in out[0]
  rows = [ln for ln in out if ln.startswith("|") and ln[2:6].strip().isdigit()]
  assert len(rows) == len(sol.stats)
  assert out[-1] == "| Solved"


def test_device_driver_without_equality_rows_and_p(require_gpu):
  """z = 0 and P = 0 (an LP): the device loop skips every P product."""
  rng = np.random.default_rng(5)
  m, n = 60, 25
  a = rng.standard_normal((m, n))
  x0 = rng.standard_normal(n)
  s0 = np.abs(rng.standard_normal(m))
  y0 = np.where(rng.uniform(size=m) < 0.5, np.abs(rng.standard_normal(m)), 0.0)
  s0[y0 > 0] = 0.0
  b = a @ x0 + s0
  c = -a.T @ y0
  ref = O.solve_qp(a, b, c, 0, np.zeros((n, n)))
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=0).solve(
      linear_solver=qtqp_amd.LinearSolver.HIP_DENSE_FP64, verbose=False)
  assert sol.status.value == ref["status"]
  if ref["status"] == "solved":
    assert abs(c @ sol.x - c @ ref["x"]) <= 1e-6 * max(1.0, abs(c @ ref["x"]))


@pytest.mark.parametrize("linear_solver", ENUMS, ids=lambda e: e.name)
def test_device_driver_hit_max_iter_matches_host_loop(require_gpu, monkeypatch, linear_solver):
  """max_iter reached on the device loop: HIT_MAX_ITER in the status and the last
  stats entry, and the tau-scaled iterate of the host loop after the same
  iterations (tests/test_qtqp.py hit-max-iter behaviour)."""
  a, b, c, p = G.gen_feasible(60, 40, 5, 1)
  sols = {}
  for dev in (True, False):
    monkeypatch.setattr(direct._HipDenseBase, "device_driver", dev)
    sols[dev] = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=5, p=sp.csc_matrix(p)).solve(
        linear_solver=linear_solver, max_iter=3, verbose=False, collect_stats=True)
  for sol in sols.values():
    assert sol.status == qtqp_amd.SolutionStatus.HIT_MAX_ITER
    assert len(sol.stats) == 3 and sol.stats[-1]["status"] == sol.status
  for k in "xys":
    assert _rel(getattr(sols[True], k), getattr(sols[False], k)) < 1e-8, k


def test_device_driver_presolve(require_gpu):
  """Rows with b >= 1e20 are dropped before the device loop and restored after
  it: the infeasibility certificate and the solved case's slack
  (tests/test_qtqp.py:561-596)."""
  rng = np.random.default_rng(1842)
  a, b, c, p = G.gen_infeasible(30, 20, 4, 1842)
  a_full = np.vstack([a, rng.normal(size=(3, 20))])
  b_full = np.concatenate([b, np.full(3, 1e21)])
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(a_full), b=b_full, c=c, z=4, p=sp.csc_matrix(p)).solve(
      linear_solver=qtqp_amd.LinearSolver.HIP_DENSE, verbose=False)
  assert sol.status == qtqp_amd.SolutionStatus.INFEASIBLE
  G.check_infeasible(sol.x, sol.y, sol.s, a_full, b_full, 4)
  a, b, c, p = G.gen_feasible(30, 20, 4, 5)
  extra = rng.normal(size=(2, 20))
  sol = qtqp_amd.QTQP(a=sp.csc_matrix(np.vstack([a, extra])), b=np.concatenate([b, [1e20, np.inf]]),
                      c=c, z=4, p=sp.csc_matrix(p)).solve(
      linear_solver=qtqp_amd.LinearSolver.HIP_DENSE, verbose=False)
  assert sol.status == qtqp_amd.SolutionStatus.SOLVED
  np.testing.assert_allclose(sol.y[-2:], 0.0)
  np.testing.assert_allclose(sol.s[-2:][0], 1e20 - extra[0] @ sol.x)


def test_device_ipm_argument_errors(require_gpu):
  """The device loop's C ABI rejects bad arguments with QTQP_ERR_ARG (ValueError)."""
  lib = direct.HipDenseSolver()._lib
  a, b, c = np.ones((4, 3)), np.ones(4), np.ones(3)
  with pytest.raises(ValueError):
    direct.DeviceIpm(lib, 0, a, None, False, b, c, 4, True)   # z must be < m
  with pytest.raises(ValueError):
    direct.DeviceIpm(lib, 7, a, None, False, b, c, 1, True)   # unknown precision
  ipm = direct.DeviceIpm(lib, 0, a, np.eye(3), True, b, c, 1, True)
  opts = qtqp_amd._native.IpmOpts(atol=1e-7, rtol=1e-8, atol_infeas=1e-8, rtol_infeas=1e-9,
                                  step_size_scale=0.99, min_static_regularization=1e-8,
                                  linear_solver_atol=1e-12, linear_solver_rtol=1e-12,
                                  max_iterative_refinement_steps=0, collect_stats=0)
  with pytest.raises(ValueError):
    ipm.iterate(opts)                                            # refinement steps >= 1
  ipm.free()
  ipm.free()  # idempotent


def _solve_with_graph(monkeypatch, linear_solver, qp, use_graph, steps=0, **kw):
  backend = linear_solver.value
  monkeypatch.setattr(backend, "device_graph", use_graph)
  monkeypatch.setattr(backend, "device_graph_steps", steps)
  return qtqp_amd.QTQP(a=sp.csc_matrix(qp[0]), b=qp[1], c=qp[2], z=qp[3],
                       p=sp.csc_matrix(qp[4])).solve(linear_solver=linear_solver, verbose=False,
                                                     collect_stats=True, **kw)


@pytest.mark.parametrize("steps", [0, 1])
@pytest.mark.parametrize("kind", ["feasible", "infeasible", "mpc"])
def test_graph_iteration_matches_host_driven_fp64(require_gpu, monkeypatch, kind, steps):
  """With the FP64 factor, the device-loop iteration replayed from a HIP graph
  (refinement and stopping decided on the device) runs the same kernels on the
  same data as the host-driven launches, so every iterate, statistic and
  status is bitwise equal.  steps=1 holds a single refinement step per solve
  in the graph, so most iterations are handed back and redone host-driven
  from the saved start, which must not change anything either."""
  if kind == "mpc":
    q = problems.mpc_qp(16, 6, 2, seed=4)
    qp = (q.a, q.b, q.c, q.z, q.p)
  else:
    gen = G.gen_feasible if kind == "feasible" else G.gen_infeasible
    a, b, c, p = gen(150, 100, 10, 42)
    qp = (a, b, c, 10, p)
  ls = qtqp_amd.LinearSolver.HIP_DENSE_FP64
  eager = _solve_with_graph(monkeypatch, ls, qp, False)
  graph = _solve_with_graph(monkeypatch, ls, qp, True, steps)
  assert graph.status == eager.status
  assert len(graph.stats) == len(eager.stats)
  for k in "xys":
    np.testing.assert_array_equal(getattr(graph, k), getattr(eager, k))
  for sg, se in zip(graph.stats, eager.stats):
    for key in ("mu", "sigma", "alpha", "tau", "pres", "dres", "gap"):
      assert sg[key] == se[key], (sg["iter"], key)
    for lin in ("q_lin_sys_stats", "predictor_lin_sys_stats", "corrector_lin_sys_stats"):
      for key in ("solves", "status", "final_residual_norm", "tolerance"):
        assert sg[lin][key] == se[lin][key], (sg["iter"], lin, key)


def test_graph_iteration_mixed_precision_matches_oracle(require_gpu, monkeypatch):
  """The mixed handle in the graph (FP32 chain and solves gated on the device,
  FP64 chain opened on a breakdown or when FP32 refinement contracts slowly):
  oracle status and x/y/s, on a QP where FP32 holds and on one where it falls
  back, with the handle's counters telling which ran."""
  for qp in (problems.mpc_qp(32, 8, 3, seed=11), problems.random_dense_qp(200, 300, 50, seed=0)):
    ref = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
    sol = _solve_with_graph(monkeypatch, qtqp_amd.LinearSolver.HIP_DENSE,
                            (qp.a, qp.b, qp.c, qp.z, qp.p), True)
    assert sol.status.value == ref["status"]
    for k in "xys":
      assert _rel(getattr(sol, k), ref[k]) < 1e-6, k
