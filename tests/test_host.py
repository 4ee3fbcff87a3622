"""Host logic of qtqp_amd on the CPU (no GPU): the QTQP driver, presolve,
status/certificate handling, stats keys, logging and the LinearSolver
protocol, driven through a test-only backend that wraps the oracle's dense
Gram solver (the product itself has no CPU backend)."""
import numpy as np
import pytest
import scipy.sparse as sp

import qtqp_amd
from qtqp_amd import direct
from oracle import qtqp_oracle as O
import golden_cases
import qp_generators as G

TRAJ = golden_cases.load_trajectories()


class OracleBackend(direct.LinearSolver):
  """Test-only LinearSolver delegating to the oracle's dense Gram backend."""

  def set_dims(self, n, m, z):
    self.n = n

  def set_kkt(self, kkt):
    super().set_kkt(kkt)
    dense = kkt.toarray()
    n = self.n
    pu = dense[:n, :n]
    self.be = O.DenseGramBackend(dense[:n, n:].T, pu + pu.T - np.diag(np.diag(pu)))

  def update_diag(self, diag):
    super().update_diag(diag)
    self.be.update_diag(diag)

  def factorize(self):
    self.be.factorize()

  def solve(self, rhs):
    return self.be.solve(rhs)

  def __matmul__(self, x):
    return self.be.matvec(x)

  def format(self):
    return "csr"


@pytest.fixture
def oracle_backend(monkeypatch):
  monkeypatch.setattr(qtqp_amd, "_resolve_linear_solver",
                      lambda ls, *shape: (qtqp_amd.LinearSolver.HIP_DENSE, OracleBackend()))


def _solve(a, b, c, z, p, **kw):
  kw.setdefault("verbose", False)
  return qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z,
                       p=None if p is None else sp.csc_matrix(p)).solve(**kw)


@pytest.mark.parametrize("key", sorted(TRAJ))
def test_driver_matches_reference_trajectory(oracle_backend, key):
  case = TRAJ[key]
  a, b, c, z, p = golden_cases.inputs(case["spec"])
  sol = _solve(a, b, c, z, p, equilibrate=case["equilibrate"], collect_stats=True)
  assert sol.status.value == case["status"]
  assert len(sol.stats) == case["iters"]
  np.testing.assert_allclose([s["mu"] for s in sol.stats][:3],
                             [t["mu"] for t in case["trajectory"]][:3], rtol=1e-8)
  ref = {k: np.array([np.nan if v is None else v for v in case[k]]) for k in "xys"}
  for k in "xys":
    got = getattr(sol, k)
    assert np.array_equal(np.isnan(got), np.isnan(ref[k]))
    ok = ~np.isnan(ref[k])
    if ok.any():
      assert np.max(np.abs(got[ok] - ref[k][ok])) / max(1, np.max(np.abs(ref[k][ok]))) < 1e-4


def test_stats_keys_and_iter_sequence(oracle_backend):
  a, b, c, p = G.gen_feasible(10, 5, 3, 42)
  sol = _solve(a, b, c, 3, p, collect_stats=True)
  base = {"iter", "pcost", "dcost", "pres", "dres", "gap", "mu", "sigma", "alpha", "tau",
          "norm_x", "norm_y", "status", "time", "prelrhs", "drelrhs", "pinfeas", "dinfeas"}
  extra = {"complementarity", "norm_s", "max_sy", "min_sy", "std_sy", "max_s_over_y",
           "min_s_over_y", "mean_s_over_y", "std_s_over_y"}
  for st in sol.stats:
    assert base <= st.keys() and extra <= st.keys()
  assert [s["iter"] for s in sol.stats] == list(range(len(sol.stats)))
  assert sol.stats[-1]["status"] == qtqp_amd.SolutionStatus.SOLVED
  assert _solve(a, b, c, 3, p, collect_stats=False).stats == []


def test_hit_max_iter(oracle_backend):
  a, b, c, p = G.gen_feasible(30, 20, 5, 1)
  sol = _solve(a, b, c, 5, p, max_iter=1, collect_stats=True)
  assert sol.status == qtqp_amd.SolutionStatus.HIT_MAX_ITER
  assert len(sol.stats) == 1 and sol.stats[-1]["status"] == sol.status
  assert np.all(np.isfinite(sol.x)) and np.all(np.isfinite(sol.y))


def test_verbose_output(oracle_backend, capsys):
  a, b, c, p = G.gen_feasible(10, 5, 3, 42)
  _solve(a, b, c, 3, p, verbose=True)
  out = capsys.readouterr().out
  assert "QTQP" in out and "iter" in out and "Solved" in out
  a, b, c, p = G.gen_infeasible(30, 20, 4, 7)
  _solve(a, b, c, 4, p, verbose=True)
  assert "Primal infeasible" in capsys.readouterr().out
  _solve(a, b, c, 4, p, verbose=False)
  assert capsys.readouterr().out == ""


def test_presolve_restores_certificate(oracle_backend):
  # tests/test_qtqp.py:561-596 (dropped rows with b >= 1e20)
  rng = np.random.default_rng(1842)
  a, b, c, p = G.gen_infeasible(30, 20, 4, 1842)
  a_full = np.vstack([a, rng.normal(size=(3, 20))])
  b_full = np.concatenate([b, np.full(3, 1e21)])
  sol = _solve(a_full, b_full, c, 4, p)
  G.check_infeasible(sol.x, sol.y, sol.s, a_full, b_full, 4)
  a, b, c, p = G.gen_unbounded(30, 20, 4, 1942)
  a_full = np.vstack([a, rng.normal(size=(3, 20))])
  b_full = np.concatenate([b, np.full(3, np.inf)])
  sol = _solve(a_full, b_full, c, 4, p)
  assert sol.status == qtqp_amd.SolutionStatus.UNBOUNDED
  assert np.all(np.isnan(sol.y)) and np.all(np.isnan(sol.s[-3:]))


def test_presolve_drop_solved_fills_slack(oracle_backend):
  a, b, c, p = G.gen_feasible(30, 20, 4, 5)
  rng = np.random.default_rng(0)
  extra = rng.normal(size=(2, 20))
  sol = _solve(np.vstack([a, extra]), np.concatenate([b, [1e20, np.inf]]), c, 4, p)
  assert sol.status == qtqp_amd.SolutionStatus.SOLVED
  np.testing.assert_allclose(sol.y[-2:], 0.0)
  np.testing.assert_allclose(sol.s[-2:][0], 1e20 - extra[0] @ sol.x)


def test_input_validation():
  a = sp.csc_matrix(np.ones((3, 2)))
  with pytest.raises(TypeError):
    qtqp_amd.QTQP(a=a.tocsr(), b=np.ones(3), c=np.ones(2), z=0)
  with pytest.raises(ValueError):
    qtqp_amd.QTQP(a=a, b=np.ones(4), c=np.ones(2), z=0)
  with pytest.raises(ValueError):
    qtqp_amd.QTQP(a=a, b=np.ones(3), c=np.ones(3), z=0)
  with pytest.raises(ValueError):
    qtqp_amd.QTQP(a=a, b=np.ones(3), c=np.ones(2), z=-1)
  with pytest.raises(ValueError, match="effective z == m"):
    qtqp_amd.QTQP(a=a, b=np.ones(3), c=np.ones(2), z=3)
  with pytest.raises(ValueError, match="Equality RHS entries"):
    qtqp_amd.QTQP(a=a, b=np.array([1.0, np.inf, 1.0]), c=np.ones(2), z=2)
  with pytest.raises(ValueError):
    qtqp_amd.QTQP(a=a, b=np.array([1.0, np.nan, 1.0]), c=np.ones(2), z=1)
  with pytest.raises(ValueError, match="symmetric"):
    qtqp_amd.QTQP(a=a, b=np.ones(3), c=np.ones(2), z=0, p=sp.csc_matrix([[1.0, 1.0], [0.0, 1.0]]))
  with pytest.raises(TypeError):
    qtqp_amd.QTQP(a=a, b=np.ones(3), c=np.ones(2), z=0, p=sp.csr_matrix(np.eye(2)))
  with pytest.raises(ValueError):
    qtqp_amd.QTQP(a=a, b=np.ones(3), c=np.ones(2), z=0, p=sp.csc_matrix(np.ones((3, 2))))


def test_backend_freed_on_exception(monkeypatch):
  # tests/test_qtqp.py:700-737
  class Failing(direct.LinearSolver):
    freed = False

    def factorize(self):
      raise RuntimeError("forced factorization failure")

    def format(self):
      return "csc"

    def free(self):
      self.freed = True

  be = Failing()
  monkeypatch.setattr(qtqp_amd, "_resolve_linear_solver", lambda ls, *shape: (ls, be))
  a, b, c, p = G.gen_feasible(20, 10, 3, 842)
  solver = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=3, p=sp.csc_matrix(p))
  with pytest.raises(RuntimeError, match="forced factorization failure"):
    solver.solve(verbose=False, linear_solver=qtqp_amd.LinearSolver.HIP_DENSE)
  assert be.freed and solver._linear_solver is None  # pylint: disable=protected-access


def test_upper_triangular_matvec_matches_full():
  # tests/test_qtqp.py:774-819 with the protocol's default matvec.
  class Tri(direct.LinearSolver):
    def factorize(self):
      pass

    def format(self):
      return "csc"

  a, _, _, p = G.gen_feasible(20, 12, 4, 2026, density=0.3)
  rng = np.random.default_rng(1)
  mu = rng.uniform()
  s, y = rng.uniform(size=20), rng.uniform(size=20)
  s[:4] = 0
  kkt = direct.DirectKktSolver(a=sp.csc_matrix(a), p=sp.csc_matrix(p), z=4,
                               min_static_regularization=1e-8,
                               max_iterative_refinement_steps=2, atol=1e-12, rtol=1e-12,
                               solver=Tri())
  kkt.update(mu=mu, s=s, y=y)
  hy = np.concatenate([np.full(4, mu), s[4:] / y[4:] + mu])
  full = np.block([[p + mu * np.eye(12), a.T], [a, -np.diag(hy)]])
  v = rng.normal(size=32)
  np.testing.assert_allclose(kkt._solver @ v, full @ v, rtol=1e-10, atol=1e-10)  # pylint: disable=protected-access


def test_tau_gating_uses_linearized_when_not_converged():
  # tests/test_qtqp.py:1155-1220
  solver = qtqp_amd.QTQP(a=sp.csc_matrix(np.eye(2)), b=np.ones(2), c=np.ones(2), z=0,
                         p=sp.csc_matrix((2, 2)))
  solver.q = np.ones(4)
  solver.kinv_q = np.zeros(4)
  for converged, want in ((True, "quadratic"), (False, "linearized")):
    class Fake:
      def solve(self, rhs, warm_start):
        return np.zeros(4), {"solves": 1, "final_residual_norm": 1e-4, "rhs_norm": 1.0,
                             "tolerance": 1e-3, "converged": converged,
                             "status": "converged" if converged else "stalled"}
    solver._linear_solver = Fake()  # pylint: disable=protected-access
    solver._solve_for_tau = lambda *a, **k: 1.0  # pylint: disable=protected-access
    solver._solve_for_tau_linearized_fallback = lambda *a, **k: 2.0  # pylint: disable=protected-access
    _, _, tau, st = solver._newton_step(  # pylint: disable=protected-access
        p=sp.csc_matrix((2, 2)), mu=1.0, mu_target=0.0, r_anchor=np.zeros(4), tau_anchor=1.0,
        x=np.zeros(2), y=np.ones(2), s=np.ones(2), tau=1.0, correction=None)
    assert st["tau_method"] == want and tau == (1.0 if converged else 2.0)


def test_driver_scalar_pieces_match_oracle():
  rng = np.random.default_rng(3)
  m, n, z = 30, 20, 4
  a, b, c, p = G.gen_feasible(m, n, z, 3)
  solver = qtqp_amd.QTQP(a=sp.csc_matrix(a), b=b, c=c, z=z, p=sp.csc_matrix(p))
  mu, alpha = 0.3, 0.7
  x, y, tau, s = rng.normal(size=n), rng.uniform(size=m), 1.3, rng.uniform(size=m)
  dx, dy, dtau, ds = rng.normal(size=n), rng.normal(size=m), 0.2, rng.normal(size=m)
  assert solver._compute_sigma(mu, x, y, tau, s, alpha, dx, dy, dtau, ds) == pytest.approx(  # pylint: disable=protected-access
      O.sigma_mehrotra(mu, x, y, tau, s, alpha, dx, dy, dtau, ds, m, z), rel=1e-14)
  assert solver._compute_step_size(y, s, dy, ds) == O.step_size(y, s, dy, ds, z)  # pylint: disable=protected-access
  xn, yn, tn, sn = solver._normalize(x.copy(), y.copy(), tau, s.copy())  # pylint: disable=protected-access
  np.testing.assert_allclose(xn @ xn + yn @ yn + tn * tn, m - z + 1, rtol=1e-12)


def test_auto_reports_unavailable_backends(monkeypatch):
  monkeypatch.setattr(qtqp_amd, "_AUTO_SOLVER_CACHE", {})

  def boom(ls):
    raise OSError(f"{ls.name} unavailable")
  monkeypatch.setattr(qtqp_amd, "_instantiate_linear_solver", boom)
  with pytest.raises(RuntimeError, match="AUTO could not initialize"):
    qtqp_amd._resolve_linear_solver(qtqp_amd.LinearSolver.AUTO)  # pylint: disable=protected-access


@pytest.mark.parametrize("m, n, nnz, first", [
    (300, 200, 300 * 200 + 200 * 200, "HIP_DENSE"),         # small dense QP (cfg1)
    (2048, 256, 2048 * 256 // 2, "HIP_DENSE"),              # condensed MPC (cfg3)
    (6000, 4000, 6000 * 4000 + 4000 * 4000, "HIP_DENSE"),   # large dense QP (cfg2)
    (6000, 4000, 60_000, "HIP_SPARSE"),                     # reference's sparse family
    (60_000, 50_000, 200_000, "HIP_SPARSE"),                # past the dense n limit
    (40_000, 20_000, 40_000 * 20_000, "HIP_SPARSE"),        # dense, but over 4 GiB
    (1000, 500, 2000, "HIP_DENSE"),                         # sparse but small
])
def test_auto_chooses_by_problem(monkeypatch, m, n, nnz, first):
  """AUTO's choice depends on each problem, and a backend that failed to
  initialise is skipped without fixing the choice for later problems."""
  monkeypatch.setattr(qtqp_amd, "_AUTO_SOLVER_CACHE", {})
  made = []

  def make(ls):
    made.append(ls.name)
    return object()
  monkeypatch.setattr(qtqp_amd, "_instantiate_linear_solver", make)
  order = qtqp_amd._auto_linear_solver_order(m, n, nnz)  # pylint: disable=protected-access
  assert order[0].name == first and len(set(order)) == 3
  ls, _ = qtqp_amd._resolve_linear_solver(qtqp_amd.LinearSolver.AUTO, (m, n, nnz))  # pylint: disable=protected-access
  assert ls.name == first and made == [first]


def test_auto_skips_unavailable_backends_per_problem(monkeypatch):
  monkeypatch.setattr(qtqp_amd, "_AUTO_SOLVER_CACHE", {})

  def make(ls):
    if ls is qtqp_amd.LinearSolver.HIP_SPARSE:
      raise OSError("no sparse library")
    return object()
  monkeypatch.setattr(qtqp_amd, "_instantiate_linear_solver", make)
  sparse_shape = (6000, 4000, 60_000)
  ls, _ = qtqp_amd._resolve_linear_solver(qtqp_amd.LinearSolver.AUTO, sparse_shape)  # pylint: disable=protected-access
  assert ls is qtqp_amd.LinearSolver.HIP_DENSE
  ls, _ = qtqp_amd._resolve_linear_solver(qtqp_amd.LinearSolver.AUTO, (300, 200, 100_000))  # pylint: disable=protected-access
  assert ls is qtqp_amd.LinearSolver.HIP_DENSE


def test_family_batch_presolve_rows_must_agree():
  """A family batch drops the rows with b >= 1e20 once for all its QPs: the
  rows must be the same in every QP, and the reference's input checks apply
  (__init__.py:248-270)."""
  from qtqp_amd.batch import BatchQP
  b = np.array([[1.0, 2.0, np.inf, 3.0], [0.5, 1.0, 1e21, 2.0]])
  keep = BatchQP._family_keep(b, 1)
  np.testing.assert_array_equal(keep, [True, True, False, True])
  with pytest.raises(ValueError, match="same rows"):
    BatchQP._family_keep(np.array([[1.0, 2.0, np.inf], [1.0, 2.0, 3.0]]), 1)
  with pytest.raises(ValueError, match="Equality RHS"):
    BatchQP._family_keep(np.array([[np.inf, 2.0], [1.0, 2.0]]), 1)
  with pytest.raises(ValueError, match="Inequality RHS"):
    BatchQP._family_keep(np.array([[1.0, -np.inf], [1.0, 2.0]]), 1)


def test_family_bc_matches_condensed_mpc():
  """Family.bc ([x0, 1] times the stacked affine maps) gives each x0's b and c
  as problems.condensed_mpc builds them for that x0."""
  from qtqp_amd import problems, warmstart
  fam = warmstart.Family(horizon=10, nx=4, nu=2, family_seed=3)
  x0s = problems.parametric_mpc_x0(5, seed=11, nx=4)
  b, c = fam.bc(x0s)
  assert b.shape == (5, fam.m) and c.shape == (5, fam.n)
  for i, x0 in enumerate(x0s):
    q = problems.condensed_mpc(fam.sys, fam.horizon, x0)
    np.testing.assert_allclose(b[i], q.b, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(c[i], q.c, rtol=1e-12, atol=1e-12)
