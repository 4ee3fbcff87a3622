"""GPU parity of the batched device IPM against the CPU oracle.

Bar: same status as the FP64 oracle on identical inputs, x/y/s within 1e-6
relative (problems with a unique solution; 1e-4 for the degenerate sparse
reference families), IPM iteration counts within one of the oracle's.
"""
import numpy as np
import pytest

import qtqp_amd
from qtqp_amd import batch, problems
from oracle import qtqp_oracle as O
import golden_cases

pytestmark = pytest.mark.gpu
TRAJ = golden_cases.load_trajectories()


def _rel(u, v):
  ok = ~np.isnan(v)
  if not ok.any():
    return 0.0
  return float(np.max(np.abs(u[ok] - v[ok])) / max(1.0, np.max(np.abs(v[ok]))))


def _compare(sol, ref, tol):
  assert sol.status.value == ref["status"], (sol.status, ref["status"])
  for k in "xys":
    got, want = getattr(sol, k), ref[k]
    assert np.array_equal(np.isnan(got), np.isnan(want)), k
    assert _rel(got, want) < tol, (k, _rel(got, want))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("equilibrate", [True, False])
def test_batch_matches_oracle_on_fixture_problems(require_gpu, precision, equilibrate):
  keys = sorted(k for k in TRAJ if TRAJ[k]["equilibrate"] == equilibrate)
  inputs = [golden_cases.inputs(TRAJ[k]["spec"]) for k in keys]
  bq = batch.BatchQP(inputs)
  res = bq.solve(equilibrate=equilibrate, precision=precision)
  bq.free()
  for k, inp, sol in zip(keys, inputs, res.solutions):
    ref = O.solve_qp(*inp, equilibrate=equilibrate)
    unique = TRAJ[k]["spec"]["kind"] in ("readme", "dense", "mpc") or ref["status"] != "solved"
    _compare(sol, ref, 1e-6 if unique else 1e-4)
    assert abs(sol.iters - ref["iters"]) <= 1, (k, sol.iters, ref["iters"])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_uneven_mpc_batch_matches_oracle(require_gpu, precision):
  qps = problems.uneven_mpc_batch(48, base_seed=500, t_range=(4, 24), nx_range=(2, 10),
                                  nu_range=(1, 4))
  bq = batch.BatchQP(qps)
  res = bq.solve(precision=precision)
  for qp, sol in zip(qps, res.solutions):
    ref = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
    _compare(sol, ref, 1e-6)
    assert abs(sol.iters - ref["iters"]) <= 1
  # re-solving the same resident batch is bit-for-bit deterministic
  res2 = bq.solve(precision=precision, use_graph=False)
  for s1, s2 in zip(res.solutions, res2.solutions):
    np.testing.assert_array_equal(s1.x, s2.x)
    assert s1.iters == s2.iters
  bq.free()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_trace_reports_the_refinement_results(require_gpu, precision):
  """The per-iteration trace carries each KKT solve's own result: solve counts
  within max_iterative_refinement_steps, non-negative residual norms, and a
  residual below the refinement tolerance wherever a solve converged.  (A
  miscompiled copy of the residual once returned an element of the solution
  vector instead, for solves with an odd number of steps.)"""
  keys = sorted(TRAJ)
  inputs = [golden_cases.inputs(TRAJ[k]["spec"]) for k in keys]
  bq = batch.BatchQP(inputs)
  bq.solve_device(precision=precision)
  raw = bq.raw_iterates()
  for q in range(len(keys)):
    for row in bq.trace(q, int(raw["iters"][q])):
      for kind in ("q", "p", "c"):
        solves, res = row[f"{kind}_solves"], row[f"{kind}_res"]
        stalled = solves >= 1000
        steps = solves - 1000 if stalled else solves
        assert 1 <= steps <= 20 and res >= 0.0, (keys[q], kind, row)
        if not stalled and steps < 20:
          assert res < 1e-6, (keys[q], kind, row)
  bq.free()


def test_large_instances_in_batch(require_gpu):
  # multi-panel factor (n > 64), long rows (m > 1000) and uneven neighbours
  qps = [problems.mpc_qp(40, 12, 4, seed=3), problems.random_dense_qp(150, 220, 30, seed=4),
         problems.mpc_qp(6, 3, 1, seed=5)]
  bq = batch.BatchQP(qps)
  res = bq.solve(precision="mixed")
  for qp, sol in zip(qps, res.solutions):
    _compare(sol, O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p), 1e-6)
  bq.free()


def test_warm_start_from_solution_cuts_iterations(require_gpu):
  qps = problems.uneven_mpc_batch(8, base_seed=900, t_range=(8, 16), nx_range=(3, 6),
                                  nu_range=(1, 3))
  bq = batch.BatchQP(qps)
  cold = bq.solve()
  # a central, strictly interior point near the optimum: (x*, y* + 1e-3, s* + 1e-3)
  xs = [s.x for s in cold.solutions]
  ys = [s.y + 1e-3 for s in cold.solutions]
  ss = [s.s + 1e-3 for s in cold.solutions]
  bq.set_start(xs, ys, ss, np.ones(len(qps)))
  warm = bq.solve()
  bq.free()
  for c, w in zip(cold.solutions, warm.solutions):
    assert w.status == c.status == qtqp_amd.SolutionStatus.SOLVED
    assert _rel(w.x, c.x) < 1e-6
  assert warm.iterations.mean() <= cold.iterations.mean()


def test_batch_input_validation(require_gpu):
  with pytest.raises(ValueError, match="effective z == m"):
    batch.BatchQP([(np.ones((2, 2)), np.ones(2), np.ones(2), 2, None)])
  with pytest.raises(ValueError):
    batch.BatchQP([(np.ones((2, 2)), np.ones(3), np.ones(2), 0, None)])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_tile_shapes_give_identical_results_and_autotune(require_gpu, precision):
  """Gram tile shapes change no result (each G entry sums the same products in
  the same order): every FP32 shape (with the FP64 shape fixed) and every
  FP64 shape (with the FP32 shape fixed) gives the same iterates bit for bit,
  and the autotuner fills one choice per occupied bucket and precision."""
  qps = problems.uneven_mpc_batch(40, base_seed=3000)
  bq = batch.BatchQP(qps)
  nb = len(batch.BatchQP.BUCKET_EDGES)
  fp64_shapes = [0, 2, 3]  # 128x128 (1) is FP32 only
  outs = []
  for shape in range(len(batch.BatchQP.TILE_SHAPES)):
    bq.set_tiles([shape] * nb, [0] * nb)
    bq.solve_device(precision=precision)
    outs.append(bq.raw_iterates())
  for shape in fp64_shapes:
    bq.set_tiles([0] * nb, [shape] * nb)
    bq.solve_device(precision=precision)
    outs.append(bq.raw_iterates())
  for o in outs[1:]:
    for k in ("x", "y", "s", "tau", "status", "iters"):
      assert np.array_equal(outs[0][k], o[k]), k
  tuned = bq.autotune(sample=6, iters=2)
  occupied = {int(np.searchsorted(batch.BatchQP.BUCKET_EDGES, q.n)) for q in qps}
  for prec, allowed in (("fp32", (0, 1, 2, 3)), ("fp64", tuple(fp64_shapes))):
    for b, choice in enumerate(tuned[prec]["table"]):
      ms = tuned[prec]["ms"][b]
      if b in occupied:
        assert choice in allowed
        assert all(ms[v] is not None and ms[v] > 0 for v in allowed)
      else:
        assert choice == -1 and all(v is None for v in ms)
    if prec == "fp64":
      assert all(row[1] is None for row in tuned[prec]["ms"])
  bq.solve_device(precision=precision)
  again = bq.raw_iterates()
  for k in ("x", "y", "s", "status", "iters"):
    assert np.array_equal(outs[0][k], again[k]), k
  with pytest.raises(ValueError):
    bq.set_tiles([5] * nb)
  with pytest.raises(ValueError):
    bq.set_tiles([0] * nb, [1] * nb)
  bq.free()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_row_pass_chunks_and_slot_boundaries(require_gpu, precision):
  """Row counts past the 1984-row staging chunk (one, two and three chunks, a
  partial last stage) and leading dimensions on both sides of the 128/256/512
  column-slot boundaries of the pipelined row pass, against the oracle."""
  qps = [problems.random_dense_qp(128, 2017, 10, seed=21),   # ld 128, 2 chunks
         problems.random_dense_qp(129, 2100, 20, seed=22),   # ld 132 (2 slots)
         problems.random_dense_qp(256, 4100, 50, seed=23),   # ld 256, 3 chunks
         problems.random_dense_qp(257, 300, 40, seed=24),    # ld 260 (4 slots)
         problems.random_dense_qp(513, 700, 30, seed=25),    # ld 516 (8 slots)
         problems.random_dense_qp(5, 33, 3, seed=26)]        # ld 8, 33 rows
  bq = batch.BatchQP(qps)
  res = bq.solve(precision=precision)
  bq.free()
  for qp, sol in zip(qps, res.solutions):
    ref = O.solve_qp(qp.a, qp.b, qp.c, qp.z, qp.p)
    _compare(sol, ref, 1e-6)
    assert abs(sol.iters - ref["iters"]) <= 1


def test_maximum_size_instance_in_batch(require_gpu):
  """n near the engine's limit (1024) beside small instances: the widest row
  passes (8 slots of 128 columns), 16 diagonal blocks, every Gram tile shape."""
  qps = [problems.random_dense_qp(1000, 1100, 100, seed=6), problems.mpc_qp(6, 3, 1, seed=7),
         problems.random_dense_qp(70, 90, 10, seed=8)]
  refs = [O.solve_qp(q.a, q.b, q.c, q.z, q.p) for q in qps]
  bq = batch.BatchQP(qps)
  nb = len(batch.BatchQP.BUCKET_EDGES)
  for s32, s64 in ((0, 0), (1, 2), (2, 3), (3, 2)):
    bq.set_tiles([s32] * nb, [s64] * nb)
    res = bq.solve(precision="mixed")
    for sol, ref in zip(res.solutions, refs):
      _compare(sol, ref, 1e-6)
  bq.free()


def test_full_size_uneven_mpc_batch(require_gpu):
  """The 4096-QP uneven MPC workload at full size: every instance ends solved or
  infeasible, solved ones meet the stopping tests on the unscaled data
  (size-independent KKT property), and a spread sample matches the oracle."""
  qps = problems.uneven_mpc_batch(4096, base_seed=1000)
  bq = batch.BatchQP(qps)
  bq.autotune(sample=16, iters=2)
  res = bq.solve(precision="mixed")
  assert {s.status for s in res.solutions} <= {qtqp_amd.SolutionStatus.SOLVED,
                                               qtqp_amd.SolutionStatus.INFEASIBLE}
  for q, s in zip(qps, res.solutions):
    if s.status != qtqp_amd.SolutionStatus.SOLVED:
      continue
    ax, aty, px = q.a @ s.x, q.a.T @ s.y, q.p @ s.x
    pres = np.max(np.abs(ax + s.s - q.b))
    dres = np.max(np.abs(px + aty + q.c))
    prel = max(np.max(np.abs(ax)), np.max(np.abs(s.s)), np.max(np.abs(q.b)))
    drel = max(np.max(np.abs(px)), np.max(np.abs(aty)), np.max(np.abs(q.c)))
    # the engine stops at atol 1e-7 + rtol 1e-8 * rel on its own (tau-scaled)
    # residuals; the recomputed ones may differ by rounding, hence 1e-6 here
    assert pres <= 1e-6 * max(1.0, prel) and dres <= 1e-6 * max(1.0, drel)
  for i in range(0, 4096, 512):
    q = qps[i]
    _compare(res.solutions[i], O.solve_qp(q.a, q.b, q.c, q.z, q.p), 1e-6)
  bq.free()


@pytest.mark.parametrize("use_graph", [True, False])
def test_persistent_and_per_iteration_schedules_agree_bitwise(require_gpu, use_graph):
  """One persistent launch per solve (each workgroup iterates a QP to
  completion) and one launch per IPM iteration run the same per-QP code on the
  same data, so every result is bitwise identical; completion times are finite
  and within the solve."""
  qps = problems.uneven_mpc_batch(300, base_seed=5000)
  bq = batch.BatchQP(qps)
  outs, cnts, inst = [], [], []
  for per_iteration in (False, True):
    cnts.append(bq.solve_device(precision="mixed", per_iteration=per_iteration,
                                use_graph=use_graph))
    outs.append(bq.raw_iterates())
    inst.append(bq.instance_ms())
  for k in ("x", "y", "s", "tau", "status", "iters"):
    assert np.array_equal(outs[0][k], outs[1][k]), k
  assert cnts[0]["iterations"] == cnts[1]["iterations"] == int(outs[0]["iters"].max())
  # the host checks for completion every second launch
  assert cnts[1]["iterations"] <= len(bq.launch_times_ms()) <= cnts[1]["iterations"] + 1
  for c, t in zip(cnts, inst):
    assert np.all(np.isfinite(t)) and np.all(t >= 0) and np.max(t) <= c["device_ms"]
  bq.free()


def _paired_qp(seed):
  """A feasible QP whose A has row pairs in four segments: equality rows
  E (the first two negated again at the end, as inequalities), inequality
  rows G and -G, then unpaired rows H."""
  rng = np.random.default_rng(seed)
  n, z, k, j = 24, 5, 30, 17
  e, g, hh = (rng.standard_normal((r, n)) for r in (z, k, j))
  a = np.concatenate([e, g, -g, hh, -e[:2]])
  x0 = rng.standard_normal(n)
  slack = rng.uniform(0.1, 1.0, a.shape[0] - z)
  b = a @ x0
  b[z:] += slack
  r = rng.standard_normal((n, n))
  p = r.T @ r / n + 0.1 * np.eye(n)
  return problems.DenseQP(a, b, rng.standard_normal(n), z, p)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_row_pairs_stored_once(require_gpu, precision):
  """Rows that negate another row are stored once: MPC's [I; -I; G; -G] keeps
  half its rows, a QP with pairs in four segments keeps m less its pairs, and
  one whose pairs need more segments (adjacent pairs) keeps every row.  Each
  matches the oracle, and the paired and unpaired layouts agree closely."""
  mpc = problems.uneven_mpc_batch(12, base_seed=700, t_range=(4, 20))
  seg = _paired_qp(3)
  rng = np.random.default_rng(4)
  g = rng.standard_normal((12, 10))
  adj = np.empty((24, 10))
  adj[0::2], adj[1::2] = g, -g  # pairs (2i, 2i+1): one segment per pair
  adj_qp = problems.DenseQP(adj, np.abs(rng.standard_normal(24)) + 0.5,
                            rng.standard_normal(10), 0, np.eye(10))
  qps = mpc + [seg, adj_qp]
  paired = batch.BatchQP(qps)
  plain = batch.BatchQP(qps, pair_rows=False)
  mc = paired.stored_rows()
  np.testing.assert_array_equal(mc[:12], [q.m // 2 for q in mpc])
  assert mc[12] == seg.a.shape[0] - 30 - 2
  assert mc[13] == 24
  np.testing.assert_array_equal(plain.stored_rows(), [q.a.shape[0] for q in qps])
  rp = paired.solve(precision=precision)
  ru = plain.solve(precision=precision)
  paired.free()
  plain.free()
  for q, sp, su in zip(qps, rp.solutions, ru.solutions):
    ref = O.solve_qp(q.a, q.b, q.c, q.z, q.p)
    _compare(sp, ref, 1e-6)
    assert abs(sp.iters - ref["iters"]) <= 1
    assert sp.status == su.status
    for k in "xys":
      assert _rel(getattr(sp, k), getattr(su, k)) < 1e-7, k


def _banded_qp(n, seed):
  """A feasible QP whose rows have zero columns before and after their
  nonzeros: bound rows [I; -I], block lower-triangular rows (as in
  condensed MPC), rows zero before a start column, and one zero row."""
  rng = np.random.default_rng(seed)
  k = max(2, n // 16)
  tri = np.tril(rng.standard_normal((n, n)))[:: max(1, n // 40)]
  tail = rng.standard_normal((k, n))
  for i in range(k):
    tail[i, : (i * n) // k] = 0.0
  a = np.concatenate([np.eye(n), -np.eye(n), tri, -tri, tail, np.zeros((1, n))])
  x0 = rng.uniform(-0.5, 0.5, n)
  b = a @ x0 + rng.uniform(0.1, 1.0, a.shape[0])
  r = rng.standard_normal((n, n))
  p = r.T @ r / n + 0.1 * np.eye(n)
  return problems.DenseQP(a, b, rng.standard_normal(n), 0, p)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_row_extents_give_the_full_row_results(require_gpu, precision):
  """Reading only each row's nonzero column extent leaves out exact zeros, so
  the iterates equal those of full-row reads in every element, for each
  row-pass shape (n up to 128, 256, 512 and 1024) and for rows with zeros on
  either side; the results also match the oracle."""
  qps = [_banded_qp(n, 40 + n) for n in (37, 128, 200, 300, 600)]
  qps += problems.uneven_mpc_batch(6, base_seed=900, t_range=(4, 30))
  outs = []
  for ext in (True, False):
    bq = batch.BatchQP(qps, row_extents=ext)
    bq.solve_device(precision=precision)
    outs.append(bq.raw_iterates())
    sols = bq.results()
    bq.free()
  for k in ("x", "y", "s", "tau", "status", "iters"):
    assert np.array_equal(outs[0][k], outs[1][k]), k
  for q, sol in zip(qps, sols):
    ref = O.solve_qp(q.a, q.b, q.c, q.z, q.p)
    _compare(sol, ref, 1e-6)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fused_solves_equal_the_solves_in_turn_bitwise(require_gpu, precision):
  """The two refined solves of an iteration that run together (K^{-1}[c; b]
  and the predictor's, once a QP factors in FP64) keep each solve's own
  arithmetic, so every iterate equals that of the solves run in turn, bit for
  bit: for both row-pass shapes of the two-solve pass (ld up to 128 and
  256), LPs and QPs, paired rows, and n above 256 where the solves run in
  turn either way.  Fewer passes over A show that the fused path ran."""
  qps = [_banded_qp(n, 70 + n) for n in (37, 128, 200, 256, 300, 600)]
  qps += problems.uneven_mpc_batch(6, base_seed=910, t_range=(4, 40))
  lp = problems.random_dense_qp(90, 160, 10, seed=31)
  qps.append((lp.a, lp.b, lp.c, lp.z, None))
  outs, work = [], []
  for fused in (True, False):
    bq = batch.BatchQP(qps, fused_solves=fused)
    bq.solve_device(precision=precision)
    outs.append(bq.raw_iterates())
    work.append(bq.work_counts())
    bq.free()
  for k in ("x", "y", "s", "tau", "status", "iters"):
    assert np.array_equal(outs[0][k], outs[1][k]), k
  ns = np.array([np.shape(q.a if hasattr(q, "a") else q[0])[1] for q in qps])
  # K products: fewer where the fused path runs (n <= 256), equal above
  small = ns <= 256
  assert np.all(work[0][small, 2] <= work[1][small, 2])
  assert work[0][small, 2].sum() < work[1][small, 2].sum()
  if precision == "fp64":
    assert np.all(work[0][small, 2] < work[1][small, 2])
  assert np.array_equal(work[0][~small, 2], work[1][~small, 2])
  # factorisations and factor solves are the same work either way
  for col in (0, 1, 3, 4):
    assert np.array_equal(work[0][:, col], work[1][:, col]), col
