"""Learned warm start on the parametric MPC family: what the trained network
gains over the default start, with identical statuses and KKT-accurate
solutions (checked with the reference's KKT assertion restated in
tests/qp_generators.py)."""
import numpy as np
import pytest

from qtqp_amd import problems, warmstart
import qp_generators as G

pytestmark = pytest.mark.gpu


def test_warm_start_cuts_iterations_with_identical_statuses(require_gpu):
  fam = warmstart.Family()
  x_tr = problems.parametric_mpc_x0(8192, seed=11)
  x_ev = problems.parametric_mpc_x0(1024, seed=22)
  cold = warmstart._solve(fam, x_ev)
  ws = warmstart.WarmStart(fam, k_star=7, seed=0)
  t, _ = ws.targets(x_tr)
  losses = ws.fit(x_tr, t, steps=2000)
  assert losses[-50:].mean() < 0.05 * losses[:5].mean()
  warm = warmstart.solve_with_warm_start(ws, x_ev, cap=int(np.percentile(cold["iters"], 99)))
  np.testing.assert_array_equal(warm["status"], cold["status"])
  cut = 1.0 - warm["iters"].mean() / cold["iters"].mean()
  assert cut >= 0.30, cut
  assert warm["fallback"].mean() < 0.05
  qps = fam.qps(x_ev[:32])
  for i, (a, b, c, z, p) in enumerate(qps):
    G.check_solution(warm["x"][i], warm["y"][i], warm["s"][i], a, b, c, p, z)
  ws.mlp.free()


def test_checkpoint_round_trip(require_gpu, tmp_path):
  """A saved warm start reloads to the same proposals."""
  fam = warmstart.Family()
  ws = warmstart.WarmStart(fam, hidden=64, k_star=3)
  x0 = problems.parametric_mpc_x0(512, seed=5)
  t, _ = ws.targets(x0)
  ws.fit(x0, t, steps=50)
  path = tmp_path / "ws.npz"
  ws.save(path)
  ws2 = warmstart.WarmStart.load(path)
  a, b = ws.propose(x0[:64]), ws2.propose(x0[:64])
  for u, v in zip(a, b):
    np.testing.assert_array_equal(u, v)
  ws.mlp.free()
  ws2.mlp.free()


def test_device_proposal_matches_the_host_decoding(require_gpu):
  """The proposal decoded on the device (qtqp_mlp_propose into the batch)
  solves like the same proposal decoded on the host by ``propose`` and set
  with ``set_start``: same statuses and iteration counts, solutions within
  1e-10 (the two decodings differ at most in the last bit of exp)."""
  fam = warmstart.Family()
  ws = warmstart.WarmStart(fam, hidden=64, k_star=4)
  x0 = problems.parametric_mpc_x0(2048, seed=7)
  t, _ = ws.targets(x0)
  ws.fit(x0, t, steps=300)
  x_ev = problems.parametric_mpc_x0(256, seed=8)
  # a non-finite input makes one proposal invalid: it gets the default start
  x_ev[3, 0] = np.nan
  solver = warmstart.FamilySolver(fam)
  dev = solver.solve(np.nan_to_num(x_ev), proposer=None)  # warm up the handle
  px, py, ps, pt, _, valid = ws.propose(x_ev)
  dx, dy, ds, dt = warmstart.default_start(fam, x_ev.shape[0])
  host = solver.solve(np.nan_to_num(x_ev),
                      start=(np.where(valid[:, None], px, dx), np.where(valid[:, None], py, dy),
                             np.where(valid[:, None], ps, ds), np.where(valid, pt, dt)))
  bq = solver._batch(np.nan_to_num(x_ev))
  got_valid = ws.mlp.propose((x_ev - ws.xm) / ws.xs, ws.tm, ws.ts, fam.n, fam.m, fam.z,
                             bq.device_buffers())
  np.testing.assert_array_equal(got_valid, valid)
  assert not valid[3] and valid.sum() == valid.size - 1
  bq.set_start_device(*bq.device_buffers())
  bq.solve_device()
  res = bq.results_arrays()
  np.testing.assert_array_equal(res["status"], host["status"])
  np.testing.assert_array_equal(res["iters"], host["iters"])
  for k in "xys":
    err = np.nanmax(np.abs(res[k] - host["sol_" + k]))
    assert err < 1e-10, (k, err)
  assert dev["status"].shape == host["status"].shape
  solver.free()
  ws.mlp.free()
