"""GPU tests of the family batch: QPs that share A and P, held once on the device.

A family batch must give the results of an ordinary batch over copies of the
same QPs bit for bit (the shared matrix work is the same code on the same
data), and those agree with the FP64 oracle to the batched engine's bar:
same status, x/y/s within 1e-6 relative.
"""
import numpy as np
import pytest

from qtqp_amd import batch, problems, warmstart
from oracle import qtqp_oracle as O

pytestmark = pytest.mark.gpu


def _family(count, seed, horizon=12, nx=4, nu=2):
  fam = warmstart.Family(horizon=horizon, nx=nx, nu=nu, family_seed=seed)
  x0s = problems.parametric_mpc_x0(count, seed=seed + 100, nx=nx)
  b, c = fam.bc(x0s)
  return fam, x0s, b, c


def _assert_same(ra, rb):
  for k in ("x", "y", "s"):
    np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)
  np.testing.assert_array_equal(ra["status"], rb["status"])
  np.testing.assert_array_equal(ra["iters"], rb["iters"])


def _assert_solutions(ra, sols):
  """results_arrays against results(): the device finish divides exactly as
  the host does (bitwise); a certificate's b'y or c'x sum may round
  differently (1e-13)."""
  st = np.array([s.status.value for s in sols])
  cert = st != "solved"
  for k in ("x", "y", "s"):
    want = np.stack([getattr(s, k) for s in sols])
    np.testing.assert_array_equal(ra[k][~cert], want[~cert], err_msg=k)
    np.testing.assert_allclose(ra[k][cert], want[cert], rtol=1e-13, atol=0, err_msg=k)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_family_batch_equals_ordinary_batch_bitwise(require_gpu, precision):
  fam, _, b, c = _family(40, seed=3)
  fb = batch.BatchQP.family(fam.a, fam.p, fam.z, b, c)
  ob = batch.BatchQP([(fam.a, b[i], c[i], fam.z, fam.p) for i in range(len(b))])
  fb.solve_device(precision=precision)
  ob.solve_device(precision=precision)
  _assert_same(fb.raw_iterates(), ob.raw_iterates())
  _assert_solutions(fb.results_arrays(), ob.results())
  fb.free()
  ob.free()


def test_family_batch_matches_oracle(require_gpu):
  fam, _, b, c = _family(24, seed=5, horizon=16, nx=6, nu=3)
  fb = batch.BatchQP.family(fam.a, fam.p, fam.z, b, c)
  fb.solve_device()
  ra = fb.results_arrays()
  names = {0: "unfinished", 1: "solved", 2: "infeasible", 3: "unbounded"}
  for i in range(len(b)):
    ref = O.solve_qp(fam.a, b[i], c[i], fam.z, fam.p)
    assert names[int(ra["status"][i])] == ref["status"], i
    assert abs(int(ra["iters"][i]) - ref["iters"]) <= 1
    for k in "xys":
      want = ref[k]
      err = np.max(np.abs(ra[k][i] - want)) / max(1.0, np.max(np.abs(want)))
      assert err < 1e-6, (i, k, err)
  fb.free()


def test_set_bc_resolves_like_a_new_batch(require_gpu):
  """A resident family batch re-solved with new b, c (and a second time with
  the first ones back) gives the results of batches created for them."""
  fam, _, b1, c1 = _family(32, seed=9)
  _, _, b2, c2 = _family(32, seed=10)
  fb = batch.BatchQP.family(fam.a, fam.p, fam.z, b1, c1)
  fb.solve_device()
  first = fb.raw_iterates()
  fb.set_bc(b2, c2)
  fb.solve_device()
  second = fb.raw_iterates()
  fresh = batch.BatchQP.family(fam.a, fam.p, fam.z, b2, c2)
  fresh.solve_device()
  _assert_same(second, fresh.raw_iterates())
  fb.set_bc(b1, c1)
  fb.solve_device()
  _assert_same(first, fb.raw_iterates())
  fresh.free()
  fb.free()


def test_set_bc_on_an_ordinary_batch(require_gpu):
  qps = problems.uneven_mpc_batch(12, base_seed=40, t_range=(4, 12), nx_range=(2, 6),
                                  nu_range=(1, 3))
  other = problems.uneven_mpc_batch(12, base_seed=40, t_range=(4, 12), nx_range=(2, 6),
                                    nu_range=(1, 3))
  rng = np.random.default_rng(0)
  b2 = [q.b + 0.1 * rng.standard_normal(q.b.shape) for q in other]
  c2 = [q.c + 0.1 * rng.standard_normal(q.c.shape) for q in other]
  ob = batch.BatchQP(qps)
  ob.set_bc(b2, c2)
  sols = ob.solve().solutions
  fresh = batch.BatchQP([(q.a, b2[i], c2[i], q.z, q.p) for i, q in enumerate(qps)])
  ref = fresh.solve().solutions
  for s1, s2 in zip(sols, ref):
    np.testing.assert_array_equal(s1.x, s2.x)
    np.testing.assert_array_equal(s1.y, s2.y)
    assert s1.status == s2.status and s1.iters == s2.iters
  ob.free()
  fresh.free()


def test_family_presolve_drops_and_restores_rows(require_gpu):
  """Rows with b = +inf in every QP are dropped once and restored per QP
  (s = b - a x of the dropped row is inf, y = 0), as the ordinary batch and
  the reference's presolve do (__init__.py:248-297)."""
  fam, _, b, c = _family(16, seed=12)
  b = b.copy()
  b[:, fam.m - 3:] = np.inf
  fb = batch.BatchQP.family(fam.a, fam.p, fam.z, b, c)
  ob = batch.BatchQP([(fam.a, b[i], c[i], fam.z, fam.p) for i in range(len(b))])
  fb.solve_device()
  ra = fb.results_arrays()
  _assert_solutions(ra, ob.solve().solutions)
  assert np.all(ra["y"][:, fam.m - 3:] == 0.0) and np.all(np.isinf(ra["s"][:, fam.m - 3:]))
  with pytest.raises(ValueError, match="other rows"):
    b3 = b.copy()
    b3[:, fam.m - 3:] = 1.0
    fb.set_bc(b3, c)
  fb.free()
  ob.free()


def test_family_autotune_keeps_results(require_gpu):
  fam, _, b, c = _family(48, seed=14)
  fb = batch.BatchQP.family(fam.a, fam.p, fam.z, b, c)
  fb.solve_device()
  want = fb.raw_iterates()
  fb.autotune(sample=8, iters=2)
  fb.solve_device()
  _assert_same(want, fb.raw_iterates())
  fb.free()
