"""On-disk formats: packed batch-instance files (CPU only)."""
import numpy as np
import pytest

from qtqp_amd import batch, problems


def test_instance_file_round_trip(tmp_path):
  qps = problems.uneven_mpc_batch(5, base_seed=40)
  qps.append((np.array([[1.0, 2.0]]), np.array([3.0]), np.array([0.5, -1.0]), 0, None))  # LP
  path = tmp_path / "inst.npz"
  batch.save_instances(path, qps, seeds=[40, 41, 42, 43, 44, -1])
  back, seeds = batch.load_instances(path)
  assert seeds.tolist() == [40, 41, 42, 43, 44, -1]
  assert len(back) == len(qps)
  for q, (a, b, c, z, p) in zip(qps, back):
    qa, qb, qc, qz, qp = (q.a, q.b, q.c, q.z, q.p) if hasattr(q, "a") else q
    np.testing.assert_array_equal(a, qa)
    np.testing.assert_array_equal(b, qb)
    np.testing.assert_array_equal(c, qc)
    assert z == qz
    np.testing.assert_array_equal(p, np.zeros((a.shape[1],) * 2) if qp is None else qp)


def test_instance_file_rejects_other_archives(tmp_path):
  path = tmp_path / "other.npz"
  np.savez(path, format=np.array("something else"))
  with pytest.raises(ValueError):
    batch.load_instances(path)
  with pytest.raises(ValueError):
    batch.save_instances(tmp_path / "x.npz", problems.uneven_mpc_batch(2, base_seed=1), seeds=[1])
