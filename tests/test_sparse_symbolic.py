"""Symbolic analysis of the sparse KKT engine (host only, no GPU): the
elimination order is a permutation and the factor's fill, flop count and
supernode tables agree with the elimination game of the oracle, itself pinned
by closed forms and by numpy's Cholesky factor."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import sparse_oracle as SO
from qtqp_amd import direct
import qp_generators as G


def _arrow(n, hub_first):
  hub = 0 if hub_first else n - 1
  others = [i for i in range(n) if i != hub]
  k = sp.lil_matrix((n, n))
  for i in range(n):
    k[i, i] = 1.0
  for i in others:
    k[min(i, hub), max(i, hub)] = 1.0
  return sp.csc_matrix(k)


def _tridiag(n):
  return sp.csc_matrix(sp.diags([np.ones(n - 1), np.ones(n)], [1, 0]))


def _kkt_scaffold(m, n, z, seed, density):
  a, b, c, p = G.gen_feasible(m, n, z, seed, density=density)
  a, p = sp.csc_matrix(a), sp.csc_matrix(p)
  return sp.bmat([[sp.triu(p, format="csc") + sp.eye(n, format="csc"), a.T],
                  [None, sp.eye(m, format="csc")]], format="csc")


# -- the oracle against closed forms and numpy --------------------------------

@pytest.mark.parametrize("n", [2, 7, 40])
def test_oracle_closed_forms(n):
  natural = np.arange(n)
  adj = SO.adjacency_from_upper(_arrow(n, hub_first=False))
  assert SO.elimination_counts(adj, natural).sum() == 2 * n - 1
  adj = SO.adjacency_from_upper(_arrow(n, hub_first=True))
  assert SO.elimination_counts(adj, natural).sum() == n * (n + 1) // 2
  adj = SO.adjacency_from_upper(_tridiag(n))
  assert SO.elimination_counts(adj, natural).sum() == 2 * n - 1
  assert list(SO.etree(adj, natural)) == list(range(1, n)) + [-1]
  dense = sp.csc_matrix(np.triu(np.ones((n, n))))
  assert SO.elimination_counts(SO.adjacency_from_upper(dense), natural).sum() == n * (n + 1) // 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_counts_match_numpy_cholesky(seed):
  rng = np.random.default_rng(seed)
  n = 30
  up = sp.triu(sp.random(n, n, density=0.08, random_state=rng), k=1)
  pattern = (up + up.T).toarray() != 0
  spd = np.where(pattern, rng.uniform(0.1, 1.0, (n, n)), 0.0)
  spd = np.triu(spd, 1)
  spd = spd + spd.T + n * np.eye(n)
  chol = np.linalg.cholesky(spd)
  counts = SO.elimination_counts(SO.adjacency_from_upper(sp.csc_matrix(np.triu(spd))),
                                 np.arange(n))
  assert counts.sum() == np.count_nonzero(np.abs(chol) > 0)
  assert list(counts) == [np.count_nonzero(chol[j:, j]) for j in range(n)]


# -- the library's analysis against the oracle -------------------------------

def _check_against_oracle(kkt):
  perm, info = direct.sparse_analyze(kkt)
  n = kkt.shape[0]
  assert sorted(perm.tolist()) == list(range(n))
  counts = SO.elimination_counts(SO.adjacency_from_upper(kkt), perm)
  assert info["nnz_l"] == counts.sum()
  below = counts - 1.0
  assert info["factor_flops"] == pytest.approx(float(np.sum(below + below * (below + 1))))
  assert 1 <= info["supernodes"] <= n and 1 <= info["levels"] <= info["supernodes"]
  assert info["max_rows"] == counts.max()
  return perm, info


@pytest.mark.parametrize("n", [5, 64, 130])
def test_analysis_closed_forms(n):
  perm, info = _check_against_oracle(_arrow(n, hub_first=True))
  assert info["nnz_l"] == 2 * n - 1  # minimum degree leaves the hub to the end
  _, info = _check_against_oracle(_tridiag(n))
  assert info["nnz_l"] == 2 * n - 1
  _, info = _check_against_oracle(sp.csc_matrix(np.triu(np.ones((n, n)))))
  assert info["nnz_l"] == n * (n + 1) // 2
  assert info["supernodes"] == -(-n // 64)  # one supernode, split every 64 columns


@pytest.mark.parametrize("mnz,density", [((15, 10, 2), 0.3), ((60, 40, 5), 0.1),
                                         ((150, 100, 10), 0.1), ((90, 120, 0), 0.05)])
def test_analysis_of_kkt_scaffolds(mnz, density):
  m, n, z = mnz
  _check_against_oracle(_kkt_scaffold(m, n, z, 3, density))


def test_analysis_csr_equals_csc():
  kkt = _kkt_scaffold(40, 30, 4, 5, 0.1)
  lib = direct._native.sparse_lib()
  out = []
  for fmt, mat in ((direct._native.SPARSE_CSC, sp.csc_matrix(kkt)),
                   (direct._native.SPARSE_CSR, sp.csr_matrix(kkt))):
    mat.sort_indices()
    ptr, ind = direct._native.as_i64(mat.indptr), direct._native.as_i64(mat.indices)
    perm = np.empty(kkt.shape[0], dtype=np.int64)
    inf = direct._native.SparseInfo()
    rc = lib.qtqp_sparse_analyze(kkt.shape[0], fmt, direct._native.iptr(ptr),
                                 direct._native.iptr(ind), direct._native.iptr(perm),
                                 direct._native.ctypes.byref(inf))
    assert rc == 0
    out.append((perm.tolist(), inf.nnz_l, inf.supernodes, inf.levels))
  assert out[0] == out[1]


def test_analysis_rejects_lower_entries():
  kkt = sp.csc_matrix(np.ones((3, 3)))
  with pytest.raises(ValueError, match="upper triangle"):
    direct.sparse_analyze(kkt)


# -- approximate minimum degree (the ordering above N = 40000) ---------------

@pytest.mark.parametrize("mnz,density", [((15, 10, 2), 0.3), ((60, 40, 5), 0.1),
                                         ((150, 100, 10), 0.1), ((90, 120, 0), 0.05)])
def test_amd_analysis_matches_the_oracle(mnz, density):
  """With the approximate minimum degree order the analysis still agrees with
  the elimination game on fill, flops and supernode tables."""
  m, n, z = mnz
  kkt = _kkt_scaffold(m, n, z, 3, density)
  perm, info = direct.sparse_analyze(kkt, ordering=2)
  assert sorted(perm.tolist()) == list(range(kkt.shape[0]))
  counts = SO.elimination_counts(SO.adjacency_from_upper(kkt), perm)
  assert info["nnz_l"] == counts.sum()
  below = counts - 1.0
  assert info["factor_flops"] == pytest.approx(float(np.sum(below + below * (below + 1))))


@pytest.mark.parametrize("mnz,density", [((1500, 1000, 100), 0.01), ((3000, 2000, 100), 0.01)])
def test_amd_fill_is_close_to_exact_minimum_degree(mnz, density):
  m, n, z = mnz
  kkt = _kkt_scaffold(m, n, z, 3, density)
  _, exact = direct.sparse_analyze(kkt, ordering=1)
  _, approx = direct.sparse_analyze(kkt, ordering=2)
  assert approx["nnz_l"] <= 1.05 * exact["nnz_l"]
  _, auto = direct.sparse_analyze(kkt)
  assert auto["nnz_l"] == exact["nnz_l"]  # exact below N = 40000


def _random_rows_kkt(n, seed=7):
  """KKT scaffold of n variables and 3n constraints with 3 random entries each,
  P tridiagonal (P^-1 is dense, so eliminating x first fills the y block)."""
  rng = np.random.default_rng(seed)
  m = 3 * n
  rows = np.repeat(np.arange(m), 3)
  cols = rng.integers(0, n, 3 * m)
  a = sp.csc_matrix((rng.standard_normal(3 * m), (rows, cols)), shape=(m, n))
  p = sp.diags([np.ones(n - 1), 4.0 * np.ones(n)], [1, 0], format="csc")
  return sp.bmat([[p + sp.eye(n, format="csc"), a.T], [None, sp.eye(m, format="csc")]],
                 format="csc")


def test_amd_fill_on_random_rows_is_close_to_exact():
  kkt = _random_rows_kkt(4000)  # N = 16000
  _, exact = direct.sparse_analyze(kkt, ordering=1)
  _, approx = direct.sparse_analyze(kkt, ordering=2)
  assert approx["nnz_l"] <= 1.05 * exact["nnz_l"]


def test_analysis_above_40000_orders_for_fill():
  """N = 48000 (12000 variables, 36000 constraints): the default analysis orders
  by approximate minimum degree.  The natural [x; y] order would fill the whole
  36000 x 36000 constraint block (6.5e8 entries); the ordering keeps nnz(L)
  under a tenth of that, and the order is a permutation."""
  n, m = 12000, 36000
  kkt = _random_rows_kkt(n)
  perm, info = direct.sparse_analyze(kkt)
  assert kkt.shape[0] == 48000
  assert np.array_equal(np.sort(perm), np.arange(48000))
  assert info["nnz_l"] < 0.1 * m * m / 2, info["nnz_l"]
