"""CPU restatement of the sparse KKT engine's symbolic analysis -- TEST
INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; the shipped solver never routes
through it.  The reference delegates sparse factorisation to third-party
libraries (QDLDL via ``qdldl``, cuDSS via ``nvmath-python``;
src/qtqp/solvers_sparse.py:74-93, src/qtqp/solvers_gpu.py:27-115), none of
them present here; their published algorithm is the textbook one restated
below: symbolic LDL' by the elimination game on the graph of K.

* ``elimination_counts(adj, order)`` -- column counts of L (diagonal included)
  when the vertices are eliminated in ``order``: eliminating a vertex joins its
  remaining neighbours into a clique.  Brute force, small N only.
* ``etree(adj, order)`` -- the elimination tree: the parent of column k is the
  smallest later column in k's structure.

Pinning (tests/test_sparse_symbolic.py): closed forms that a plausible slip
(counting the diagonal twice, dropping fill, eliminating in the wrong order)
fails -- an arrow matrix eliminated hub-last has nnz(L) = 2N-1 and hub-first
N(N+1)/2; a tridiagonal matrix has 2N-1; a dense one N(N+1)/2; and on random
patterns the count equals the nonzero count of numpy's Cholesky factor of an
SPD matrix with that pattern (generic values, no cancellation).
"""

from __future__ import annotations

import numpy as np


def adjacency_from_upper(kkt) -> list[set[int]]:
  """Symmetric off-diagonal adjacency of an upper-triangular scipy matrix."""
  coo = kkt.tocoo()
  n = kkt.shape[0]
  adj = [set() for _ in range(n)]
  for i, j in zip(coo.row.tolist(), coo.col.tolist()):
    if i != j:
      adj[i].add(j)
      adj[j].add(i)
  return adj


def elimination_counts(adj: list[set[int]], order) -> np.ndarray:
  """Column counts of L, in elimination order, by the elimination game."""
  g = [set(a) for a in adj]
  done = set()
  counts = []
  for v in order:
    v = int(v)
    nb = g[v] - done
    counts.append(len(nb) + 1)
    for u in nb:
      g[u] |= nb
      g[u].discard(u)
    done.add(v)
  return np.array(counts, dtype=np.int64)


def etree(adj: list[set[int]], order) -> np.ndarray:
  """Parent (in elimination positions) of each column; -1 for roots."""
  pos = {int(v): k for k, v in enumerate(order)}
  g = [set(a) for a in adj]
  done = set()
  parent = np.full(len(order), -1, dtype=np.int64)
  for k, v in enumerate(order):
    v = int(v)
    nb = g[v] - done
    if nb:
      parent[k] = min(pos[u] for u in nb)
    for u in nb:
      g[u] |= nb
      g[u].discard(u)
    done.add(v)
  return parent
