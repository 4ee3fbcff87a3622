"""Seeded synthetic QP families used by the benchmarks and the parity tests.

All generators draw from ``numpy.random.default_rng(seed)`` (PCG64), which is
bit-reproducible across platforms, so an instance is fully named by
(family, sizes, seed).  They return dense arrays:
``(a[m,n], b[m], c[n], z, p[n,n])`` with qtqp's convention
``min 1/2 x'Px + c'x  s.t.  Ax + s = b, s[:z] = 0, s[z:] >= 0``.
"""

from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class DenseQP:
  a: np.ndarray
  b: np.ndarray
  c: np.ndarray
  z: int
  p: np.ndarray
  # Known optimum when the generator constructs one (else None).
  x_star: np.ndarray | None = None
  y_star: np.ndarray | None = None
  s_star: np.ndarray | None = None

  @property
  def n(self) -> int:
    return self.a.shape[1]

  @property
  def m(self) -> int:
    return self.a.shape[0]


def random_dense_qp(n: int, m: int, z: int, seed: int = 0) -> DenseQP:
  """Random convex QP with a planted strictly complementary optimum.

  P = M M'/n + 1e-3 I (M ~ N(0,1), n x n), A ~ N(0,1) (m x n), x0 ~ N(0,1).
  Equality rows: s0 = 0, y0 ~ N(0,1).  Inequality rows: a random half is
  active (s0 = 0, y0 = |N(0,1)|), the rest inactive (s0 = |N(0,1)|, y0 = 0).
  b = A x0 + s0, c = -P x0 - A' y0, so (x0, y0, s0) satisfies the KKT
  conditions; P > 0 makes x0 unique and, with fewer active rows than n,
  y0 and s0 are unique almost surely.
  """
  rng = np.random.default_rng(seed)
  mm = rng.standard_normal((n, n))
  p = mm @ mm.T / n + 1e-3 * np.eye(n)
  a = rng.standard_normal((m, n))
  x0 = rng.standard_normal(n)
  n_ineq = m - z
  active = np.zeros(n_ineq, dtype=bool)
  active[rng.permutation(n_ineq)[: n_ineq // 2]] = True
  mag_s = np.abs(rng.standard_normal(n_ineq))
  mag_y = np.abs(rng.standard_normal(n_ineq))
  y_eq = rng.standard_normal(z)
  s0 = np.zeros(m)
  y0 = np.zeros(m)
  y0[:z] = y_eq
  s0[z:] = np.where(active, 0.0, mag_s)
  y0[z:] = np.where(active, mag_y, 0.0)
  b = a @ x0 + s0
  c = -p @ x0 - a.T @ y0
  return DenseQP(a, b, c, z, p, x0, y0, s0)


@dataclasses.dataclass
class MpcSystem:
  ad: np.ndarray  # nx x nx, spectral radius 0.95
  bd: np.ndarray  # nx x nu


def mpc_system(nx: int, nu: int, rng: np.random.Generator) -> MpcSystem:
  ad = rng.standard_normal((nx, nx))
  rho = np.max(np.abs(np.linalg.eigvals(ad)))
  ad = ad * (0.95 / rho)
  bd = rng.standard_normal((nx, nu)) / np.sqrt(nx)
  return MpcSystem(ad, bd)


def condensed_mpc(sys_: MpcSystem, horizon: int, x0: np.ndarray, r_weight=0.1,
                  u_max=1.0, x_max=5.0) -> DenseQP:
  """Condensed linear MPC: decision U = (u_0..u_{T-1}), states x_1..x_T.

  x_{t+1} = A x_t + B u_t, so X = Phi x0 + Gamma U with Phi = [A; ...; A^T] and
  Gamma[t, k] = A^(t-k) B for k <= t.  Objective 1/2 (X'X + r U'U), i.e.
  P = Gamma'Gamma + r I and c = Gamma' Phi x0 (Q = I, R = r I).  Constraints
  |u| <= u_max and |x_t| <= x_max as rows [I; -I; Gamma; -Gamma], z = 0.
  """
  ad, bd = sys_.ad, sys_.bd
  nx, nu = bd.shape
  t_h = horizon
  pows = [np.eye(nx)]
  for _ in range(t_h):
    pows.append(ad @ pows[-1])
  phi = np.concatenate([pows[t + 1] for t in range(t_h)], axis=0)
  # Block-Toeplitz Gamma: block (t, k) = A^(t-k) B for k <= t, filled one
  # block diagonal at a time.
  blocks = np.stack([pw @ bd for pw in pows[:t_h]])  # (T, nx, nu)
  g4 = np.zeros((t_h, nx, t_h, nu))
  for lag in range(t_h):
    idx = np.arange(t_h - lag)
    g4[idx + lag, :, idx, :] = blocks[lag]
  gam = g4.reshape(t_h * nx, t_h * nu)
  n = t_h * nu
  p = gam.T @ gam + r_weight * np.eye(n)
  p = 0.5 * (p + p.T)
  fx = phi @ x0
  c = gam.T @ fx
  eye = np.eye(n)
  a = np.concatenate([eye, -eye, gam, -gam], axis=0)
  b = np.concatenate([np.full(n, u_max), np.full(n, u_max), x_max - fx, x_max + fx])
  return DenseQP(a, b, c, 0, p)


def mpc_qp(horizon: int, nx: int, nu: int, seed: int = 0) -> DenseQP:
  """Single condensed MPC QP: system and x0 ~ N(0, I) drawn from ``seed``."""
  rng = np.random.default_rng(seed)
  sys_ = mpc_system(nx, nu, rng)
  x0 = rng.standard_normal(nx)
  return condensed_mpc(sys_, horizon, x0)


def uneven_mpc_batch(count: int, base_seed: int = 1000, t_range=(8, 64),
                     nx_range=(4, 16), nu_range=(1, 6)) -> list[DenseQP]:
  """Uneven MPC batch: instance i uses seed base_seed+i for T, nx, nu and data."""
  out = []
  for i in range(count):
    rng = np.random.default_rng(base_seed + i)
    t_h = int(rng.integers(t_range[0], t_range[1] + 1))
    nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
    nu = int(rng.integers(nu_range[0], nu_range[1] + 1))
    sys_ = mpc_system(nx, nu, rng)
    x0 = rng.standard_normal(nx)
    out.append(condensed_mpc(sys_, t_h, x0))
  return out


def parametric_mpc_family(family_seed: int = 7, horizon: int = 32, nx: int = 8,
                          nu: int = 3) -> MpcSystem:
  """Fixed (A_d, B_d) of the warm-start family; only x0 varies per instance."""
  return mpc_system(nx, nu, np.random.default_rng(family_seed))


def parametric_mpc_x0(count: int, seed: int, nx: int = 8, std: float = 0.5) -> np.ndarray:
  return std * np.random.default_rng(seed).standard_normal((count, nx))


@dataclasses.dataclass
class SparseQP:
  """A QP with scipy CSC data (``a`` m x n, ``p`` n x n) and a planted KKT point."""
  a: "scipy.sparse.csc_matrix"
  b: np.ndarray
  c: np.ndarray
  z: int
  p: "scipy.sparse.csc_matrix"
  x0: np.ndarray
  y0: np.ndarray
  s0: np.ndarray

  @property
  def n(self) -> int:
    return self.a.shape[1]

  @property
  def m(self) -> int:
    return self.a.shape[0]


def random_sparse_qp(m: int, n: int, z: int, seed: int = 0, density: float = 0.1) -> SparseQP:
  """The reference test suite's feasible sparse family (tests/test_qtqp.py:211-235),
  kept sparse: A ~ N(0,1) entries at ``density``, P = R'R / 100 with R ~ N(0,1) at
  density 0.01, and b, c built from a random primal-dual pair (x, y, s) that is
  complementary on the inequality rows."""
  import scipy.sparse as sp
  rng = np.random.default_rng(seed)
  w = rng.normal(size=m)
  x = rng.normal(size=n)
  y = w.copy()
  y[z:] = 0.5 * (w[z:] + np.abs(w[z:]))
  s = y - w
  a = sp.random(m, n, density=density, format="csc", random_state=rng,
                data_rvs=lambda k: rng.normal(size=k))
  r = sp.random(n, n, density=0.01, format="csc", random_state=rng,
                data_rvs=lambda k: rng.normal(size=k))
  p = sp.csc_matrix(r.T @ r * 0.01)
  return SparseQP(sp.csc_matrix(a), a @ x + s, -(a.T @ y), z, p, x, y, s)
