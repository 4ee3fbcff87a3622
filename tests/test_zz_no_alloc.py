"""qtqp_kkt_matvec_noalloc (the entry point that allocates nothing) against
qtqp_kkt_matvec on the same K_reg and x: the same kernels in the same order on
the same data, so the products must be bitwise equal.  Two calls in a row on
one GPU, the second with a smaller K, share the fixed scratch array; an
oversized K is rejected before anything runs.  The test's own device buffers
come from the HIP runtime the library is linked against (hipMalloc through
ctypes); the entry point under test allocates none.
"""
import ctypes

import numpy as np
import pytest

from qtqp_amd import _native

pytestmark = pytest.mark.gpu


def _system(n, m, seed, with_p=True):
  rng = np.random.default_rng(seed)
  a = rng.standard_normal((m, n))
  r = rng.standard_normal((n, n))
  p = r @ r.T / n if with_p else None
  diag = np.concatenate([rng.uniform(0.1, 2.0, n), -rng.uniform(0.05, 2.0, m)])
  x = rng.standard_normal(n + m)
  return a, p, diag, x


def _handle_matvec(lib, a, p, diag, x):
  m, n = a.shape
  h = ctypes.c_void_p()
  _native.check(lib, lib.qtqp_kkt_create(n, m, _native.dptr(a),
                                         None if p is None else _native.dptr(p),
                                         _native.PREC_FP64, ctypes.byref(h)))
  try:
    _native.check(lib, lib.qtqp_kkt_update_diag(h, _native.dptr(diag)))
    out = np.empty(n + m)
    _native.check(lib, lib.qtqp_kkt_matvec(h, _native.dptr(x), _native.dptr(out)))
  finally:
    lib.qtqp_kkt_free(h)
  return out


class _Hip:
  """hipMalloc / hipMemcpy / hipFree of the library's HIP runtime (test-side buffers)."""

  H2D, D2H = 1, 2

  def __init__(self):
    self.rt = ctypes.CDLL("libamdhip64.so.7")
    self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    self.rt.hipFree.argtypes = [ctypes.c_void_p]
    self.bufs = []

  def upload(self, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    p = ctypes.c_void_p()
    assert self.rt.hipMalloc(ctypes.byref(p), max(v.nbytes, 8)) == 0
    self.bufs.append(p)
    assert self.rt.hipMemcpy(p, v.ctypes.data, v.nbytes, self.H2D) == 0
    return p

  def download(self, p, count):
    out = np.empty(count)
    assert self.rt.hipDeviceSynchronize() == 0
    assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
    return out

  def free(self):
    assert self.rt.hipDeviceSynchronize() == 0
    for p in self.bufs:
      self.rt.hipFree(p)
    self.bufs = []


def _noalloc_matvec(lib, hip, a, p, diag, x):
  m, n = a.shape
  poff = None
  if p is not None:
    poff = p.copy()
    np.fill_diagonal(poff, 0.0)
  a_d, d_d, x_d = hip.upload(a), hip.upload(diag), hip.upload(x)
  poff_d = None if poff is None else hip.upload(poff)
  out_d = hip.upload(np.full(n + m, np.nan))
  rc = lib.qtqp_kkt_matvec_noalloc(n, m, a_d, poff_d, d_d, x_d, out_d, None)
  _native.check(lib, rc)
  return hip.download(out_d, n + m)


def test_noalloc_matvec_matches_handle_matvec(require_gpu):
  lib = _native.kkt_lib()
  hip = _Hip()
  # the larger call first, then a smaller one and a P = 0 one on the same scratch
  for n, m, seed, with_p in ((700, 1500, 1, True), (130, 260, 2, True), (40, 90, 3, False)):
    print(f"[case] noalloc matvec n={n} m={m} P={'yes' if with_p else 'no'}", flush=True)
    a, p, diag, x = _system(n, m, seed, with_p)
    want = _handle_matvec(lib, a, p, diag, x)
    got = _noalloc_matvec(lib, hip, a, p, diag, x)
    hip.free()
    np.testing.assert_array_equal(got, want)
    poff = p - np.diag(np.diag(p)) if with_p else np.zeros((n, n))
    k = np.block([[poff + np.diag(diag[:n]), a.T], [a, np.diag(diag[n:])]])
    np.testing.assert_allclose(got, k @ x, rtol=1e-12, atol=1e-12)


def test_noalloc_matvec_rejects_oversized_k(require_gpu):
  lib = _native.kkt_lib()
  # (n + m) * ceil(n / 256) + ceil(m / 32) * n + n > 2^21 for n = m = 20000
  rc = lib.qtqp_kkt_matvec_noalloc(20000, 20000, ctypes.c_void_p(256), None, ctypes.c_void_p(256),
                                   ctypes.c_void_p(256), ctypes.c_void_p(256), None)
  assert rc == _native.QTQP_ERR_ARG
  assert b"16 MiB" in lib.qtqp_last_error()
