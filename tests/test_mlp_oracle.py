"""Pins the FP64 MLP oracle: central finite differences and bf16 rounding."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
  vals = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-9, -2.5, 0.0, 3.0e38], dtype=np.float32)
  got = M.bf16(vals)
  # 1+2^-8 is a tie between 1 and 1+2^-7 -> even (1.0); 1+3*2^-9 rounds up.
  np.testing.assert_array_equal(got[:5], [1.0, 1.0, 1.0 + 2.0**-7, -2.5, 0.0])
  assert np.isfinite(got[5]) or got[5] == np.inf


def test_gradient_matches_central_differences():
  rng = np.random.default_rng(0)
  n_in, hid, n_out, b = 3, 7, 5, 6
  p = rng.standard_normal(n_in * hid + hid + hid * hid + hid + hid * n_out + n_out) * 0.5
  x = rng.standard_normal((b, n_in))
  t = rng.standard_normal((b, n_out))
  _, g = M.loss_grad(p, x, t, n_in, hid, n_out, bf16_operands=False)
  h = 1e-6
  for k in rng.choice(p.size, 25, replace=False):
    e = np.zeros_like(p)
    e[k] = h
    lp, _ = M.loss_grad(p + e, x, t, n_in, hid, n_out, bf16_operands=False)
    lm, _ = M.loss_grad(p - e, x, t, n_in, hid, n_out, bf16_operands=False)
    np.testing.assert_allclose(g[k], (lp - lm) / (2 * h), rtol=1e-6, atol=1e-9)


def test_planted_gelu_derivative_mistake_is_caught():
  rng = np.random.default_rng(1)
  n_in, hid, n_out = 2, 4, 3
  p = rng.standard_normal(n_in * hid + hid + hid * hid + hid + hid * n_out + n_out)
  x, t = rng.standard_normal((4, n_in)), rng.standard_normal((4, n_out))
  good = M._gelu_grad
  M._gelu_grad = lambda z: 0.5 * (1.0 + M.erf(z / np.sqrt(2.0)))  # drops the z*phi(z) term
  try:
    _, g = M.loss_grad(p, x, t, n_in, hid, n_out, bf16_operands=False)
  finally:
    M._gelu_grad = good
  e = np.zeros_like(p)
  e[0] = 1e-6
  lp, _ = M.loss_grad(p + e, x, t, n_in, hid, n_out, bf16_operands=False)
  lm, _ = M.loss_grad(p - e, x, t, n_in, hid, n_out, bf16_operands=False)
  assert abs(g[0] - (lp - lm) / 2e-6) > 1e-6 * max(1.0, abs(g[0]))
