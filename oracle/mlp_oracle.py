"""FP64 oracle of the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module.  It restates the network of
qtqp_amd/csrc/mlp.hip in numpy float64: three layers n_in -> hidden ->
hidden -> n_out, exact GELU x Phi(x) (Hendrycks & Gimpel 2016) between them,
mean-squared-error loss, and the backward pass.  With ``bf16_operands`` every
matrix-product operand is rounded to bfloat16 (round to nearest even) at the
same points as the GPU kernels, so the only remaining difference is the
GPU's FP32 accumulation order.  Pinned by central finite differences
(tests/test_mlp_oracle.py).  The network has no counterpart in the reference.
"""
import numpy as np
from scipy.special import erf


def bf16(a):
  """Round float32 values to bfloat16 (nearest even); returns float64."""
  u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
  u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
  return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def _gelu(z):
  return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def _gelu_grad(z):
  return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(params, n_in, hid, n_out):
  p = np.asarray(params, dtype=np.float64)
  sizes = [n_in * hid, hid, hid * hid, hid, hid * n_out, n_out]
  out, o = [], 0
  for sz in sizes:
    out.append(p[o:o + sz])
    o += sz
  w1, b1, w2, b2, w3, b3 = out
  return w1.reshape(n_in, hid), b1, w2.reshape(hid, hid), b2, w3.reshape(hid, n_out), b3


def forward(params, x, n_in, hid, n_out, bf16_operands=True):
  q = bf16 if bf16_operands else (lambda a: np.asarray(a, dtype=np.float64))
  w1, b1, w2, b2, w3, b3 = unpack(params, n_in, hid, n_out)
  z1 = q(x) @ q(w1) + b1
  h1 = _gelu(z1)
  z2 = q(h1) @ q(w2) + b2
  h2 = _gelu(z2)
  o = q(h2) @ q(w3) + b3
  return o, (z1, h1, z2, h2)


def loss_grad(params, x, t, n_in, hid, n_out, bf16_operands=True):
  """Mean-squared error over all outputs and its gradient (packed like params)."""
  q = bf16 if bf16_operands else (lambda a: np.asarray(a, dtype=np.float64))
  w1, b1, w2, b2, w3, b3 = unpack(params, n_in, hid, n_out)
  o, (z1, h1, z2, h2) = forward(params, x, n_in, hid, n_out, bf16_operands)
  d = o - np.asarray(t, dtype=np.float64)
  loss = float(np.mean(d * d))
  do = 2.0 * d / d.size
  gw3 = q(h2).T @ q(do)
  gb3 = do.sum(0)
  dz2 = (q(do) @ q(w3).T) * _gelu_grad(z2)
  gw2 = q(h1).T @ q(dz2)
  gb2 = dz2.sum(0)
  dz1 = (q(dz2) @ q(w2).T) * _gelu_grad(z1)
  gw1 = q(x).T @ q(dz1)
  gb1 = dz1.sum(0)
  g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])
  return loss, g
