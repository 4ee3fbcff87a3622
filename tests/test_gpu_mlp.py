"""GPU MLP kernels (BF16 MFMA, FP32 accumulation) against the FP64 oracle.

The oracle rounds every matrix-product operand to bfloat16 at the same points
as the kernels, so forward values and gradients must agree to FP32
accumulation error: relative 1e-4 (forward) and 1e-3 per parameter block
(gradients).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from qtqp_amd import warmstart

pytestmark = pytest.mark.gpu

SIZES = [(8, 256, 1506, 64), (3, 20, 70, 33)]


def _blocks(g, n_in, hid, n_out):
  sizes = [n_in * hid, hid, hid * hid, hid, hid * n_out, n_out]
  out, o = [], 0
  for sz in sizes:
    out.append(g[o:o + sz])
    o += sz
  return out


@pytest.mark.parametrize("n_in,hid,n_out,b", SIZES)
def test_forward_matches_oracle(require_gpu, n_in, hid, n_out, b):
  rng = np.random.default_rng(1)
  net = warmstart.MLP(n_in, hid, n_out, max_batch=64, seed=3)
  p = net.get_params().astype(np.float64)
  p[n_in * hid:n_in * hid + hid] = rng.standard_normal(hid) * 0.1  # nonzero biases
  net.set_params(p)
  x = rng.standard_normal((b, n_in)).astype(np.float32)
  y = net.forward(x)
  ref, _ = M.forward(net.get_params(), x, n_in, hid, n_out, bf16_operands=True)
  err = np.max(np.abs(y - ref)) / max(1e-30, np.max(np.abs(ref)))
  assert err < 1e-4, err
  net.free()


@pytest.mark.parametrize("n_in,hid,n_out,b", SIZES)
def test_gradient_matches_oracle(require_gpu, n_in, hid, n_out, b):
  rng = np.random.default_rng(2)
  net = warmstart.MLP(n_in, hid, n_out, max_batch=64, seed=5)
  x = rng.standard_normal((b, n_in)).astype(np.float32)
  t = rng.standard_normal((b, n_out)).astype(np.float32)
  loss, g = net.loss_grad(x, t)
  rloss, rg = M.loss_grad(net.get_params(), x, t, n_in, hid, n_out, bf16_operands=True)
  assert abs(loss - rloss) < 1e-5 * max(1.0, rloss)
  for k, (a, r) in enumerate(zip(_blocks(g, n_in, hid, n_out), _blocks(rg, n_in, hid, n_out))):
    rel = np.linalg.norm(a - r) / max(1e-30, np.linalg.norm(r))
    assert rel < 1e-3, (k, rel)
  net.free()


def test_training_reduces_loss(require_gpu):
  rng = np.random.default_rng(4)
  x = rng.standard_normal((4096, 8)).astype(np.float32)
  w = rng.standard_normal((8, 40))
  t = np.tanh(x @ w / 3.0).astype(np.float32)
  net = warmstart.MLP(8, 64, 40, max_batch=256, seed=0)
  losses = net.train(x, t, batch=256, steps=600, lr=3e-3)
  assert np.all(np.isfinite(losses))
  assert losses[-20:].mean() < 0.1 * losses[:20].mean()
  net.free()
