"""CPU: pins of the proposal network's oracle (oracle/cnn_oracle.py). The network has no counterpart
in the reference; these checks pin the restatement itself.

* bfloat16 rounding: exact values stay, halfway cases round to even, overflow to infinity
* the convolution equals scipy's 2-D correlation with zero padding, then the stride's subsampling
  (a flipped kernel, a wrong tap offset or stride phase fails it)
* im2col / col2im are adjoint (<im2col x, y> = <x, col2im y>)
* the backward pass (float64) equals central finite differences of the loss for every parameter
  of a small network, for both loss terms, the positive weight and cells with and without a bin
* closed forms of the loss at zero logits: the location term is mean(w y + (1 - y)) log 2, the
  orientation term log(n_bins)
"""
import numpy as np
import pytest
from scipy import signal

from oracle import cnn_oracle as C


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.4e38, 1 + 2 ** -7], np.float32)
    r = C.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 ** -6 and r[3] == -2.5 and r[5] == 1 + 2 ** -7
    assert np.isinf(r[4])


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_is_correlation(stride):
    rng = np.random.default_rng(0)
    S = 12
    img = rng.standard_normal((S, S))
    k = rng.standard_normal((3, 3))
    cfg = C.Config(S, [1], [stride], 2)
    flat = np.concatenate([k.ravel(), [0.0], np.zeros(3), np.zeros(3)])
    acts = C.forward(cfg, flat, img[None], use_bf16=False)
    x = C.standardize(img[None])[0]
    want = signal.correlate2d(x, k, mode="same", boundary="fill")[::stride, ::stride]
    assert np.allclose(acts[1][0, :, :, 0], np.maximum(want, 0), atol=1e-12)


def test_im2col_adjoint():
    rng = np.random.default_rng(1)
    for stride in (1, 2):
        X = rng.standard_normal((2, 8, 8, 3))
        col = C.im2col(X, stride)
        Y = rng.standard_normal(col.shape)
        assert np.isclose((col * Y).sum(), (X * C.col2im(Y, 8, stride)).sum())


def _tiny():
    cfg = C.Config(8, [3, 4], [2, 1], 5)
    rng = np.random.default_rng(2)
    flat = C.init_params(cfg, 3).astype(np.float64)
    flat += 0.05 * rng.standard_normal(flat.size)   # non-zero biases
    images = rng.standard_normal((2, 8, 8))
    G = cfg.grid
    heat = np.zeros((2, G, G), np.float32)
    bins = -np.ones((2, G, G), np.int32)
    heat[0, 1, 2] = 1
    heat[1, 3, 0] = 1
    bins[0, 1, 2] = 4
    bins[1, 3, 0] = 1
    bins[1, 2, 2] = 0
    return cfg, flat, images, heat, bins


def test_backward_equals_finite_differences():
    cfg, flat, images, heat, bins = _tiny()
    pw = 3.0
    (_, _), g = C.backward(cfg, flat, images, heat, bins, pw, use_bf16=False)
    eps = 1e-6
    fd = np.zeros_like(flat)
    for i in range(flat.size):
        a = flat.copy()
        a[i] += eps
        b = flat.copy()
        b[i] -= eps
        fd[i] = (C.loss(cfg, a, images, heat, bins, pw) - C.loss(cfg, b, images, heat, bins, pw)) / (2 * eps)
    assert np.abs(g - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())
    assert np.abs(g).max() > 1e-3


def test_loss_closed_forms():
    cfg = C.Config(8, [2], [2], 7)
    G = cfg.grid
    logits = np.zeros((1, G, G, 8))
    heat = np.zeros((1, G, G))
    heat[0, 0, 0] = 1
    bins = -np.ones((1, G, G), int)
    bins[0, 0, 0] = 3
    lh, lb = C.loss_terms(logits, heat, bins, 4.0)
    assert lh == pytest.approx((4.0 + (G * G - 1)) / (G * G) * np.log(2.0))
    assert lb == pytest.approx(np.log(7.0))
