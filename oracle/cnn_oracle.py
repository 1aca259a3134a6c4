"""CPU restatement of the proposal network of include/tm_cnn.h (test infrastructure only: the tests
compare the device network against it; nothing in the product imports it).

The network has no counterpart in the reference. Its definition: per image standardisation (FP64
mean and standard deviation), n 3x3 convolutions (padding 1, stride 1 or 2, bias, ReLU) with
channels-last activations and weights [cout][ky][kx][cin], then a 1x1 head with 1 + n_bins
outputs per cell (location logit, orientation logits). Loss: mean over cells of the location
logit's binary cross-entropy (positives weighted by pos_weight) + mean over cells with a bin of the
orientation softmax cross-entropy.

Arithmetic is float64. With bf16=True every product operand is first rounded to bfloat16 (round
to nearest even) at the points where the device rounds: the im2col'd layer input and the weights
in the forward pass, the pre-activation gradient and the weights (data gradient) or the layer
input (weight gradient) in the backward pass. With bf16=False the restatement is plain float64,
which the finite-difference checks use.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def bf16(x):
    """Rounds float32 values to bfloat16 (nearest even) and returns them as float64."""
    a = np.ascontiguousarray(x, np.float32).copy()
    u = a.view(np.uint32)
    finite = (u & 0x7F800000) != 0x7F800000
    r = u + (0x7FFF + ((u >> 16) & 1))
    u[:] = np.where(finite, r & 0xFFFF0000, u & 0xFFFF0000)
    return a.astype(np.float64)


@dataclass
class Config:
    input_size: int
    channels: list
    strides: list
    n_bins: int

    @property
    def grid(self):
        s = self.input_size
        for st in self.strides:
            s //= st
        return s

    def shapes(self):
        """[(w_shape, b_shape)] per conv layer, then the head."""
        out, cin = [], 1
        for c in self.channels:
            out.append(((c, 3, 3, cin), (c,)))
            cin = c
        out.append(((1 + self.n_bins, cin), (1 + self.n_bins,)))
        return out

    def n_params(self):
        return int(sum(np.prod(w) + np.prod(b) for w, b in self.shapes()))


def unflatten(cfg: Config, flat):
    flat = np.asarray(flat, np.float64)
    out, o = [], 0
    for ws, bs in cfg.shapes():
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        out.append((flat[o:o + nw].reshape(ws), flat[o + nw:o + nw + nb].reshape(bs)))
        o += nw + nb
    assert o == flat.size
    return out


def flatten(parts):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in parts])


def standardize(images):
    x = np.asarray(images, np.float32).astype(np.float64)
    m = x.mean(axis=(1, 2), keepdims=True)
    v = np.maximum((x * x).mean(axis=(1, 2), keepdims=True) - m * m, 1e-30)
    # the device computes mean and 1/std in FP64 and applies them in float32
    return ((x.astype(np.float32) - m.astype(np.float32)) * (1.0 / np.sqrt(v)).astype(np.float32)).astype(np.float64)


def im2col(X, stride):
    """X [B, S, S, C] -> col [B, So, So, 3, 3, C]: col[..., ky, kx, c] = X[oy*stride+ky-1, ox*stride+kx-1, c]."""
    B, S, _, C = X.shape
    So = S // stride
    P = np.zeros((B, S + 2, S + 2, C), X.dtype)
    P[:, 1:S + 1, 1:S + 1] = X
    col = np.zeros((B, So, So, 3, 3, C), X.dtype)
    for ky in range(3):
        for kx in range(3):
            col[:, :, :, ky, kx] = P[:, ky:ky + stride * So:stride, kx:kx + stride * So:stride]
    return col


def col2im(dcol, S, stride):
    """Adjoint of im2col: dcol [B, So, So, 3, 3, C] -> dX [B, S, S, C]."""
    B, So, _, _, _, C = dcol.shape
    dP = np.zeros((B, S + 2, S + 2, C), dcol.dtype)
    for ky in range(3):
        for kx in range(3):
            dP[:, ky:ky + stride * So:stride, kx:kx + stride * So:stride] += dcol[:, :, :, ky, kx]
    return dP[:, 1:S + 1, 1:S + 1]


def _r(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def forward(cfg: Config, flat, images, use_bf16=True):
    """Returns (activations: [input, layer outputs..., logits], im2col inputs per conv layer)."""
    P = unflatten(cfg, flat)
    acts = [standardize(images)[..., None]]
    for (w, b), st in zip(P[:-1], cfg.strides):
        col = im2col(acts[-1], st)
        B, So = col.shape[0], col.shape[1]
        y = _r(col.reshape(B * So * So, -1), use_bf16) @ _r(w.reshape(w.shape[0], -1), use_bf16).T + b
        y = np.maximum(y, 0.0)
        if use_bf16:
            y = y.astype(np.float32).astype(np.float64)
        acts.append(y.reshape(B, So, So, -1))
    wh, bh = P[-1]
    x = acts[-1]
    B, G = x.shape[0], x.shape[1]
    z = _r(x.reshape(B * G * G, -1), use_bf16) @ _r(wh, use_bf16).T + bh
    if use_bf16:
        z = z.astype(np.float32).astype(np.float64)
    acts.append(z.reshape(B, G, G, -1))
    return acts


def _softplus(x):
    return np.maximum(x, 0.0) + np.log1p(np.exp(-np.abs(x)))


def loss_terms(logits, heat_t, bin_t, pos_weight):
    z = logits[..., 0].ravel()
    y = np.asarray(heat_t, np.float64).ravel()
    t = np.asarray(bin_t).ravel()
    lh = pos_weight * y * _softplus(-z) + (1 - y) * _softplus(z)
    o = logits[..., 1:].reshape(z.size, -1)
    mx = o.max(axis=1, keepdims=True)
    lse = (mx + np.log(np.exp(o - mx).sum(axis=1, keepdims=True)))[:, 0]
    sel = t >= 0
    lb = (lse - o[np.arange(z.size), np.where(sel, t, 0)])[sel]
    return float(lh.mean()), float(lb.sum() / max(1, sel.sum()))


def loss(cfg, flat, images, heat_t, bin_t, pos_weight, use_bf16=False):
    lh, lb = loss_terms(forward(cfg, flat, images, use_bf16)[-1], heat_t, bin_t, pos_weight)
    return lh + lb


def backward(cfg: Config, flat, images, heat_t, bin_t, pos_weight, use_bf16=True):
    """Returns ((heat loss, bin loss), flat gradient)."""
    P = unflatten(cfg, flat)
    acts = forward(cfg, flat, images, use_bf16)
    logits = acts[-1]
    B, G = logits.shape[0], logits.shape[1]
    cells = B * G * G
    z = logits[..., 0].ravel()
    y = np.asarray(heat_t, np.float64).ravel()
    t = np.asarray(bin_t).ravel()
    sig = 1.0 / (1.0 + np.exp(-z))
    dz0 = (pos_weight * y * (sig - 1.0) + (1 - y) * sig) / cells
    o = logits[..., 1:].reshape(cells, -1)
    sel = t >= 0
    nsel = max(1, int(sel.sum()))
    p = np.exp(o - o.max(axis=1, keepdims=True))
    p /= p.sum(axis=1, keepdims=True)
    p[np.arange(cells), np.where(sel, t, 0)] -= 1.0
    p[~sel] = 0.0
    dlog = np.concatenate([dz0[:, None], p / nsel], axis=1)
    if use_bf16:
        dlog = dlog.astype(np.float32).astype(np.float64)
    grads = [None] * len(P)
    # head
    wh, bh = P[-1]
    xL = acts[-2].reshape(cells, -1)
    grads[-1] = (_r(dlog, use_bf16).T @ _r(xL, use_bf16), dlog.sum(axis=0))
    d = (_r(dlog, use_bf16) @ _r(wh, use_bf16)).reshape(acts[-2].shape)
    for l in range(len(cfg.channels) - 1, -1, -1):
        w, b = P[l]
        a_out = acts[l + 1]
        dZ = np.where(a_out > 0, d, 0.0)
        if use_bf16:
            dZ = dZ.astype(np.float32).astype(np.float64)
        Bq, So = dZ.shape[0], dZ.shape[1]
        dZ2 = dZ.reshape(Bq * So * So, -1)
        col = im2col(acts[l], cfg.strides[l]).reshape(Bq * So * So, -1)
        grads[l] = ((_r(dZ2, use_bf16).T @ _r(col, use_bf16)).reshape(w.shape), dZ2.sum(axis=0))
        if l > 0:
            dcol = _r(dZ2, use_bf16) @ _r(w.reshape(w.shape[0], -1), use_bf16)
            d = col2im(dcol.reshape(Bq, So, So, 3, 3, -1), acts[l].shape[1], cfg.strides[l])
    return loss_terms(logits, heat_t, bin_t, pos_weight), flatten(grads)


def init_params(cfg: Config, seed: int):
    """He-normal weights, zero biases (a CPU draw; the device's own init uses its own generator)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ws, bs in cfg.shapes():
        fan = int(np.prod(ws[1:]))
        parts.append((rng.normal(0.0, np.sqrt(2.0 / fan), ws), np.zeros(bs)))
    return flatten(parts).astype(np.float32)
