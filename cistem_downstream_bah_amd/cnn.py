"""ctypes binding of the proposal network (include/tm_cnn.h): a small convolutional network whose
forward pass, backward pass and Adam step run as hand-written HIP kernels (bf16 MFMA products,
FP32 accumulation). It has no counterpart in the reference.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import core

F32P = core.F32P
I32P = core.I32P
MAX_LAYERS = 16
CNN_EXPORTS = ["tm_cnn_create", "tm_cnn_destroy", "tm_cnn_shape", "tm_cnn_init_params", "tm_cnn_set_params",
               "tm_cnn_get_params", "tm_cnn_forward", "tm_cnn_backward", "tm_cnn_adam_step", "tm_cnn_get_activation",
               "tm_cnn_debug_gemm", "tm_cnn_last_error", "tm_cnn_heat", "tm_cnn_cell_logits", "tm_cnn_forward_device",
               "tm_cnn_backward_device", "tm_cnn_loss", "tm_cnn_synchronize"]


class CnnConfig(ctypes.Structure):
    """tm_cnn_config (include/tm_cnn.h)."""
    _fields_ = [("input_size", ctypes.c_int), ("n_layers", ctypes.c_int), ("channels", ctypes.c_int * MAX_LAYERS),
                ("strides", ctypes.c_int * MAX_LAYERS), ("n_bins", ctypes.c_int), ("max_batch", ctypes.c_int)]


def _lib():
    L = core.lib()
    if not getattr(L, "_tm_cnn_bound", False):
        vp = ctypes.c_void_p
        L.tm_cnn_last_error.restype = ctypes.c_char_p
        L.tm_cnn_create.argtypes = [ctypes.c_int, ctypes.POINTER(CnnConfig), ctypes.POINTER(vp)]
        L.tm_cnn_destroy.argtypes = [vp]
        L.tm_cnn_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)]
        L.tm_cnn_init_params.argtypes = [vp, ctypes.c_ulonglong]
        L.tm_cnn_set_params.argtypes = [vp, F32P]
        L.tm_cnn_get_params.argtypes = [vp, F32P]
        L.tm_cnn_forward.argtypes = [vp, F32P, ctypes.c_int, F32P]
        L.tm_cnn_backward.argtypes = [vp, F32P, I32P, ctypes.c_float, F32P, F32P]
        L.tm_cnn_adam_step.argtypes = [vp] + [ctypes.c_float] * 4
        L.tm_cnn_forward_device.argtypes = [vp, vp, ctypes.c_int]
        L.tm_cnn_backward_device.argtypes = [vp, vp, vp, ctypes.c_longlong, ctypes.c_float]
        L.tm_cnn_loss.argtypes = [vp, F32P]
        L.tm_cnn_synchronize.argtypes = [vp]
        L.tm_cnn_heat.argtypes = [vp, F32P]
        L.tm_cnn_cell_logits.argtypes = [vp, I32P, ctypes.c_int, F32P]
        L.tm_cnn_get_activation.argtypes = [vp, ctypes.c_int, F32P]
        L.tm_cnn_debug_gemm.argtypes = [ctypes.c_int, F32P, F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int, F32P]
        L._tm_cnn_bound = True
    return L


def _check(rc):
    if rc != 0:
        raise core.TMError(rc, _lib().tm_cnn_last_error().decode())


def _f(a):
    return None if a is None else a.ctypes.data_as(F32P)


def debug_gemm(A: np.ndarray, B: np.ndarray, device: int = 0) -> np.ndarray:
    """C = A . B^T on bf16-rounded operands with FP32 accumulation (tm_cnn_debug_gemm)."""
    a = np.ascontiguousarray(A, np.float32)
    b = np.ascontiguousarray(B, np.float32)
    M, K = a.shape
    N = b.shape[0]
    c = np.zeros((M, N), np.float32)
    _check(_lib().tm_cnn_debug_gemm(device, _f(a), _f(b), M, N, K, _f(c)))
    return c


@dataclass
class NetSpec:
    """Layer plan: 3x3 convolutions (channels, strides), n_bins orientation bins, input edge."""
    input_size: int = 512
    channels: list = field(default_factory=lambda: [32, 32, 64, 64, 128, 128, 128])
    strides: list = field(default_factory=lambda: [2, 1, 2, 1, 2, 1, 1])
    n_bins: int = 600


class ProposalNet:
    """One network on one GPU (tm_cnn)."""

    def __init__(self, spec: NetSpec, max_batch: int = 16, device: int = 0, seed: int | None = 0):
        self.spec = spec
        n = len(spec.channels)
        if n != len(spec.strides) or not 1 <= n <= MAX_LAYERS:
            raise ValueError("channels and strides must have the same length, 1..16")
        cfg = CnnConfig(spec.input_size, n, (ctypes.c_int * MAX_LAYERS)(*spec.channels),
                        (ctypes.c_int * MAX_LAYERS)(*spec.strides), spec.n_bins, max_batch)
        self.max_batch = max_batch
        self.h = ctypes.c_void_p()
        _check(_lib().tm_cnn_create(device, ctypes.byref(cfg), ctypes.byref(self.h)))
        npar = ctypes.c_longlong(0)
        g = ctypes.c_int(0)
        _check(_lib().tm_cnn_shape(self.h, ctypes.byref(npar), ctypes.byref(g)))
        self.n_params, self.grid = npar.value, g.value
        self.cells = self.grid * self.grid
        if seed is not None:
            _check(_lib().tm_cnn_init_params(self.h, seed))

    def close(self):
        if getattr(self, "h", None):
            _lib().tm_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        _check(_lib().tm_cnn_get_params(self.h, _f(out)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, np.float32)
        if a.size != self.n_params:
            raise ValueError(f"{a.size} parameters given, {self.n_params} expected")
        _check(_lib().tm_cnn_set_params(self.h, _f(a)))

    def forward(self, images: np.ndarray, want_logits: bool = True):
        x = np.ascontiguousarray(images, np.float32)
        if x.ndim == 2:
            x = x[None]
        B = x.shape[0]
        out = np.zeros((B, self.grid, self.grid, 1 + self.spec.n_bins), np.float32) if want_logits else None
        _check(_lib().tm_cnn_forward(self.h, _f(x), B, _f(out)))
        return out

    def forward_device(self, images_ptr: int, batch: int) -> None:
        """Forward pass over batch images already in device memory (a device address of batch x S x S
        FP32 values), enqueued without a copy or a wait (tm_cnn_forward_device)."""
        _check(_lib().tm_cnn_forward_device(self.h, ctypes.c_void_p(images_ptr), batch))

    def heat(self, batch: int = 1) -> np.ndarray:
        """Location logits of the last forward pass, (batch, grid, grid)."""
        out = np.zeros((batch, self.grid, self.grid), np.float32)
        _check(_lib().tm_cnn_heat(self.h, _f(out)))
        return out

    def cell_logits(self, cells) -> np.ndarray:
        """Logits (1 + n_bins) of the given flat cells (b grid^2 + y grid + x)."""
        c = np.ascontiguousarray(cells, np.int32)
        out = np.zeros((c.size, 1 + self.spec.n_bins), np.float32)
        _check(_lib().tm_cnn_cell_logits(self.h, c.ctypes.data_as(I32P), c.size, _f(out)))
        return out

    def backward(self, heat_target: np.ndarray, bin_target: np.ndarray, pos_weight: float = 1.0,
                 want_grads: bool = False):
        ht = np.ascontiguousarray(heat_target, np.float32)
        bt = np.ascontiguousarray(bin_target, np.int32)
        loss = np.zeros(2, np.float32)
        g = np.zeros(self.n_params, np.float32) if want_grads else None
        _check(_lib().tm_cnn_backward(self.h, _f(ht), bt.ctypes.data_as(I32P), pos_weight, _f(loss), _f(g)))
        return (float(loss[0]), float(loss[1])), g

    def backward_device(self, heat_ptr: int, bins_ptr: int, n_labelled: int, pos_weight: float = 1.0) -> None:
        """Backward pass from targets in device memory (FP32 heat and int32 bins, batch x grid x grid),
        enqueued without a wait; n_labelled = the number of cells with a bin (tm_cnn_backward_device)."""
        _check(_lib().tm_cnn_backward_device(self.h, ctypes.c_void_p(heat_ptr), ctypes.c_void_p(bins_ptr), n_labelled,
                                             pos_weight))

    def loss(self):
        """Both loss terms of the last backward pass (waits for it)."""
        out = np.zeros(2, np.float32)
        _check(_lib().tm_cnn_loss(self.h, _f(out)))
        return float(out[0]), float(out[1])

    def synchronize(self) -> None:
        _check(_lib().tm_cnn_synchronize(self.h))

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        _check(_lib().tm_cnn_adam_step(self.h, lr, beta1, beta2, eps))

    def activation(self, layer: int, batch: int) -> np.ndarray:
        s = self.spec.input_size
        for st in self.spec.strides[: layer + 1]:
            s //= st
        c = self.spec.channels[layer] if layer < len(self.spec.channels) else 1 + self.spec.n_bins
        out = np.zeros((batch, s, s, c), np.float32)
        _check(_lib().tm_cnn_get_activation(self.h, layer, _f(out)))
        return out
