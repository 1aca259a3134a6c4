"""GPU: the proposal network's kernels (include/tm_cnn.h) against the oracle (oracle/cnn_oracle.py).

* bf16 MFMA product C = A . B^T: ragged M, N, K, a 601-column head shape, each block shape
  (128 x 128 on 16x16x32 MFMAs for M, N >= 128, with a 16-wide k remainder and split over K; 64 x 64,
  128 x 32 for N <= 32, 32 x 128 for M <= 32) and split-K cases of 2, 8 and 19 splits, against float64
  on the same bf16-rounded operands, within 1e-5 of sum |a| |b| per element
* five networks: channels 8, 16, 16 (strides 2, 2, 1), 16, 32, 48 (strides 2, 1, 2), 16, 48, 64, 32
  (strides 2, 1, 2, 1), 8, 24 (strides 2, 2) and 16, 160, 32 (strides 2, 1, 2: the 160-channel
  layers of the config-1-scale network, a partial 64-channel tile in the forward and data-gradient
  products and 32-channel weight-gradient tiles), so that the implicit-product forward convolution (Cin a multiple of 16) runs
  with one, two and three k steps per tap and 16-, 32-, 48- and 64-channel wave tiles beside the
  im2col path (Cin 1 and 8); the implicit data gradient (Cout a multiple of 16, the previous ReLU's
  backward in its epilogue) runs at stride 1 and over the four pixel-parity classes at stride 2, into
  8 (a partial tile), 16, 32, 48 and 64 channels; and the implicit weight gradient (Cin and Cout
  multiples of 16) runs with output x input channel tiles of 16 x 16, 32 x 16, 48 x 32, 48 x 16,
  64 x 48 and 32 x 64 at both strides; the head (21 outputs, a ragged 16-channel tile) runs as 1x1
  implicit products (forward, weight gradient over the zero-padded dL/dlogits rows, data gradient
  with the last ReLU's backward), except after the 24-channel layer, where the head and that layer
  take the general paths (im2col, transposes, split products, col2im)
* forward pass: every layer's activations and the logits within 2e-4 (relative L2) of the oracle
  with the device's bf16 rounding points (the remainder is FP32 accumulation order, and the odd
  input that lies within that noise of a bf16 rounding edge and rounds the other way, one bf16 ulp
  of 2^-8 on that element; the 48- and 64-channel layers have the most such elements), and each
  layer within 1e-6 of the oracle's layer applied to the device's own input; the location
  logits and the logits of chosen cells read back equal the full logit volume
* backward pass: the loss terms within 1e-4 relative and every parameter tensor's gradient within
  1e-4 (relative L2) of the oracle; the oracle's own gradient is pinned by finite differences on the CPU
* training: 40 Adam steps on a fixed batch lower the loss by more than half; parameters round-trip
  through set/get; the seeded initialisation is deterministic; set_params right after an enqueued
  Adam step takes effect with fresh moments
* the device-input entry points (tm_cnn_forward_device, tm_cnn_backward_device, tm_cnn_loss) give
  bit-identical location logits, loss terms and (after one Adam step each) parameters to the host
  entry points on the same images and targets; a bin past the last counts as none there
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import cnn
from oracle import cnn_oracle as C

pytestmark = pytest.mark.gpu


def _bf(x):
    return C.bf16(x)


@pytest.mark.parametrize("M,N,K", [(37, 50, 23), (130, 70, 64), (200, 601, 128), (5, 3, 5000), (64, 64, 16),
                                   (1000, 32, 300), (16, 200, 64), (20, 300, 4100), (32, 96, 40000),
                                   (300, 200, 4100), (1000, 130, 48), (129, 601, 128), (601, 128, 20000)])
def test_gemm(M, N, K):
    print(f"case: gemm {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M * N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((N, K)).astype(np.float32)
    got = cnn.debug_gemm(A, B)
    a, b = _bf(A), _bf(B)
    ref = a @ b.T
    bound = np.abs(a) @ np.abs(b).T
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + 1e-30)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


CASES = {"narrow": ([8, 16, 16], [2, 2, 1]), "wide": ([16, 32, 48], [2, 1, 2]), "deep": ([16, 48, 64, 32], [2, 1, 2, 1]),
         "odd": ([8, 24], [2, 2]), "c160": ([16, 160, 32], [2, 1, 2])}


def _case(kind="narrow"):
    ch, st = CASES[kind]
    spec = cnn.NetSpec(input_size=64, channels=list(ch), strides=list(st), n_bins=20)
    cfg = C.Config(64, spec.channels, spec.strides, spec.n_bins)
    rng = np.random.default_rng(5)
    flat = C.init_params(cfg, 7)
    flat[:] += 0.02 * rng.standard_normal(flat.size).astype(np.float32)
    images = rng.standard_normal((3, 64, 64)).astype(np.float32)
    images[:, 20:30, 30:40] += 2.0
    G = cfg.grid
    heat = np.zeros((3, G, G), np.float32)
    bins = -np.ones((3, G, G), np.int32)
    for b in range(3):
        for (y, x, t) in [(3, 5, 2 + b), (10, 12, 19 - b), (14, 1, 7)]:
            heat[b, y, x] = 1
            bins[b, y, x] = t
    return spec, cfg, flat, images, heat, bins


@pytest.mark.parametrize("kind", list(CASES))
def test_forward_matches_oracle(kind):
    print(f"case: cnn forward 64^2, layers {CASES[kind]}, 20 bins, batch 3", flush=True)
    spec, cfg, flat, images, *_ = _case(kind)
    net = cnn.ProposalNet(spec, max_batch=4, seed=None)
    try:
        net.set_params(flat)
        logits = net.forward(images)
        acts = C.forward(cfg, flat, images, use_bf16=True)
        P = C.unflatten(cfg, flat)
        prev = acts[0]
        for l in range(len(spec.channels)):
            dev = net.activation(l, 3)
            e = _rel(dev, acts[l + 1])
            # the layer alone, on the device's own input: no rounding flips carried in from before
            w, b = P[l]
            col = C.im2col(prev, spec.strides[l])
            Bn, So = col.shape[0], col.shape[1]
            one = np.maximum(C.bf16(col.reshape(Bn * So * So, -1)) @ C.bf16(w.reshape(w.shape[0], -1)).T + b, 0.0)
            e1 = _rel(dev, one.reshape(dev.shape))
            print(f"layer {l}: rel {e:.2e} end to end, {e1:.2e} on the device's own input", flush=True)
            assert e <= 2e-4 and e1 <= 1e-6
            prev = dev.astype(np.float64)
        e = _rel(logits, acts[-1])
        print(f"logits: rel {e:.2e}", flush=True)
        assert e <= 2e-4
        # the proposal step's reads: the location logits and the logits of chosen cells
        assert np.array_equal(net.heat(3), logits[..., 0])
        G = net.grid
        cells = [0, 5, G * G - 1, G * G + 17, 3 * G * G - 1]
        assert np.array_equal(net.cell_logits(cells), logits.reshape(-1, logits.shape[-1])[cells])
    finally:
        net.close()


@pytest.mark.parametrize("kind", list(CASES))
def test_backward_matches_oracle(kind):
    print(f"case: cnn backward 64^2, layers {CASES[kind]}, 20 bins, batch 3", flush=True)
    spec, cfg, flat, images, heat, bins = _case(kind)
    net = cnn.ProposalNet(spec, max_batch=4, seed=None)
    try:
        net.set_params(flat)
        net.forward(images, want_logits=False)
        (lh, lb), g = net.backward(heat, bins, pos_weight=5.0, want_grads=True)
        (rh, rb), rg = C.backward(cfg, flat, images, heat, bins, 5.0, use_bf16=True)
        assert lh == pytest.approx(rh, rel=1e-4) and lb == pytest.approx(rb, rel=1e-4)
        o = 0
        worst = 0.0
        for ws, bs in cfg.shapes():
            for shp in (ws, bs):
                n = int(np.prod(shp))
                e = _rel(g[o:o + n], rg[o:o + n])
                worst = max(worst, e)
                assert e <= 1e-4, (shp, e)
                o += n
        print(f"gradients: worst tensor rel {worst:.2e}, overall {_rel(g, rg):.2e}", flush=True)
        assert _rel(g, rg) <= 1e-5
    finally:
        net.close()


def test_training_lowers_the_loss_and_params_round_trip():
    print("case: cnn training 40 Adam steps", flush=True)
    spec, cfg, flat, images, heat, bins = _case()
    net = cnn.ProposalNet(spec, max_batch=4, seed=11)
    other = cnn.ProposalNet(spec, max_batch=4, seed=11)
    try:
        assert np.array_equal(net.params(), other.params())
        losses = []
        for _ in range(40):
            net.forward(images, want_logits=False)
            (lh, lb), _ = net.backward(heat, bins, pos_weight=5.0)
            losses.append(lh + lb)
            net.adam_step(lr=3e-3)
        print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
        assert losses[-1] < 0.5 * losses[0]
        p = net.params()
        other.set_params(p)
        assert np.array_equal(other.params(), p)
        assert np.array_equal(other.forward(images), net.forward(images))
    finally:
        net.close()
        other.close()


def _hip():
    import ctypes
    import os
    H = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [ctypes.c_void_p]
    return H


def _upload(H, a):
    import ctypes
    p = ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert H.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
    return p.value


def test_device_entry_points_match_host_path():
    print("case: cnn device-input forward/backward against the host entry points, wide network, batch 3", flush=True)
    spec, cfg, flat, images, heat, bins = _case("wide")
    H = _hip()
    host = cnn.ProposalNet(spec, max_batch=4, seed=None)
    dev = cnn.ProposalNet(spec, max_batch=4, seed=None)
    bufs = []
    try:
        host.set_params(flat)
        dev.set_params(flat)
        logits = host.forward(images)
        (lh, lb), _ = host.backward(heat, bins, pos_weight=5.0)
        host.adam_step(lr=3e-3)
        bufs = [_upload(H, np.ascontiguousarray(a)) for a in (images.astype(np.float32), heat, bins)]
        dev.forward_device(bufs[0], 3)
        assert np.array_equal(dev.heat(3), logits[..., 0])
        dev.backward_device(bufs[1], bufs[2], int((bins >= 0).sum()), pos_weight=5.0)
        assert dev.loss() == (lh, lb)
        dev.adam_step(lr=3e-3)
        dev.synchronize()
        assert np.array_equal(dev.params(), host.params())
        # a bin past the last one carries no orientation term on the device path
        wild = bins.copy()
        wild[0, 3, 5] = spec.n_bins + 7
        tame = bins.copy()
        tame[0, 3, 5] = -1
        bufs.append(_upload(H, wild))
        dev.set_params(flat)
        host.set_params(flat)
        dev.forward_device(bufs[0], 3)
        dev.backward_device(bufs[1], bufs[3], int((tame >= 0).sum()), pos_weight=5.0)
        host.forward(images, want_logits=False)
        (th, tb), _ = host.backward(heat, tame, pos_weight=5.0)
        assert dev.loss() == (th, tb)
    finally:
        for b in bufs:
            H.hipFree(b)
        host.close()
        dev.close()


def test_set_params_after_an_enqueued_adam_step():
    """tm_cnn_adam_step only enqueues its update: tm_cnn_set_params right after it leaves exactly the
    given parameters with fresh moments (the next step equals a new handle's first step bit for bit)."""
    print("case: cnn set_params right after an enqueued Adam step", flush=True)
    spec, cfg, flat, images, heat, bins = _case()
    net = cnn.ProposalNet(spec, max_batch=4, seed=3)
    other = cnn.ProposalNet(spec, max_batch=4, seed=None)
    try:
        net.forward(images, want_logits=False)
        net.backward(heat, bins, pos_weight=5.0)
        net.adam_step(lr=1e-2)
        net.set_params(flat)
        assert np.array_equal(net.params(), flat)
        other.set_params(flat)
        for n in (net, other):
            n.forward(images, want_logits=False)
            n.backward(heat, bins, pos_weight=5.0)
            n.adam_step(lr=1e-2)
        assert np.array_equal(net.params(), other.params())
    finally:
        net.close()
        other.close()
