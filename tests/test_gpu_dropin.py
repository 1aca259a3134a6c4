"""GPU: the reference-signature C++ drop-in (include/template_matching_core.hpp) driven by a caller
shaped like match_template's GPU branch (tests/dropin/reference_caller.cpp, match_template.cpp:858-985:
Init with the reference's 25 arguments, one RunInnerLoop(Image&, float, float, int, long&) per defocus
plane, and the host merge of each plane from d_max_intensity_projection, d_best_*, d_sum3, d_sumSq3 and
histogram) gives the C ABI's multi-plane search: MIP, psi, theta, phi and defocus maps and histogram
identical, sums within FP32 rounding (the caller adds each plane's sums as floats, as the reference's
caller does), CCF counts equal. Config 0, positions 100-103, three defocus planes."""
import os
import subprocess

import numpy as np
import pytest

from cistem_downstream_bah_amd import core, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "dropin", "reference_caller")


def test_reference_shaped_caller_equals_the_c_abi(tmp_path):
    s = synth.config0(0)
    mt = O.MatchTemplateOracle(O.MTParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                                          defocus_search_range=100.0, defocus_step=100.0,
                                          first_search_position=100, last_search_position=103),
                               s.micrograph, s.volume)
    I = mt.info
    d = str(tmp_path)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{I.P} {I.W} {I.H} {I.T} {I.n_pos} {I.first_pos} {I.last_pos} 512 {I.n_def_half} "
                f"0 360 10 {I.defocus_step} {float(np.float32(I.histogram_min_scaled)):.9g} "
                f"{float(np.float32(I.histogram_step_scaled)):.9g}\n")
    mt.template_spectrum.astype(np.complex64).tofile(os.path.join(d, "template.bin"))
    mt.input_spectrum.astype(np.complex64).tofile(os.path.join(d, "input.bin"))
    mt.grid.astype(np.float32).tofile(os.path.join(d, "grid.bin"))
    planes = list(range(-I.n_def_half, I.n_def_half + 1))
    np.concatenate([mt.projection_filter(k).astype(np.complex64).ravel() for k in planes]).tofile(
        os.path.join(d, "filters.bin"))
    print(f"case: reference-shaped caller, config0 positions 100-103, {len(planes)} planes", flush=True)
    r = subprocess.run([CALLER, d], capture_output=True, text=True, timeout=240)
    print(r.stdout + r.stderr, flush=True)
    assert r.returncode == 0, open(os.path.join(d, "meta.txt")).read()
    W, H = I.W, I.H
    raw = open(os.path.join(d, "out.bin"), "rb").read()
    n = (W + 2) * H
    o = 0
    out = {}
    for name in ("mip", "psi", "theta", "phi", "defocus"):
        out[name] = np.frombuffer(raw, np.float32, n, o).reshape(H, W + 2)[:, :W]
        o += 4 * n
    for name in ("sum", "sumsq"):
        out[name] = np.frombuffer(raw, np.float64, n, o).reshape(H, W + 2)[:, :W]
        o += 8 * n
    hist = np.frombuffer(raw, np.int64, 512, o)
    o += 8 * 512
    nccf, position = np.frombuffer(raw, np.int64, 2, o)

    # the same search through the C ABI: the planes accumulate on the device
    g = core.TemplateMatchingCore(0)
    try:
        g.init(mt.template_spectrum, mt.input_spectrum, I.T, mt.grid, mt.psi, I.first_pos, I.last_pos,
               np.float32(I.histogram_min_scaled), np.float32(I.histogram_step_scaled), 512, 0, I.defocus_step,
               I.pixel_size_step)
        total = sum(g.run_inner_loop(mt.projection_filter(k), 0, k) for k in planes)
        ref = g.results()
    finally:
        g.close()
    assert nccf == position == total == 3 * 4 * 37
    for name in ("mip", "psi", "theta", "phi", "defocus"):
        assert np.array_equal(out[name], getattr(ref, name)), name
    assert np.array_equal(hist, ref.hist)
    for name in ("sum", "sumsq"):
        err = np.abs(out[name] - getattr(ref, name)).max() / np.abs(getattr(ref, name)).max()
        print(f"{name}: rel {err:.2e}", flush=True)
        assert err <= 1e-6
