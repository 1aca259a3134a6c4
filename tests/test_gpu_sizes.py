"""GPU parity for image sides that are not powers of two: the mixed radix-2/3/5 FFT plans of the
preparation, the column pass and the row pass, against the oracle on the same inputs.

* 384 x 324 (2^7 3 x 2^2 3^4), a 260 x 260 micrograph that the size policy pads to 270 x 270
  (2 3^3 5, an odd half-length row transform), and a 384 x 256 micrograph that it rotates to
  256 x 384 (Rotate2DInPlaceBy90Degrees(true), match_template.cpp:595): the input preparation
  (rotation included) and a search over 10 grid positions to the bars of parity_util
* the non-power-of-two image sizes of the multi-micrograph workload after the reference's size
  policy: 2048 x 3072, 4096 x 2916, 2048 x 1536 and 2592 x 2592,
  and the K3 sizes 4096 x 5832 (a 5760 x 4092 frame after the policy, match_template.cpp:537-538) and
  5832 x 5832 (2^3 3^6: a radix-8 stage and six radix-3 stages per line):
  input preparation, single CCF planes (max |gpu - oracle| / max |oracle| <= 1e-4) and a search over
  one grid position
Each micrograph holds copies of the 64^3 pseudo-atom template (seeded), so the oracle stays cheap.
"""
import os

import numpy as np
import pytest

from cistem_downstream_bah_amd import core, synth
from cistem_downstream_bah_amd import match_template as MT
from oracle import oracle as O
from parity_util import compare_search, rel

pytestmark = pytest.mark.gpu

THREADS = int(os.environ.get("TM_ORACLE_THREADS", str(min(16, os.cpu_count() or 1))))


def _micrograph(W, H, seed, copies):
    return synth.make_synthetic(seed, W, H, 64, 200, 24.0, copies, 0.1, 15000.0, min_dist=64)


def _setup(s, params):
    mt = O.MatchTemplateOracle(params, s.micrograph, s.volume)
    I = mt.info
    g = core.TemplateMatchingCore(0)
    g.init(mt.template_spectrum, mt.input_spectrum, I.T, mt.grid, mt.psi, I.first_pos, I.last_pos,
           I.histogram_min_scaled, I.histogram_step_scaled, 512, I.max_padding, I.defocus_step, I.pixel_size_step,
           batch_size=0)
    return mt, g


def _check_preparation(s, mt, label):
    spec, cx, cy, W, H, rot, mp = MT.prepare_input_image(s.micrograph, s.volume.shape[0])
    assert (W, H) == (mt.info.W, mt.info.H)
    ox, oy = mt.whitening_curve
    assert np.array_equal(cx, ox)
    ec, ei = rel(cy, oy), rel(spec, mt.input_spectrum)
    print(f"{label}: whitening curve rel {ec:.2e} input spectrum rel {ei:.2e}", flush=True)
    assert ec <= 1e-5 and ei <= 1e-5


@pytest.mark.parametrize("mx,my,W,H", [(384, 324, 384, 324), (260, 260, 270, 270), (384, 256, 256, 384)])
def test_small_mixed_radix_search(mx, my, W, H):
    s = _micrograph(mx, my, 7000 + mx, 3)
    mt, g = _setup(s, O.MTParams(angular_step=15.0, in_plane_angular_step=15.0, padding=1.0,
                                 first_search_position=0, last_search_position=9))
    try:
        I = mt.info
        assert (I.W, I.H, I.T) == (W, H, 64) and I.rotated == (mx == 384 and my == 256)
        _check_preparation(s, mt, f"{W}x{H}")
        ref = mt.run(threads=THREADS)
        n = g.run_inner_loop(mt.projection_filter(0), 0, 0)
        got = g.results()
        assert n == ref["nccf"] == 10 * I.n_psi
        compare_search(mt, ref, got, f"{W}x{H}")
    finally:
        g.close()


@pytest.mark.parametrize("W,H", [(2048, 3072), (4096, 2916), (2048, 1536), (2592, 2592), (4096, 5832), (5832, 5832)])
def test_multi_micrograph_sizes(W, H):
    s = _micrograph(W, H, 8000 + H, 12)
    mt, g = _setup(s, O.MTParams(angular_step=10.0, in_plane_angular_step=15.0, padding=1.0,
                                 first_search_position=40, last_search_position=40))
    try:
        I = mt.info
        assert (I.W, I.H) == (W, H)
        _check_preparation(s, mt, f"{W}x{H}")
        pf = mt.projection_filter(0)
        for k in (0, 11):
            _, ref_c = mt.projection_and_ccf(40, k, 0)
            e = rel(g.debug_ccf(pf, 40, k), ref_c)
            print(f"{W}x{H} psi {k}: ccf rel {e:.2e}", flush=True)
            assert e <= 1e-4
        ref = mt.run(threads=THREADS)
        n = g.run_inner_loop(pf, 0, 0)
        got = g.results()
        assert n == ref["nccf"] == I.n_psi
        compare_search(mt, ref, got, f"{W}x{H}", max_tie_orientations=12)
    finally:
        g.close()
