"""GPU: the pixel-size loop of match_template (match_template.cpp:837-914) against the oracle.

* each plane's template (tm_prepare_template_scaled: Image::ChangePixelSize by
  (pixel_size + size_i * step) / pixel_size, image.cpp:8705-8792, then DC 0 and the quadrant swap)
  for a 64^3 template, pixel size 1.0, step 0.02: factors 0.98 (spectrum padded 1000 -> 1020) and
  1.02 (cropped 1000 -> 980), both through the device's transforms of any length: within 2e-5 of
  the spectrum's maximum
* a search over the three pixel-size planes (one defocus plane, positions 100-103 of config 0) with
  the template replaced between planes (tm_set_template): MIP, best index (global, plane =
  size-major call order), pixel-size map, sums and histogram to the search bars
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import match_template as MT
from cistem_downstream_bah_amd import synth
from oracle import oracle as O
from parity_util import compare_search, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    O.set_threads(16)
    s = synth.config0(0)
    prm = O.MTParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0, pixel_size_search_range=0.02,
                     pixel_size_step=0.02, first_search_position=100, last_search_position=103)
    print("case: oracle setup with three pixel-size planes (ChangePixelSize of a 64^3 template)", flush=True)
    mt = O.MatchTemplateOracle(prm, s.micrograph, s.volume)
    assert mt.info.n_size_half == 1
    yield s, prm, mt
    O.set_threads(1)


@pytest.mark.parametrize("size_i", [-1, 1])
def test_scaled_template(setup, size_i):
    s, prm, mt = setup
    print(f"case: scaled template size_i {size_i}", flush=True)
    f = float((np.float32(prm.pixel_size) + np.float32(size_i) * np.float32(prm.pixel_size_step)) /
              np.float32(prm.pixel_size))
    got = MT.prepare_template(s.volume, prm.padding, 0, f)
    ref = mt.template_for_size(size_i)
    e = rel(got, ref)
    print(f"size_i {size_i} factor {f:.6f}: template max err / max {e:.2e}", flush=True)
    assert e <= 2e-5


def test_search_over_pixel_size_planes(setup):
    s, prm, mt = setup
    print("case: search over three pixel-size planes", flush=True)
    ref = mt.run(threads=16)
    sp = MT.SearchParams(angular_step=prm.angular_step, in_plane_angular_step=prm.in_plane_angular_step,
                         padding=prm.padding, pixel_size_search_range=prm.pixel_size_search_range,
                         pixel_size_step=prm.pixel_size_step, first_search_position=100, last_search_position=103)
    search = MT.MatchTemplateSearch(s.micrograph, s.volume, sp)
    assert search.n_size_half == 1
    n = search.run()
    got = search.core.results()
    search.core.close()
    assert n == ref["nccf"] == 3 * 4 * mt.info.n_psi
    compare_search(mt, ref, got, "pixel-size planes")
    same = got.best_index == ref["best_index"]
    assert np.array_equal(got.pixel_size[same], ref["pixel_size"][same])
    assert set(np.unique(ref["pixel_size"])) <= {np.float32(-0.02), np.float32(0.0), np.float32(0.02)}
