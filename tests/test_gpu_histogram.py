"""GPU: the search's histogram, bin for bin, equals the reference expression applied to the GPU's own
correlation planes.

match_template.cpp:1124 bins each value v of the logical W x H region as
int(double(v - histogram_min_scaled) / histogram_step_scaled), with the min and step held in double
(:443-444), and counts bins 0..n-1. tm_debug_ccf returns the plane the search's row pass produces
for one orientation (the same kernels, one orientation at a time), so binning those planes on the
host with that expression (numpy's float64 division is the correctly rounded one) must give the
search's histogram exactly: no value may move to a neighbouring bin.

Cases: config 0 (512^2, two grid positions, 74 orientations) and the 4096^2 config-1 shape
(position 3, the first 8 psi angles of the 1.5 deg list) with FP32 and FP16 plane storage.
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import core, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def reference_bins(plane: np.ndarray, hmin: float, hstep: float, nbins: int = 512) -> np.ndarray:
    q = np.trunc((plane.astype(np.float64) - hmin) / hstep)
    q = q[(q >= 0) & (q < nbins)].astype(np.int64)
    return np.bincount(q, minlength=nbins)


def _check(mt, pos_lo, pos_hi, psi, plane_storage="fp32"):
    I = mt.info
    g = core.TemplateMatchingCore(0)
    try:
        g.init(mt.template_spectrum, mt.input_spectrum, I.T, mt.grid, psi, pos_lo, pos_hi, I.histogram_min_scaled,
               I.histogram_step_scaled, 512, I.max_padding, I.defocus_step, I.pixel_size_step, batch_size=16,
               plane_storage=plane_storage)
        pf = mt.projection_filter(0)
        n = g.run_inner_loop(pf, 0, 0)
        got = g.results().hist
        want = np.zeros(512, np.int64)
        for pos in range(pos_lo, pos_hi + 1):
            for k in range(len(psi)):
                want += reference_bins(g.debug_ccf(pf, pos, k), I.histogram_min_scaled, I.histogram_step_scaled)
    finally:
        g.close()
    assert n == (pos_hi - pos_lo + 1) * len(psi)
    diff = np.nonzero(got != want)[0]
    print(f"{I.W}x{I.H} {plane_storage}: {n} CCFs, {int(want.sum())} counts in range, {diff.size} bins differ",
          flush=True)
    assert got.sum() == want.sum() and diff.size == 0, [(int(b), int(got[b]), int(want[b])) for b in diff[:10]]


def test_histogram_bins_config0():
    s = synth.config0(0)
    mt = O.MatchTemplateOracle(O.MTParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                                          first_search_position=100, last_search_position=101),
                               s.micrograph, s.volume)
    _check(mt, 100, 101, mt.psi)


@pytest.fixture(scope="module")
def config1_pole():
    s = synth.config1(1000)
    return O.MatchTemplateOracle(O.MTParams(angular_step=2.5, in_plane_angular_step=1.5, padding=2.0,
                                            first_search_position=3, last_search_position=3), s.micrograph, s.volume)


@pytest.mark.parametrize("plane_storage", ["fp32", "fp16"])
def test_histogram_bins_4096(config1_pole, plane_storage):
    mt = config1_pole
    assert (mt.info.W, mt.info.H) == (4096, 4096)
    _check(mt, 3, 3, np.ascontiguousarray(mt.psi[:8]), plane_storage)
