"""CPU: peak extraction and expected threshold (match_template.cpp:1763-1980)."""
import numpy as np
from scipy.special import erfcinv

from cistem_downstream_bah_amd import peaks


def test_expected_threshold_close_to_exact():
    for (w, h, n) in [(512, 512, 15466), (4096, 4096, 2200000)]:
        exact = np.sqrt(2) * erfcinv(2.0 / (w * h * n))
        assert abs(peaks.expected_threshold(w, h, n) - exact) / exact < 2e-3


def test_find_peaks_on_known_map():
    rng = np.random.default_rng(0)
    z = rng.standard_normal((256, 256)).astype(np.float32)
    truth = [(40, 50, 12.0), (200, 100, 9.0), (120, 130, 8.5), (5, 5, 20.0)]  # last one inside the border
    for x, y, v in truth:
        z[y, x] = v
    z[131, 121] = 8.4   # neighbour of (120, 130) within the exclusion radius
    found = peaks.find_peaks(z, 8.0, template_size=64, min_peak_radius=10)
    got = [(p["x"], p["y"]) for p in found]
    assert got == [(40, 50), (200, 100), (120, 130)]   # strongest first, border and neighbour excluded
