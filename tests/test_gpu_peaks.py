"""GPU: the device peak search (tm_find_peaks) against the host peak search of the driver
(peaks.find_peaks, the form of match_template.cpp:1839-1980) on the same search.

* config 0 (whole grid) at its expected threshold and at a lower threshold that also takes noise
  peaks (tens of peaks, masks that overlap the border and each other): identical peak lists,
  positions, z values and the angles and defocus at each peak
* capacity cuts the list; a threshold above every pixel gives no peak
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import peaks as PK, synth
from cistem_downstream_bah_amd import match_template as MT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def searched():
    s = synth.config0(0)
    se = MT.MatchTemplateSearch(s.micrograph, s.volume,
                                MT.SearchParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                                                defocus_search_range=100.0, defocus_step=100.0))
    n = se.run()
    return se, n


@pytest.mark.parametrize("thr", [None, 4.5])
def test_device_peaks_equal_host_peaks(searched, thr):
    se, n = searched
    threshold = PK.expected_threshold(se.W, se.H, n) if thr is None else thr
    print(f"case: config0 peaks above {threshold:.3f}", flush=True)
    fin = se.core.finalize(n)
    r = se.core.results()
    host = PK.find_peaks(fin["zscore"], threshold, se.T, 10,
                         maps={"psi": r.psi, "theta": r.theta, "phi": r.phi, "defocus": r.defocus})
    dev = se.core.find_peaks(n, threshold, se.T // 4 + 1, 10.0)
    print(f"{len(dev)} peaks on the device, {len(host)} on the host", flush=True)
    assert len(dev) == len(host) > 0
    for a, b in zip(dev, host):
        assert (a["x"], a["y"]) == (b["x"], b["y"])
        assert a["peak_height"] == pytest.approx(b["peak_height"], rel=0, abs=0)
        for k in ("psi", "theta", "phi", "defocus"):
            assert a[k] == b[k], k


def test_peak_info_fields(searched):
    """tm_peak_info follows TemplateMatchFoundPeakInfo (template_matching.h:4-18): positions in
    Angstroms (pixel x pixel size, match_template.cpp:1883-1884), peak numbers in finding order."""
    se, n = searched
    a = se.core.find_peaks(n, 4.5, se.T // 4 + 1, 10.0, capacity=20)
    b = se.core.find_peaks(n, 4.5, se.T // 4 + 1, 10.0, capacity=20, pixel_size=1.3)
    assert len(a) == len(b) == 20
    for k, (p, q) in enumerate(zip(a, b)):
        assert p["peak_number"] == p["original_peak_number"] == p["new_peak_number"] == k + 1
        assert (p["x_pos"], p["y_pos"]) == (p["x"], p["y"])
        assert q["x_pos"] == np.float32(np.float32(p["x"]) * np.float32(1.3))
        assert q["y_pos"] == np.float32(np.float32(p["y"]) * np.float32(1.3))
        assert (q["x"], q["y"], q["peak_height"], q["psi"]) == (p["x"], p["y"], p["peak_height"], p["psi"])


def test_capacity_and_high_threshold(searched):
    se, n = searched
    assert len(se.core.find_peaks(n, 4.5, se.T // 4 + 1, 10.0, capacity=3)) == 3
    assert se.core.find_peaks(n, 1e9, se.T // 4 + 1, 10.0) == []
