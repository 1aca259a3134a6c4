"""CPU: pins of the oracle's Image::ChangePixelSize (image.cpp:8705-8792) and the pixel-size planes
of the search (match_template.cpp:837-845), by closed forms.

* a band-limited Gaussian blob resamples to a Gaussian with its width scaled by new/pad (the
  ratio of the transform edges, ~1/factor), peak value and centre kept, integral scaled by
  (new/pad)^d: a reversed crop/pad, a wrong scale or a half-pixel shift fails it. Factors 1.0417
  (crop to 960) and 1.02 (crop to 980, radix 7), 0.98 (pad to 1020, radix 17) and 0.96 (pad to 1042,
  the 521-point prime transform), in 2-D; 1.02 in 3-D
* the padding value is the mean inside 0.45 of the edge: a constant image stays constant
* sizes: pad 1000 (1 / tolerance) and new = int(pad / factor), made even like the input
Not pinned by a closed form: the positive-Nyquist planes ClipInto fills in when a spectrum grows
(image.cpp:8626-8690); they carry no energy for these band-limited inputs.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _moments(img, c, R=12):
    n = img.shape[-1]
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    w = img.astype(np.float64) * (((xx - c) ** 2 + (yy - c) ** 2) <= R * R)
    m = w.sum()
    cx, cy = (w * xx).sum() / m, (w * yy).sum() / m
    return cx, cy, np.sqrt((w * ((xx - cx) ** 2 + (yy - cy) ** 2)).sum() / m / 2), m


@pytest.fixture(scope="module", autouse=True)
def threads():
    O.set_threads(8)
    yield
    O.set_threads(1)


@pytest.mark.parametrize("factor,new", [(1.0417, 960), (1.02, 980), (0.98, 1020), (0.96, 1042)])
def test_gaussian_resamples_by_the_edge_ratio(factor, new):
    n = 64
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    img = np.exp(-((xx - 32) ** 2 + (yy - 32) ** 2) / (2 * 3.0 ** 2)).astype(np.float32)
    out, (pad, nw) = O.change_pixel_size(img, factor)
    assert pad[:2] == (1000, 1000) and nw[:2] == (new, new)
    s = new / 1000.0
    cx0, cy0, sig0, m0 = _moments(img, 32)
    cx, cy, sig, m = _moments(out, 32)
    assert abs(out.max() - 1.0) < 1e-4
    assert abs(cx - 32) < 2e-3 and abs(cy - 32) < 2e-3
    assert abs(sig / sig0 - s) < 3e-3 * s
    assert abs(m / m0 - s ** 2) < 3e-3 * s ** 2


@pytest.mark.parametrize("factor", [1.02, 0.98])
def test_constant_stays_constant(factor):
    img = np.full((48, 48), 2.5, np.float32)
    out, _ = O.change_pixel_size(img, factor)
    assert np.abs(out - 2.5).max() < 1e-5


def test_gaussian_3d():
    n = 32
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    vol = np.exp(-((xx - 16) ** 2 + (yy - 16) ** 2 + (zz - 16) ** 2) / (2 * 2.5 ** 2)).astype(np.float32)
    out, (pad, nw) = O.change_pixel_size(vol, 1.02)
    assert pad == (1000, 1000, 1000) and nw == (980, 980, 980)
    s = 0.98
    w0, w1 = vol.astype(np.float64), out.astype(np.float64)
    r2 = (xx - 16) ** 2 + (yy - 16) ** 2 + (zz - 16) ** 2 <= 10 ** 2
    def mom(w):
        w = w * r2
        m = w.sum()
        return m, np.sqrt((w * ((xx - 16) ** 2 + (yy - 16) ** 2 + (zz - 16) ** 2)).sum() / m / 3)
    m0, sg0 = mom(w0)
    m1, sg1 = mom(w1)
    assert abs(out.max() - 1.0) < 1e-4
    assert abs(sg1 / sg0 - s) < 3e-3 * s
    assert abs(m1 / m0 - s ** 3) < 4e-3 * s ** 3
