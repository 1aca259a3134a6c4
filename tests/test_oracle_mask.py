"""CPU: pins of the oracle's CosineMask (image.cpp:3934-4209; the centred real-space branch that
refine_template.cpp:716 uses for the particle window) and of the masked window of its refine_template
restatement.

Pins, chosen so that a dropped term, a wrong sign or a strict/inclusive slip fails one of them:
* a constant image is unchanged (the band mean is the constant)
* closed-form points on the x axis (radius 10, edge 4: band 8..12): d = 8 keeps the pixel (weight 1),
  d = 10 gives the average of pixel and band mean (weight 1/2), d = 12 and beyond give the band mean
* the band mean is a brute-force average over the pixels whose squared distance lies in
  [8^2, 12^2], both ends included (a planted value on d = 8 and on d = 12 moves it)
* pixels inside the band's inner radius are untouched; affine maps commute with the mask
* the mask volume of a constant image approaches the area integral pi r0^2 + int e(s)^2 2 pi (r0 + s) ds
* radius below half the edge: the inner radius clamps at 0
* refine window: the masked window is CosineMask(radius / pixel, 20 / pixel) of the unmasked clip,
  with the radius clamped to (t/2 - 1) pixels as refine_template.cpp:520-521 does
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O


def _img(n=48, seed=0):
    return np.random.default_rng(seed).standard_normal((n, n)).astype(np.float32)


def _band_mean(a, r0, r1):
    n = a.shape[0]
    c = n // 2
    jj, ii = np.mgrid[0:n, 0:n]
    d2 = ((ii - c) ** 2 + (jj - c) ** 2).astype(np.float64)
    sel = (d2 >= r0 * r0) & (d2 <= r1 * r1)
    return a[sel].astype(np.float64).mean(), d2


def test_constant_image_unchanged():
    a = np.full((40, 40), 2.5, np.float32)
    m, _ = O.cosine_mask(a, 11.0, 5.0)
    assert np.array_equal(m, a)


def test_closed_form_points_and_band_mean():
    a = _img()
    n = a.shape[0]
    c = n // 2
    a[c, c + 8] = 40.0     # inner end of the band (d^2 = r0^2, included)
    a[c + 12, c] = -30.0   # outer end (d^2 = (r0 + edge)^2, included)
    m, _ = O.cosine_mask(a, 10.0, 4.0)
    mean, d2 = _band_mean(a, 8.0, 12.0)
    # the planted end values are part of the mean: without either one it differs clearly
    n_band = int(((d2 >= 64) & (d2 <= 144)).sum())
    assert abs(40.0 / n_band) > 1e-3 and abs(30.0 / n_band) > 1e-3
    assert m[c, c + 8] == pytest.approx(a[c, c + 8], abs=1e-5)
    assert m[c, c + 10] == pytest.approx(0.5 * (a[c, c + 10] + mean), abs=1e-5)
    assert m[c, c + 12] == pytest.approx(mean, abs=1e-6)
    outside = d2 > 144
    assert np.allclose(m[outside], np.float32(mean), rtol=0, atol=0)
    inside = d2 < 64
    assert np.array_equal(m[inside], a[inside])
    # band: between pixel and mean, weight falling with distance
    band = (d2 >= 64) & (d2 <= 144)
    e = (m[band].astype(np.float64) - mean) / (a[band].astype(np.float64) - mean)
    dist = np.sqrt(d2[band])
    ok = np.abs(a[band] - mean) > 1e-2
    assert np.all((e[ok] > -1e-5) & (e[ok] < 1 + 1e-5))
    order = np.argsort(dist[ok])
    assert np.all(np.diff(e[ok][order]) < 1e-4)


def test_affine_equivariance():
    a = _img(seed=2)
    m1, _ = O.cosine_mask(a, 14.0, 6.0)
    m2, _ = O.cosine_mask(3.0 * a - 1.5, 14.0, 6.0)
    assert np.allclose(m2, 3.0 * m1 - 1.5, atol=2e-5)


def test_mask_volume_area_integral():
    r, edge = 40.0, 12.0
    _, vol = O.cosine_mask(np.ones((128, 128), np.float32), r, edge)
    r0 = r - edge / 2
    band, _ = integrate.quad(lambda s: ((1 + np.cos(np.pi * s / edge)) / 2) ** 2 * 2 * np.pi * (r0 + s), 0, edge)
    assert vol == pytest.approx(np.pi * r0 * r0 + band, rel=0.01)


def test_inner_radius_clamps_at_zero():
    a = _img(seed=3)
    n = a.shape[0]
    c = n // 2
    m, _ = O.cosine_mask(a, 2.0, 10.0)   # r0 = 2 - 5 < 0 -> 0; band 0..10
    mean, d2 = _band_mean(a, 0.0, 10.0)
    assert m[c, c] == pytest.approx(a[c, c], abs=1e-6)   # d = 0: weight 1
    assert np.allclose(m[d2 > 100], np.float32(mean), rtol=0, atol=0)


def _window_real(rt):
    t = rt.t
    w = rt.window_spectrum.astype(np.complex128)
    return np.roll(np.fft.irfft2(w, s=(t, t)) * t * t, (t // 2, t // 2), axis=(0, 1))


@pytest.mark.parametrize("mask_radius,pixel_size", [(12.0, 1.0), (30.0, 1.5), (500.0, 1.0)])
def test_refine_window_is_masked_clip(mask_radius, pixel_size):
    rng = np.random.default_rng(11)
    img = rng.standard_normal((192, 256)).astype(np.float32)
    t = 32
    vol = np.zeros((t, t, t), np.float32)
    vol[t // 2, t // 2, t // 2] = 1
    plain = O.RefineTemplateOracle(O.RTParams(pixel_size=pixel_size), img, vol)
    masked = O.RefineTemplateOracle(O.RTParams(pixel_size=pixel_size, mask_radius=mask_radius), img, vol)
    plain.window(10, -7)
    masked.window(10, -7)
    clip = _window_real(plain).astype(np.float32)
    radius = min(mask_radius, (t / 2 - 1) * pixel_size)
    want, _ = O.cosine_mask(clip, np.float32(radius) / np.float32(pixel_size),
                            np.float32(20.0) / np.float32(pixel_size))
    got = _window_real(masked)
    assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max()
    assert np.abs(got - clip).max() > 0.1 * np.abs(clip).max()
