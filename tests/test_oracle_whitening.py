"""CPU: the oracle's whitening curves pinned by closed forms and brute-force enumeration.

match_template whitens the input (match_template.cpp:703-724) and refine_template its micrograph
(refine_template.cpp:526-546) with Image::Compute1DPowerSpectrumCurve (image.cpp:7135-7205): every
Fourier component of the half-plane except the explicit Hermitian mates adds its intensity |F|^2 to
the curve by linear interpolation at its spatial frequency, a second curve counts the components the
same way, power / counts is the mean intensity per point, and the filter is its SquareRoot,
Reciprocal (0 stays 0), scaled to a maximum of 1, on int((W/2+1) sqrt 2 + 1) points over [0, sqrt 2/2].

* plane waves on a floor: a few cosines of chosen amplitudes at chosen frequencies plus one spike
  (whose flat spectrum keeps every point of the curve well above FP32 rounding, so the Reciprocal is
  well conditioned). The spectrum is numpy's FFT of the same image scaled by 1/N (ForwardFFT,
  image.cpp:5085-5102), checked against its closed form (|F| = A/2 at each wave's component over the
  spike's flat h/N); the expected curve comes from enumerating every half-plane component of the
  image by hand and splitting its intensity and its count between the two neighbouring points.
  Taking amplitudes instead of intensities, nearest-point binning, counting the Hermitian mates or an
  off-by-one point count each change the curve far beyond the tolerance (checked below)
* coloured noise: white noise multiplied in Fourier space by a known radial gain g(r) = exp(-r/0.25)
  gives a filter proportional to 1/g (log slope +4 per unit frequency, not the +2 of an amplitude
  curve), flat after dividing by 1/g
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

WAVES = [  # (u, v, amplitude, phase): u along x (0..W/2), v along y (signed)
    (3, 2, 1.0, 0.3),
    (5, -7, 3.0, 1.1),
    (0, 4, 2.0, -0.4),
    (9, 1, 5.0, 2.0),
    (12, -10, 4.0, 0.7),
]


SPIKE = 12.0


def _plane_wave_image(W, H):
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    img = np.full((H, W), 0.5)
    for u, v, a, ph in WAVES:
        img += a * np.cos(2 * np.pi * (u * x / W + v * y / H) + ph)
    # the spike sits where the waves are lowest, far inside the 5-sigma outlier bounds
    y0, x0 = np.unravel_index(np.argmin(img), img.shape)
    img[y0, x0] += SPIKE
    assert abs(img[y0, x0] - img.mean()) < 2.0 * img.std()
    return img.astype(np.float32), (x0, y0)


def _intensities(img, intensity=True):
    """|F|^2 (or |F|) of the half-plane, F = FFT / N with the DC term zeroed (ZeroCentralPixel)."""
    H, W = img.shape
    F = np.fft.rfft2(img.astype(np.float64)) / (W * H)
    F[0, 0] = 0
    return np.abs(F) ** 2 if intensity else np.abs(F)


def _expected_curve(W, H, spec, nearest=False, count_mates=False, extra_point=0):
    """Brute force over the half-plane components (i = 0..W/2, every j) of a W x H image; spec[j, i]
    is the value each component adds."""
    n = int((W / 2.0 + 1.0) * math.sqrt(2.0) + 1.0) + extra_point
    x = np.array([c * (0.5 * math.sqrt(2.0)) / (n - 1) for c in range(n)])
    step = x[1] - x[0]
    power = np.zeros(n)
    counts = np.zeros(n)
    for j in range(H):
        ly = j - H if j >= H // 2 else j
        for i in range(W // 2 + 1):
            if i == 0 and j >= H // 2 and not count_mates:
                continue  # explicit Hermitian mate (image.cpp:5063-5068)
            r = math.sqrt((i / W) ** 2 + (ly / H) ** 2)
            val = spec[j, i]
            k = int(r / step)
            if nearest or k == n - 1:
                k = min(int(r / step + 0.5), n - 1) if nearest else k
                power[k] += val
                counts[k] += 1.0
            else:
                d = r / step - k
                power[k] += val * (1 - d)
                power[k + 1] += val * d
                counts[k] += 1 - d
                counts[k + 1] += d
    mean = np.where(counts > 0, power / np.maximum(counts, 1e-300), 0.0)
    f = np.where(mean > 0, 1.0 / np.sqrt(np.maximum(mean, 1e-300)), 0.0)
    return x, f / f.max()


@pytest.fixture(scope="module")
def plane_waves():
    W, H = 32, 24
    img, (x0, y0) = _plane_wave_image(W, H)
    vol = np.zeros((8, 8, 8), np.float32)
    vol[4, 4, 4] = 1
    # numpy's spectrum against the closed form: the spike's h/N e^{-2 pi i k.r0} everywhere, plus A/2 e^{i phi}
    # at each wave's half-plane component (a wave with u = 0 also at (0, -v), a Hermitian mate)
    F = np.fft.rfft2(img.astype(np.float64)) / (W * H)
    jj, ii = np.mgrid[0:H, 0:W // 2 + 1]
    closed = SPIKE / (W * H) * np.exp(-2j * np.pi * (ii * x0 / W + jj * y0 / H))
    for u, v, a, ph in WAVES:
        closed[v % H, u] += a / 2 * np.exp(1j * ph)
        if u == 0:
            closed[(-v) % H, 0] += a / 2 * np.exp(-1j * ph)
    closed[0, 0] = F[0, 0]
    assert np.abs(F - closed).max() < 1e-6
    return W, H, img, vol


def _close(a, b, tol=1e-4):
    return a.shape == b.shape and np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("which", ["match_template", "refine_template"])
def test_whitening_of_plane_waves(plane_waves, which):
    W, H, img, vol = plane_waves
    if which == "match_template":
        o = O.MatchTemplateOracle(O.MTParams(), img, vol)
        assert (o.info.W, o.info.H, o.info.rotated) == (W, H, False)
    else:
        o = O.RefineTemplateOracle(O.RTParams(), img, vol)
    x, y = (np.asarray(a, np.float64) for a in o.whitening_curve)
    xe, ye = _expected_curve(W, H, _intensities(img))
    assert _close(x, xe, 1e-6)
    err = np.abs(y - ye).max()
    print(f"{which}: max |filter - brute force| {err:.2e}")
    assert _close(y, ye), np.c_[y, ye]
    # each plausible mistake moves the curve far outside that tolerance
    mistakes = {"amplitudes": _expected_curve(W, H, _intensities(img, intensity=False))[1],
                "nearest point": _expected_curve(W, H, _intensities(img), nearest=True)[1],
                "Hermitian mates counted": _expected_curve(W, H, _intensities(img), count_mates=True)[1]}
    for name, bad in mistakes.items():
        assert np.abs(bad - ye).max() > 100 * err + 1e-3, name
    assert _expected_curve(W, H, _intensities(img), extra_point=1)[0].size != x.size


def _coloured_noise(W, H, seed=4, length=0.25):
    rng = np.random.default_rng(seed)
    f = np.fft.rfft2(rng.standard_normal((H, W)))
    fy = np.fft.fftfreq(H)[:, None]
    fx = np.arange(W // 2 + 1)[None, :] / W
    r = np.sqrt(fx * fx + fy * fy)
    return np.fft.irfft2(f * np.exp(-r / length), s=(H, W)).astype(np.float32)


@pytest.mark.parametrize("which", ["match_template", "refine_template"])
def test_whitening_inverts_a_radial_gain(which):
    W, H = 512, 384
    img = _coloured_noise(W, H)
    vol = np.zeros((16, 16, 16), np.float32)
    vol[8, 8, 8] = 1
    o = (O.MatchTemplateOracle(O.MTParams(), img, vol) if which == "match_template"
         else O.RefineTemplateOracle(O.RTParams(), img, vol))
    x, y = (np.asarray(a, np.float64) for a in o.whitening_curve)
    band = (x > 0.05) & (x < 0.6)
    # filter proportional to 1/g = exp(r / 0.25): log-linear in r with slope 4 (an amplitude curve
    # would give 1/sqrt(g), slope 2), flat once divided by it
    slope, _ = np.polyfit(x[band], np.log(y[band]), 1)
    flat = y[band] * np.exp(-x[band] / 0.25)
    print(f"{which}: log slope {slope:.3f}, flatness {np.std(flat) / np.mean(flat):.3f}")
    assert abs(slope - 4.0) < 0.2
    assert np.std(flat) / np.mean(flat) < 0.08
