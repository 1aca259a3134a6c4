"""Peak extraction from the z-score map and the expected-threshold formula.

Follows the master side of the reference's match_template (match_template.cpp:1763-1980):
  * expected threshold = sqrt(2) * erfcinv(2 / (W * H * N_ccf)) with cisTEM's erfcinv (Winitzki's
    approximation, a = 0.147; src/core/functions.cpp:1282-1299) and the original image size, computed
    by the library's tm_expected_threshold;
  * peaks: repeatedly take the integer maximum of the scaled MIP away from a border of T/4 + 1
    pixels (Image::FindPeakWithIntegerCoordinates, image.cpp:9961-10100), stop below the threshold,
    record x, y (Angstrom), the best angles, defocus and pixel-size offsets and the peak height,
    then mask a disc of min_peak_radius pixels (square window [c - r, c + r)); at most 1000 peaks
    (MAX_ALLOWED_NUMBER_OF_PEAKS, match_template.cpp:7).
"""
from __future__ import annotations

import math

import numpy as np

MAX_PEAKS = 1000
PI = 3.14159265359


def expected_threshold(image_x: int, image_y: int, n_ccf: int) -> float:
    """tm_expected_threshold of the library (include/tm_results.h)."""
    from . import results
    return results.expected_threshold(image_x, image_y, n_ccf)


def find_peaks(scaled_mip: np.ndarray, threshold: float, template_size: int, min_peak_radius: int = 10,
               pixel_size: float = 1.0, maps: dict | None = None) -> list:
    z = np.array(scaled_mip, dtype=np.float32, copy=True)
    oy, ox = z.shape
    edge = template_size // 4 + 1
    r = int(min_peak_radius)
    r2 = np.float32(r) ** 2
    out = []
    inner = np.full_like(z, -np.inf)
    while len(out) < MAX_PEAKS:
        inner[edge:oy - edge, edge:ox - edge] = z[edge:oy - edge, edge:ox - edge]
        flat = int(np.argmax(inner))
        by, bx = divmod(flat, ox)
        if inner[by, bx] < threshold:
            break
        rec = {"x": bx * pixel_size, "y": by * pixel_size, "peak_height": float(scaled_mip[by, bx])}
        for k, m in (maps or {}).items():
            rec[k] = float(m[by, bx])
        out.append(rec)
        ys = np.arange(max(by - r, 0), min(by + r, oy))
        xs = np.arange(max(bx - r, 0), min(bx + r, ox))
        dy2 = (ys.astype(np.float32) - by) ** 2
        dx2 = (xs.astype(np.float32) - bx) ** 2
        mask = (dy2[:, None] + dx2[None, :]) <= r2
        sub = z[ys[0]:ys[-1] + 1, xs[0]:xs[-1] + 1]
        sub[mask] = -np.float32(np.finfo(np.float32).max)
    return out
