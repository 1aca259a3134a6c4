"""Oracle (test infrastructure only): numpy restatement of the worker/master result records of the
reference's job system. Only tests/ may import it.

* worker record: match_template.cpp:1371-1531 (meta values, trim / pad of every map with its central
  average, maps in padded real-space rows, histogram longs as float pairs)
* central average: Image::ReturnAverageOfRealValues(radius, false), image.cpp:6263-6345
* resize: Image::Resize -> ClipInto, image.cpp:8825-8838, 8522-8575 (centred copy, fill value)
* master merge: AggregatedTemplateResult::AddResult, match_template.cpp:2015-2103
* master scaling: match_template.cpp:1617-1760 (float arithmetic; MIP and sums times sqrt(W H) in double)

Pinned by (tests/test_results_cpu.py): known-answer crops and pads, constant maps, the disc's pixel
count against the lattice points of the disc, and, through the oracle's own CPU search, the merge of
two workers' records over disjoint position ranges against one search over their union (every map
inside the valid area, the histogram and the CCF count), with the master's z-score against the local
path's closed form.
"""
from __future__ import annotations

import numpy as np

META, MAPS = 6, 8


def row_jump(x: int) -> int:
    return 2 if x % 2 == 0 else 1


def map_floats(x: int, y: int) -> int:
    return (x + row_jump(x)) * y


def record_size(x: int, y: int, bins: int) -> int:
    return META + MAPS * map_floats(x, y) + 2 * bins


def disc_average(img: np.ndarray, radius: np.float32) -> np.float32:
    """ReturnAverageOfRealValues(radius, false) on a dense (ny, nx) image (image.cpp:6280-6326)."""
    ny, nx = img.shape
    radius = np.float32(radius)
    if radius > 0:
        r2 = radius * radius
        dy = (np.arange(ny) - ny // 2).astype(np.float32)
        dx = (np.arange(nx) - nx // 2).astype(np.float32)
        d2 = (dx[None, :] * dx[None, :]) + (dy[:, None] * dy[:, None])
        inside = d2 <= r2
        n = int(inside.sum())
        if n == 0:
            return np.float32(0.0)
        # sequential FP64 sum in row-major order, as the reference's loop
        s = 0.0
        for v in img[inside].astype(np.float64):
            s += v
        return np.float32(s / n)
    s = 0.0
    for v in img.ravel().astype(np.float64):
        s += v
    return np.float32(s / (nx * ny))


def resize(img: np.ndarray, ox: int, oy: int, fill: np.float32) -> np.ndarray:
    """Image::Resize (ClipInto centred, image.cpp:8545-8575)."""
    ny, nx = img.shape
    if (nx, ny) == (ox, oy):
        return img.copy()
    out = np.full((oy, ox), fill, np.float32)
    for jj in range(oy):
        j = ny // 2 + jj - oy // 2
        if not 0 <= j < ny:
            continue
        for ii in range(ox):
            i = nx // 2 + ii - ox // 2
            if 0 <= i < nx:
                out[jj, ii] = img[j, i]
    return out


def pack(maps: list, out_x: int, out_y: int, remove: int, histogram: np.ndarray, n_ccfs: float,
         sqrt_input_pixels: float) -> np.ndarray:
    """The worker's record (match_template.cpp:1389-1531); maps: 8 dense (ny, nx) float32 arrays."""
    ny, nx = maps[0].shape
    trim_x, trim_y = out_x - remove, out_y - remove
    if out_x > nx:
        trim_x = nx - remove
    if out_y > ny:
        trim_y = ny - remove
    central = np.float32(0.35 * (nx + ny - 2 * remove))
    n = map_floats(out_x, out_y)
    j = row_jump(out_x)
    bins = len(histogram)
    rec = np.zeros(record_size(out_x, out_y, bins), np.float32)
    rec[:META] = [out_x, out_y, n, bins, n_ccfs, np.float32(sqrt_input_pixels)]
    for k, m in enumerate(maps):
        m = np.asarray(m, np.float32)
        avg = disc_average(m, central)
        m = resize(resize(m, trim_x, trim_y, avg), out_x, out_y, avg)
        rows = np.zeros((out_y, out_x + j), np.float32)
        rows[:, :out_x] = m
        rec[META + k * n: META + (k + 1) * n] = rows.ravel()
    rec[META + MAPS * n:] = np.asarray(histogram, np.int64).view(np.float32)
    return rec


def add(collated, record: np.ndarray, total_ccfs: np.float32):
    """AddResult (match_template.cpp:2015-2103). collated None starts a new collated record.
    Returns (collated, total_ccfs)."""
    r = np.asarray(record, np.float32)
    n, bins = int(r[2]), int(r[3])
    if collated is None:
        collated = np.zeros_like(r)
        collated[:4] = r[:4]
        collated[5] = r[5]
        total_ccfs = np.float32(0.0)
    total_ccfs = np.float32(total_ccfs + r[4])
    cm = collated[META:META + MAPS * n].reshape(MAPS, n)
    rm = r[META:META + MAPS * n].reshape(MAPS, n)
    win = rm[0] > cm[0]
    cm[:6, win] = rm[:6, win]
    cm[6] += rm[6]
    cm[7] += rm[7]
    h = collated[META + MAPS * n:].view(np.int64)
    h += r[META + MAPS * n: META + MAPS * n + 2 * bins].view(np.int64)
    return collated, total_ccfs


def finalize(collated: np.ndarray, total_ccfs: np.float32) -> dict:
    """The master's scaling (match_template.cpp:1704-1760): dense maps."""
    c = np.asarray(collated, np.float32)
    x, y, n, bins = int(c[0]), int(c[1]), int(c[2]), int(c[3])
    sqrt_px = np.float64(c[5])
    total = np.float32(total_ccfs)
    m = c[META:META + MAPS * n].reshape(MAPS, y, n // y)[:, :, :x]
    mip = m[0]
    mean = (m[6] / total).astype(np.float32)
    with np.errstate(invalid="ignore"):
        sd = np.sqrt((m[7] / total).astype(np.float32) - np.power(mean, np.float32(2.0))).astype(np.float32)
    ok = sd > 0
    z = np.where(ok, (mip - mean) / np.where(ok, sd, np.float32(1.0)), np.float32(0.0)).astype(np.float32)
    sd = np.where(ok, sd, np.float32(0.0)).astype(np.float32)
    out = {"mip_scaled": (mip.astype(np.float64) * sqrt_px).astype(np.float32), "zscore": z,
           "average": (mean.astype(np.float64) * sqrt_px).astype(np.float32),
           "std_dev": (sd.astype(np.float64) * sqrt_px).astype(np.float32)}
    for k, name in enumerate(("psi", "theta", "phi", "defocus", "pixel_size")):
        out[name] = m[k + 1].copy()
    out["histogram"] = c[META + MAPS * n: META + MAPS * n + 2 * bins].view(np.int64).copy()
    return out
