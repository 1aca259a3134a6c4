"""ctypes binding of the worker/master result records (include/tm_results.h): the float array a
match_template worker sends to the master (match_template.cpp:1371-1531), the master's merge
(AggregatedTemplateResult::AddResult, match_template.cpp:2015-2103) and its final scaling
(match_template.cpp:1617-1760). Host code; no GPU needed.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import core

F32P = core.F32P
I64P = core.I64P
META = 6
MAPS = 8
MAP_NAMES = ("mip", "psi", "theta", "phi", "defocus", "pixel_size", "sum", "sumsq")
RESULT_EXPORTS = ["tm_result_size", "tm_result_pack", "tm_result_add", "tm_result_finalize", "tm_result_last_error",
                  "tm_expected_threshold", "tm_histogram_table"]


def _lib():
    L = core.lib()
    if not getattr(L, "_tm_results_bound", False):
        L.tm_result_last_error.restype = ctypes.c_char_p
        L.tm_result_size.restype = ctypes.c_longlong
        L.tm_result_size.argtypes = [ctypes.c_int] * 3
        L.tm_result_pack.argtypes = [ctypes.POINTER(F32P), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, I64P, ctypes.c_int, ctypes.c_float, ctypes.c_double, F32P,
                                     ctypes.c_longlong]
        L.tm_result_add.argtypes = [F32P, F32P, ctypes.c_longlong, ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.tm_result_finalize.argtypes = [F32P, ctypes.c_longlong, ctypes.c_float, F32P, F32P, F32P, F32P, F32P, I64P]
        L.tm_expected_threshold.restype = ctypes.c_double
        L.tm_expected_threshold.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong]
        L.tm_histogram_table.argtypes = [I64P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_double,
                                         ctypes.c_longlong, F32P, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
        L._tm_results_bound = True
    return L


def _check(rc):
    if rc != 0:
        raise core.TMError(rc, _lib().tm_result_last_error().decode())


def result_size(x: int, y: int, bins: int = 512) -> int:
    """Floats in the record of an x by y map with `bins` histogram bins."""
    return int(_lib().tm_result_size(x, y, bins))


def pack(maps: dict, out_x: int, out_y: int, remove_npix_from_edge: int, histogram: np.ndarray, n_ccfs: float,
         sqrt_input_pixels: float) -> np.ndarray:
    """The worker's record from dense maps in the micrograph's orientation (MAP_NAMES keys, each
    (ny, nx); sums as floats)."""
    arrs = [np.ascontiguousarray(maps[k], np.float32) for k in MAP_NAMES]
    ny, nx = arrs[0].shape
    if any(a.shape != (ny, nx) for a in arrs):
        raise ValueError("maps differ in shape")
    h = np.ascontiguousarray(histogram, np.int64)
    out = np.zeros(result_size(out_x, out_y, h.size), np.float32)
    ptrs = (F32P * MAPS)(*[a.ctypes.data_as(F32P) for a in arrs])
    _check(_lib().tm_result_pack(ptrs, nx, ny, out_x, out_y, remove_npix_from_edge, h.ctypes.data_as(I64P), h.size,
                                 n_ccfs, sqrt_input_pixels, out.ctypes.data_as(F32P), out.size))
    return out


class Aggregate:
    """The master's collated record over any number of worker records (AddResult)."""

    def __init__(self):
        self.collated = None
        self.total_ccfs = ctypes.c_float(0.0)

    def add(self, record: np.ndarray) -> None:
        r = np.ascontiguousarray(record, np.float32)
        first = self.collated is None
        if first:
            self.collated = np.zeros(r.size, np.float32)
        _check(_lib().tm_result_add(self.collated.ctypes.data_as(F32P), r.ctypes.data_as(F32P), r.size,
                                    ctypes.byref(self.total_ccfs), int(first)))

    def finalize(self) -> dict:
        """Dense maps: mip_scaled, zscore, average, std_dev, psi, theta, phi, defocus, pixel_size and the
        histogram."""
        c = self.collated
        x, y, bins = int(c[0]), int(c[1]), int(c[3])
        out = {k: np.zeros((y, x), np.float32) for k in ("mip_scaled", "zscore", "average", "std_dev")}
        angles = np.zeros((5, y, x), np.float32)
        hist = np.zeros(bins, np.int64)
        _check(_lib().tm_result_finalize(c.ctypes.data_as(F32P), c.size, self.total_ccfs.value,
                                         *[out[k].ctypes.data_as(F32P) for k in ("mip_scaled", "zscore", "average",
                                                                                 "std_dev")],
                                         angles.ctypes.data_as(F32P), hist.ctypes.data_as(I64P)))
        for i, k in enumerate(("psi", "theta", "phi", "defocus", "pixel_size")):
            out[k] = angles[i]
        out["histogram"] = hist
        return out


def unpack(record: np.ndarray) -> dict:
    """Header values and dense maps of a record (the layout of include/tm_results.h)."""
    r = np.asarray(record, np.float32)
    x, y, n, bins = int(r[0]), int(r[1]), int(r[2]), int(r[3])
    row = n // y
    out = {"x": x, "y": y, "n": n, "bins": bins, "n_ccfs": float(r[4]), "sqrt_input_pixels": float(r[5])}
    for k, name in enumerate(MAP_NAMES):
        out[name] = r[META + k * n: META + (k + 1) * n].reshape(y, row)[:, :x].copy()
    out["histogram"] = r[META + MAPS * n: META + MAPS * n + 2 * bins].view(np.int64).copy()
    return out


def expected_threshold(image_x: int, image_y: int, n_ccf: int) -> float:
    """tm_expected_threshold: sqrtf(2) erfcinv(2 / (x y N)) with cisTEM's erfcinv (match_template.cpp:1799)."""
    return float(_lib().tm_expected_threshold(int(image_x), int(image_y), int(n_ccf)))


def histogram_table(histogram: np.ndarray, histogram_min: float, histogram_step: float, sqrt_input_pixels: float,
                    total_ccfs: int):
    """tm_histogram_table: the histogram file's columns (SNR, survival, expected survival;
    match_template.cpp:1333-1364)."""
    h = np.ascontiguousarray(histogram, np.int64)
    snr = np.zeros(h.size, np.float32)
    surv = np.zeros(h.size, np.float64)
    exp_ = np.zeros(h.size, np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    _check(_lib().tm_histogram_table(h.ctypes.data_as(I64P), int(h.size), float(histogram_min), float(histogram_step),
                                     float(sqrt_input_pixels), int(total_ccfs), snr.ctypes.data_as(F32P),
                                     surv.ctypes.data_as(dp), exp_.ctypes.data_as(dp)))
    return snr, surv, exp_
