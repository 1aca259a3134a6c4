"""Host driver mirroring the reference's ``refine_template`` program for one image on the GPU.

``TemplateRefiner`` does what RefineTemplateApp::DoCalculation does
(src/programs/refine_template/refine_template.cpp:264-1229): whiten the image and transform it back
to real space, pad and transform the template, take the peaks of the scaled MIP, window each peak
and run the coordinate descent over defocus and the Euler angles with TemplateScore. Everything
runs through the C ABI of libtm_hip.so (include/tm_refine.h, include/tm_prepare.h); there is no CPU
path. ``mask_radius`` > 0 masks each particle window with the program's CosineMask (:716).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import core
from .match_template import _check as _pcheck, _lib as _plib, prepare_template

F32P = core.F32P

MAP_NAMES = ("mip", "scaled_mip", "psi", "theta", "phi", "defocus", "pixel_size")
OUTPUT_NAMES = ("best_mip", "best_scaled_mip", "best_psi", "best_theta", "best_phi", "best_defocus",
                "best_pixel_size")
REFINE_EXPORTS = ["tm_prepare_refine_image", "tm_refine_create", "tm_refine_destroy", "tm_refine_init",
                  "tm_refine_run", "tm_refine_peak_list", "tm_refine_score", "tm_refine_set_timing", "tm_refine_set_fused",
                  "tm_refine_get_timing", "tm_refine_last_error", "tm_refine_score_batch", "tm_refine_set_image",
                  "tm_refine_set_micrograph", "tm_refine_get_image"]


class RefineParams(ctypes.Structure):
    """tm_refine_params (include/tm_refine.h)."""
    _fields_ = ([("image_x", ctypes.c_int), ("image_y", ctypes.c_int), ("template_size", ctypes.c_int),
                 ("padded_size", ctypes.c_int)] +
                [(n, ctypes.c_float) for n in ("pixel_size", "voltage_kV", "spherical_aberration_mm",
                                               "amplitude_contrast", "defocus1", "defocus2", "defocus_angle",
                                               "phase_shift_deg", "angular_step", "in_plane_angular_step",
                                               "defocus_search_range", "defocus_search_step", "wanted_threshold",
                                               "min_peak_radius", "xy_change_threshold")] +
                [("exclude_above_xy_threshold", ctypes.c_int), ("shared_local_maps", ctypes.c_int),
                 ("max_in_flight", ctypes.c_int), ("pixel_size_search_range", ctypes.c_float),
                 ("pixel_size_search_step", ctypes.c_float), ("mask_radius", ctypes.c_float)])


class RefinedPeak(ctypes.Structure):
    """tm_refined_peak (include/tm_refine.h)."""
    _fields_ = [(n, ctypes.c_float) for n in
                ("x", "y", "value", "psi", "theta", "phi", "defocus", "pixel_size", "x_pos", "y_pos", "shift_x",
                 "shift_y", "r_psi", "r_theta", "r_phi", "r_defocus", "r_pixel_size", "peak_height", "mip",
                 "best_score", "starting_score", "offset")] + \
              [("best_address", ctypes.c_longlong), ("evaluations", ctypes.c_int), ("reserved", ctypes.c_int)]


def _lib():
    L = _plib()
    if not getattr(L, "_tm_refine_bound", False):
        vp = ctypes.c_void_p
        ip = ctypes.POINTER(ctypes.c_int)
        L.tm_prepare_refine_image.argtypes = [ctypes.c_int, F32P, ctypes.c_int, ctypes.c_int, F32P, F32P, F32P, ip]
        L.tm_refine_last_error.restype = ctypes.c_char_p
        L.tm_refine_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.tm_refine_destroy.argtypes = [vp]
        L.tm_refine_init.argtypes = [vp, ctypes.POINTER(RefineParams), F32P, F32P, F32P, F32P, ctypes.c_int]
        L.tm_refine_run.argtypes = [vp, ctypes.POINTER(F32P), ctypes.POINTER(F32P), ctypes.POINTER(RefinedPeak),
                                    ctypes.c_int, ip]
        L.tm_refine_peak_list.argtypes = [vp, F32P, F32P, ctypes.c_int, ip]
        L.tm_refine_score.argtypes = [vp, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 6 + [F32P]
        L.tm_refine_score_batch.argtypes = [vp, ctypes.c_int, ip, F32P, ctypes.c_float, F32P]
        L.tm_refine_set_image.argtypes = [vp, F32P, F32P, F32P, ctypes.c_int]
        L.tm_refine_set_micrograph.argtypes = [vp, F32P]
        L.tm_refine_get_image.argtypes = [vp, F32P, F32P, F32P]
        L.tm_refine_set_timing.argtypes = [vp, ctypes.c_int]
        L.tm_refine_set_fused.argtypes = [vp, ctypes.c_int]
        i64p = ctypes.POINTER(ctypes.c_longlong)
        L.tm_refine_get_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double), i64p,
                                           i64p, ctypes.c_int]
        L._tm_refine_bound = True
    return L


def _check(rc):
    if rc != 0:
        raise core.TMError(rc, _lib().tm_refine_last_error().decode())


def _f(a):
    return a.ctypes.data_as(F32P)


def prepare_refine_image(micrograph: np.ndarray, device: int = 0):
    """Whitened real-space image and whitening curve (refine_template.cpp:526-546)."""
    L = _lib()
    m = np.ascontiguousarray(micrograph, np.float32)
    my, mx = m.shape
    n = ctypes.c_int(0)
    _pcheck(L.tm_prepare_refine_image(device, _f(m), mx, my, None, None, None, ctypes.byref(n)))
    out = np.zeros((my, mx), np.float32)
    cx = np.zeros(n.value, np.float32)
    cy = np.zeros(n.value, np.float32)
    _pcheck(L.tm_prepare_refine_image(device, _f(m), mx, my, _f(out), _f(cx), _f(cy), ctypes.byref(n)))
    return out, cx, cy


@dataclass
class RefineSettings:
    """refine_template arguments used by the path (refine_template.cpp:264-312)."""
    pixel_size: float = 1.0
    voltage_kV: float = 300.0
    spherical_aberration_mm: float = 2.7
    amplitude_contrast: float = 0.07
    defocus1: float = 15000.0
    defocus2: float = 15000.0
    defocus_angle: float = 0.0
    phase_shift_deg: float = 0.0
    angular_step: float = 1.5
    in_plane_angular_step: float = 1.5
    defocus_search_range: float = 0.0
    defocus_search_step: float = 0.0
    padding: float = 1.0
    wanted_threshold: float = 7.5
    min_peak_radius: float = 10.0
    xy_change_threshold: float = 10.0
    exclude_above_xy_threshold: bool = False
    shared_local_maps: bool = False
    max_in_flight: int = 0
    pixel_size_search_range: float = 0.0
    pixel_size_search_step: float = 0.0
    mask_radius: float = 0.0


class TemplateRefiner:
    """Refinement of the peaks of one image on one GPU."""

    def __init__(self, micrograph: np.ndarray, volume: np.ndarray, s: RefineSettings, device: int = 0,
                 template_spectrum: np.ndarray | None = None, whitened=None):
        self.s = s
        my, mx = micrograph.shape
        t = volume.shape[0]
        self.mx, self.my, self.t = mx, my, t
        img, cx, cy = whitened if whitened is not None else prepare_refine_image(micrograph, device)
        tpl = prepare_template(volume, s.padding, device) if template_spectrum is None else template_spectrum
        self.whitened, self.curve_x, self.curve_y = img, cx, cy
        self.P = tpl.shape[0]
        L = _lib()
        self.h = ctypes.c_void_p()
        _check(L.tm_refine_create(device, ctypes.byref(self.h)))
        p = RefineParams(mx, my, t, self.P, s.pixel_size, s.voltage_kV, s.spherical_aberration_mm,
                         s.amplitude_contrast, s.defocus1, s.defocus2, s.defocus_angle, s.phase_shift_deg,
                         s.angular_step, s.in_plane_angular_step, s.defocus_search_range, s.defocus_search_step,
                         s.wanted_threshold, s.min_peak_radius, s.xy_change_threshold,
                         int(s.exclude_above_xy_threshold), int(s.shared_local_maps), s.max_in_flight,
                         s.pixel_size_search_range, s.pixel_size_search_step, s.mask_radius)
        t32 = np.ascontiguousarray(tpl, np.complex64)
        _check(L.tm_refine_init(self.h, ctypes.byref(p), _f(np.ascontiguousarray(img, np.float32)),
                                _f(t32.view(np.float32)), _f(np.ascontiguousarray(cx, np.float32)),
                                _f(np.ascontiguousarray(cy, np.float32)), int(cx.size)))

    def close(self):
        if getattr(self, "h", None):
            _lib().tm_refine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_image(self, micrograph: np.ndarray, whitened=None, device: int = 0) -> None:
        """Replaces the micrograph (same size), keeping the template: tm_refine_set_image."""
        if micrograph.shape != (self.my, self.mx):
            raise ValueError(f"micrograph shape {micrograph.shape} != {(self.my, self.mx)}")
        img, cx, cy = whitened if whitened is not None else prepare_refine_image(micrograph, device)
        self.whitened, self.curve_x, self.curve_y = img, cx, cy
        _check(_lib().tm_refine_set_image(self.h, _f(np.ascontiguousarray(img, np.float32)),
                                          _f(np.ascontiguousarray(cx, np.float32)),
                                          _f(np.ascontiguousarray(cy, np.float32)), int(cx.size)))

    def set_micrograph(self, micrograph: np.ndarray) -> None:
        """Replaces the micrograph (same size), prepared on the device: tm_refine_set_micrograph."""
        m = np.ascontiguousarray(micrograph, np.float32)
        if m.shape != (self.my, self.mx):
            raise ValueError(f"micrograph shape {m.shape} != {(self.my, self.mx)}")
        _check(_lib().tm_refine_set_micrograph(self.h, _f(m)))

    def image(self):
        """The handle's whitened image and whitening curve (x, y): tm_refine_get_image."""
        img = np.zeros((self.my, self.mx), np.float32)
        cx = np.zeros(self.curve_x.size, np.float32)
        cy = np.zeros(self.curve_x.size, np.float32)
        _check(_lib().tm_refine_get_image(self.h, _f(img), _f(cx), _f(cy)))
        return img, cx, cy

    def peak_list(self, scaled_mip: np.ndarray, cap: int = 100000) -> np.ndarray:
        sm = np.ascontiguousarray(scaled_mip, np.float32)
        out = np.zeros((cap, 3), np.float32)
        n = ctypes.c_int(0)
        _check(_lib().tm_refine_peak_list(self.h, _f(sm), _f(out), cap, ctypes.byref(n)))
        return out[: min(n.value, cap)].copy()

    def score(self, peak_x: int, peak_y: int, phi, theta, psi, current_defocus=0.0, step=0.0, pixel_size_factor=1.0):
        """TemplateScore for the window at the peak: (x, y, value)."""
        out = np.zeros(3, np.float32)
        _check(_lib().tm_refine_score(self.h, int(peak_x), int(peak_y), current_defocus, step, phi, theta, psi,
                                      pixel_size_factor, _f(out)))
        return float(out[0]), float(out[1]), float(out[2])

    def score_batch(self, peak_xy: np.ndarray, angles: np.ndarray, current_defocus: float = 0.0) -> np.ndarray:
        """TemplateScore for each (window, orientation) request: peak_xy (n, 2) int pixels from the
        image centre, angles (n, 3) phi, theta, psi. Returns (n, 3): x, y offset and value."""
        xy = np.ascontiguousarray(peak_xy, np.int32).reshape(-1, 2)
        ang = np.ascontiguousarray(angles, np.float32).reshape(-1, 3)
        if xy.shape[0] != ang.shape[0]:
            raise ValueError("peak_xy and angles differ in length")
        out = np.zeros((xy.shape[0], 3), np.float32)
        _check(_lib().tm_refine_score_batch(self.h, xy.shape[0], xy.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                            _f(ang), current_defocus, _f(out)))
        return out

    def set_timing(self, on: bool) -> None:
        _check(_lib().tm_refine_set_timing(self.h, int(on)))

    def set_fused(self, on: bool) -> None:
        """One-kernel TemplateScore batches for t = 64 (tm_refine_set_fused; on by default)."""
        _check(_lib().tm_refine_set_fused(self.h, int(on)))

    def timing(self) -> dict:
        """{stage: (total ms, launches, items)} over the last run."""
        cap = 8
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        nl = (ctypes.c_longlong * cap)()
        it = (ctypes.c_longlong * cap)()
        n = _lib().tm_refine_get_timing(self.h, names, ms, nl, it, cap)
        return {names[i].decode(): (ms[i], nl[i], it[i]) for i in range(n)}

    def run(self, maps: dict, cap: int = 100000, want_maps: bool = True):
        """Refines every peak. maps: (my, mx) arrays named as MAP_NAMES (the search's outputs).
        Returns (list of peak dicts, dict of output maps named as OUTPUT_NAMES or None)."""
        ins = [np.ascontiguousarray(maps[n], np.float32) for n in MAP_NAMES]
        for a in ins:
            if a.shape != (self.my, self.mx):
                raise ValueError(f"map shape {a.shape} != {(self.my, self.mx)}")
        outs = [np.zeros((self.my, self.mx), np.float32) for _ in MAP_NAMES] if want_maps else None
        in_ptrs = (F32P * 7)(*[_f(a) for a in ins])
        out_ptrs = (F32P * 7)(*[_f(a) for a in outs]) if want_maps else None
        peaks = (RefinedPeak * cap)()
        n = ctypes.c_int(0)
        _check(_lib().tm_refine_run(self.h, in_ptrs, out_ptrs, peaks, cap, ctypes.byref(n)))
        recs = [{f: getattr(p, f) for f, _ in RefinedPeak._fields_ if f != "reserved"} for p in peaks[: min(n.value, cap)]]
        return recs, (dict(zip(OUTPUT_NAMES, outs)) if want_maps else None)
