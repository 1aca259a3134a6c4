"""ctypes bindings for oracle/libtm_oracle.so — TEST INFRASTRUCTURE ONLY.

Thin wrappers over the C API at the bottom of ``oracle/tm_oracle.cpp``. Arrays use the
reference's FFTW "padded" real layout (row stride ``nx + 2`` floats for even ``nx``,
src/core/image.cpp:4761-4846) unless a docstring says otherwise.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtm_oracle.so")
_lib = None

F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle library with its Makefile (g++ only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        L.tmo_last_error.restype = ctypes.c_char_p
        L.tmo_fft_c2c.argtypes = [F32P, ctypes.c_int, ctypes.c_int]
        L.tmo_fft_c2c_f64.argtypes = [F64P, ctypes.c_int, ctypes.c_int]
        L.tmo_fft_real_nd.argtypes = [F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tmo_phase_shift_real.argtypes = [F32P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.tmo_cross_correlate_real.argtypes = [F32P, F32P, ctypes.c_int, ctypes.c_int]
        L.tmo_euler_grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, F32P, ctypes.c_int]
        L.tmo_psi_list.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, F32P, ctypes.c_int]
        L.tmo_factorize_axis.argtypes = [ctypes.c_int, ctypes.c_int]
        L.tmo_factorized_upper.argtypes = [ctypes.c_int] * 4
        L.tmo_factorized_lower.argtypes = [ctypes.c_int] * 4
        L.tmo_erfcinv.argtypes = [ctypes.c_double]
        L.tmo_erfcinv.restype = ctypes.c_double
        L.tmo_euler_matrix.argtypes = [ctypes.c_float] * 3 + [F32P]
        L.tmo_ctf_image.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 8 + [F32P]
        L.tmo_extract_slice.argtypes = [F32P, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, F32P]
        L.tmo_mt_create.argtypes = [F32P, ctypes.POINTER(ctypes.c_int), F32P, ctypes.c_int, ctypes.c_int, F32P,
                                    ctypes.c_int, ctypes.c_int]
        L.tmo_mt_create.restype = vp
        L.tmo_mt_destroy.argtypes = [vp, ctypes.c_int]
        L.tmo_mt_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long), F64P]
        L.tmo_mt_get.argtypes = [vp, ctypes.c_int, ctypes.c_int, F32P]
        L.tmo_mt_projection_filter.argtypes = [vp, ctypes.c_int, ctypes.c_int, F32P]
        L.tmo_mt_projection_ccf.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, F32P,
                                             F32P]
        L.tmo_mt_template_for_size.argtypes = [vp, ctypes.c_int, ctypes.c_int, F32P]
        L.tmo_change_pixel_size.argtypes = [F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, F32P, ctypes.POINTER(ctypes.c_int)]
        L.tmo_set_threads.argtypes = [ctypes.c_int]
        L.tmo_cosine_mask.argtypes = [F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                      ctypes.POINTER(ctypes.c_float)]
        L.tmo_mt_run.argtypes = [vp] + [ctypes.c_int] * 8
        L.tmo_mt_results.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.tmo_mt_results.restype = ctypes.c_long
        L.tmo_mt_finalise.argtypes = [vp, ctypes.c_int, ctypes.c_long, F32P, F32P, F32P, F32P]
        L.tmo_mt_finalise.restype = ctypes.c_float
        L.tmo_rt_create.argtypes = [F32P, F32P, ctypes.c_int, ctypes.c_int, F32P, ctypes.c_int, ctypes.c_int]
        L.tmo_rt_create.restype = vp
        L.tmo_rt_destroy.argtypes = [vp, ctypes.c_int]
        L.tmo_rt_info.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_long)]
        L.tmo_rt_get.argtypes = [vp, ctypes.c_int, ctypes.c_int, F32P]
        L.tmo_rt_window.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tmo_rt_window_from_real.argtypes = [vp, ctypes.c_int, F32P]
        L.tmo_rt_score.argtypes = [vp, ctypes.c_int] + [ctypes.c_float] * 6 + [F32P, F32P]
        L.tmo_rt_peak_list.argtypes = [vp, ctypes.c_int, F32P, F32P, ctypes.c_long]
        L.tmo_rt_peak_list.restype = ctypes.c_long
        L.tmo_rt_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(F32P), ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                 ctypes.POINTER(F32P), ctypes.POINTER(RefinedPeak), ctypes.c_long]
        L.tmo_rt_run.restype = ctypes.c_long
        _lib = L
    return _lib


class RefinedPeak(ctypes.Structure):
    """One refined peak: the found peak and the search's values there, then the refined values
    (refine_template.cpp:1092-1113). Positions are pixels relative to the image centre."""
    _fields_ = [(n, ctypes.c_float) for n in
                ("x", "y", "value", "psi", "theta", "phi", "defocus", "pixel_size", "x_pos", "y_pos", "shift_x",
                 "shift_y", "r_psi", "r_theta", "r_phi", "r_defocus", "r_pixel_size", "peak_height", "mip",
                 "best_score", "starting_score", "offset")] + \
              [("best_address", ctypes.c_longlong), ("evaluations", ctypes.c_int), ("pad_", ctypes.c_int)]


def peaks_to_dicts(arr) -> list:
    return [{n: getattr(p, n) for n, _ in RefinedPeak._fields_ if n != "pad_"} for p in arr]


def set_threads(n: int) -> None:
    """Worker threads of the oracle's large transforms (results do not depend on it)."""
    lib().tmo_set_threads(int(n))


def cosine_mask(img: np.ndarray, radius: float, edge: float):
    """Image::CosineMask (image.cpp:3934-4209; centred real space, not inverted) of a dense 2-D or 3-D
    image. Returns (masked image, mask volume)."""
    a = np.array(img, np.float32, order="C", copy=True)
    shp = a.shape if a.ndim == 3 else (1,) + a.shape
    vol = ctypes.c_float(0)
    _check(lib().tmo_cosine_mask(_fp(a), shp[2], shp[1], shp[0], radius, edge, ctypes.byref(vol)))
    return a, vol.value


def change_pixel_size(img: np.ndarray, factor: float, out_shape=None, tolerance: float = 0.001,
                      return_fft: bool = False, fp64: bool = False):
    """Image::ChangePixelSize (image.cpp:8705-8792) of a dense real 2-D (ny, nx) or 3-D (nz, ny, nx)
    image into an image of out_shape (default: the input's). Returns the dense real result, or the
    (.., n/2+1) complex spectrum with return_fft; and the (pad, new) edge sizes used."""
    a = np.ascontiguousarray(img, np.float32)
    shp = a.shape if a.ndim == 3 else (1,) + a.shape
    out_shape = tuple(out_shape or a.shape)
    oshp = out_shape if len(out_shape) == 3 else (1,) + out_shape
    nz, ny, nx = shp
    oz, oy, ox = oshp
    pad = 2 if ox % 2 == 0 else 1
    out = np.zeros(oz * oy * (ox + pad), np.float32)
    dims = (ctypes.c_int * 6)()
    _check(lib().tmo_change_pixel_size(_fp(a), nx, ny, nz, factor, tolerance, ox, oy, oz, int(return_fft), int(fp64),
                                       _fp(out), dims))
    o = out.reshape(oz, oy, ox + pad)
    if return_fft:
        res = o.view(np.complex64)
    else:
        res = np.ascontiguousarray(o[..., :ox])
    if len(out_shape) == 2:
        res = res[0]
    return res, (tuple(dims[:3]), tuple(dims[3:]))


def _fp(a):
    return a.ctypes.data_as(F32P)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().tmo_last_error().decode())


# ---------------------------------------------------------------- layout helpers
def to_padded(a: np.ndarray) -> np.ndarray:
    """(ny, nx) or (nz, ny, nx) dense real -> padded layout with nx+2 (even) columns."""
    nx = a.shape[-1]
    pad = 2 if nx % 2 == 0 else 1
    out = np.zeros(a.shape[:-1] + (nx + pad,), np.float32)
    out[..., :nx] = a
    return out


def from_padded(p: np.ndarray, nx: int) -> np.ndarray:
    return np.ascontiguousarray(p[..., :nx])


def spectrum_view(p: np.ndarray) -> np.ndarray:
    """padded float array (..., nx+2) -> complex64 view (..., nx/2+1)."""
    return p.view(np.complex64)


# ---------------------------------------------------------------- FFT
def fft_c2c(x: np.ndarray, sign: int = -1) -> np.ndarray:
    if x.dtype == np.complex128:
        y = np.ascontiguousarray(x).copy()
        _check(lib().tmo_fft_c2c_f64(y.view(np.float64).ctypes.data_as(F64P), y.size, sign))
        return y
    y = np.ascontiguousarray(x, np.complex64).copy()
    _check(lib().tmo_fft_c2c(y.view(np.float32).ctypes.data_as(F32P), y.size, sign))
    return y


def rfftn(real: np.ndarray) -> np.ndarray:
    """Unscaled R2C of a dense (ny, nx) / (nz, ny, nx) float array; returns complex (.., nx/2+1)."""
    shape = real.shape
    nz, ny, nx = (1,) * (3 - len(shape)) + shape
    p = to_padded(np.asarray(real, np.float32).reshape(nz, ny, nx))
    _check(lib().tmo_fft_real_nd(_fp(p), nx, ny, nz, -1))
    return p.view(np.complex64).reshape(shape[:-1] + (nx // 2 + 1,))


def irfftn(spec: np.ndarray, nx: int) -> np.ndarray:
    """Unscaled C2R (FFTW semantics) of a complex (.., nx/2+1) array; returns dense float."""
    shape = spec.shape
    nz, ny, _ = (1,) * (3 - len(shape)) + shape
    p = np.ascontiguousarray(spec, np.complex64).reshape(nz, ny, nx // 2 + 1).copy().view(np.float32)
    _check(lib().tmo_fft_real_nd(_fp(p), nx, ny, nz, +1))
    return from_padded(p, nx).reshape(shape[:-1] + (nx,))


# ---------------------------------------------------------------- Image-level pins
def phase_shift_real(img: np.ndarray, sx: float, sy: float) -> np.ndarray:
    ny, nx = img.shape
    p = to_padded(img)
    _check(lib().tmo_phase_shift_real(_fp(p), nx, ny, sx, sy))
    return from_padded(p, nx)


def cross_correlate_real(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ny, nx = a.shape
    pa, pb = to_padded(a), to_padded(b)
    _check(lib().tmo_cross_correlate_real(_fp(pa), _fp(pb), nx, ny))
    return from_padded(pa, nx)


def euler_grid(sym="C1", step=10.0) -> np.ndarray:
    t, n = sym[0].upper(), int(sym[1:] or 0)
    cap = 1 << 22
    out = np.zeros(2 * cap, np.float32)
    cnt = lib().tmo_euler_grid(ord(t), n, step, _fp(out), cap)
    return out[: 2 * cnt].reshape(cnt, 2).copy()


def psi_list(start=0.0, mx=360.0, step=10.0) -> np.ndarray:
    out = np.zeros(100000, np.float32)
    n = lib().tmo_psi_list(start, mx, step, _fp(out), out.size)
    return out[:n].copy()


def factorize_axis(original: int, reference_size: int) -> int:
    return lib().tmo_factorize_axis(original, reference_size)


def erfcinv(x: float) -> float:
    return lib().tmo_erfcinv(x)


def euler_matrix(phi, theta, psi) -> np.ndarray:
    m = np.zeros(9, np.float32)
    lib().tmo_euler_matrix(phi, theta, psi, _fp(m))
    return m.reshape(3, 3)


def ctf_image(n, kv=300.0, cs=2.7, amp=0.07, df1=15000.0, df2=15000.0, astig=0.0, pixel_size=1.0, phase_deg=0.0):
    out = np.zeros((n, n // 2 + 1), np.complex64)
    _check(lib().tmo_ctf_image(n, n, kv, cs, amp, df1, df2, astig, pixel_size, phase_deg, _fp(out.view(np.float32))))
    return out


def extract_slice(volume_spectrum: np.ndarray, phi, theta, psi) -> np.ndarray:
    P = volume_spectrum.shape[0]
    v = np.ascontiguousarray(volume_spectrum, np.complex64)
    out = np.zeros((P, P // 2 + 1), np.complex64)
    _check(lib().tmo_extract_slice(_fp(v.view(np.float32)), P, phi, theta, psi, _fp(out.view(np.float32))))
    return out


# ---------------------------------------------------------------- match_template pipeline
@dataclass
class MTParams:
    """match_template arguments (match_template.cpp:346-387) used by the path."""
    pixel_size: float = 1.0
    voltage_kV: float = 300.0
    spherical_aberration_mm: float = 2.7
    amplitude_contrast: float = 0.07
    defocus1: float = 15000.0
    defocus2: float = 15000.0
    defocus_angle: float = 0.0
    phase_shift_deg: float = 0.0
    high_resolution_limit_search: float = 2.0
    angular_step: float = 10.0
    in_plane_angular_step: float = 10.0
    defocus_search_range: float = 0.0
    defocus_step: float = 0.0
    pixel_size_search_range: float = 0.0
    pixel_size_step: float = 0.0
    padding: float = 1.0
    particle_radius_angstroms: float = 0.0
    symmetry: str = "C1"
    first_search_position: int = -1
    last_search_position: int = -1

    def packed(self):
        f = np.array([self.pixel_size, self.voltage_kV, self.spherical_aberration_mm, self.amplitude_contrast,
                      self.defocus1, self.defocus2, self.defocus_angle, self.phase_shift_deg,
                      self.high_resolution_limit_search, self.angular_step, self.in_plane_angular_step,
                      self.defocus_search_range, self.defocus_step, self.pixel_size_search_range,
                      self.pixel_size_step, self.padding, self.particle_radius_angstroms], np.float32)
        i = np.array([ord(self.symmetry[0].upper()), int(self.symmetry[1:] or 0),
                      self.first_search_position, self.last_search_position], np.int32)
        return f, i


@dataclass
class MTInfo:
    W: int
    H: int
    T: int
    P: int
    n_pos: int
    n_psi: int
    first_pos: int
    last_pos: int
    n_def_half: int
    rotated: bool
    max_padding: int
    numel: int
    n_curve: int
    total_correlation_positions: int
    sqrt_input_pixels: float
    histogram_min_scaled: float
    histogram_step_scaled: float
    defocus_step: float
    pixel_size_step: float
    n_size_half: int = 0


class MatchTemplateOracle:
    """The reference CPU match_template path (setup, hot loop, finalise) on the CPU."""

    def __init__(self, params: MTParams, micrograph: np.ndarray, volume: np.ndarray, fp64: bool = False):
        L = lib()
        f, i = params.packed()
        m = np.ascontiguousarray(micrograph, np.float32)
        v = np.ascontiguousarray(volume, np.float32)
        self.fp64 = int(fp64)
        self.h = L.tmo_mt_create(_fp(f), i.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _fp(m), m.shape[1],
                                 m.shape[0], _fp(v), v.shape[0], self.fp64)
        if not self.h:
            raise RuntimeError("oracle: " + L.tmo_last_error().decode())
        info = np.zeros(16, np.int64)
        dinfo = np.zeros(8, np.float64)
        L.tmo_mt_info(self.h, self.fp64, info.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                      dinfo.ctypes.data_as(F64P))
        self.info = MTInfo(*[int(x) for x in info[:14]], *[float(x) for x in dinfo[:5]])
        self.info.n_size_half = int(info[14])
        self.info.rotated = bool(self.info.rotated)
        self.params = params

    def __del__(self):
        if getattr(self, "h", None):
            lib().tmo_mt_destroy(self.h, self.fp64)
            self.h = None

    def _get(self, which, n):
        out = np.zeros(n, np.float32)
        lib().tmo_mt_get(self.h, self.fp64, which, _fp(out))
        return out

    @property
    def input_spectrum(self) -> np.ndarray:
        """(H, W/2+1) complex64 whitened, normalised, quadrant-swapped input spectrum."""
        I = self.info
        return self._get(0, I.numel).view(np.complex64).reshape(I.H, I.W // 2 + 1)

    @property
    def template_spectrum(self) -> np.ndarray:
        """(P, P, P/2+1) complex64, DC 0, quadrant-swapped."""
        P = self.info.P
        return self._get(1, P * P * (P + 2)).view(np.complex64).reshape(P, P, P // 2 + 1)

    @property
    def grid(self) -> np.ndarray:
        return self._get(2, 2 * self.info.n_pos).reshape(-1, 2)

    @property
    def psi(self) -> np.ndarray:
        return self._get(3, self.info.n_psi)

    @property
    def whitening_curve(self):
        n = self.info.n_curve
        return self._get(5, n), self._get(4, n)

    def projection_filter(self, defocus_i: int = 0) -> np.ndarray:
        T = self.info.T
        out = np.zeros((T, T // 2 + 1), np.complex64)
        _check(lib().tmo_mt_projection_filter(self.h, self.fp64, defocus_i, _fp(out.view(np.float32))))
        return out

    def template_for_size(self, size_i: int) -> np.ndarray:
        """(P, P, P/2+1) complex64 template spectrum of pixel-size plane size_i (match_template.cpp:837-845)."""
        P = self.info.P
        out = np.zeros(P * P * (P + 2), np.float32)
        _check(lib().tmo_mt_template_for_size(self.h, self.fp64, size_i, _fp(out)))
        return out.view(np.complex64).reshape(P, P, P // 2 + 1)

    def projection_and_ccf(self, pos: int, psi_idx: int, defocus_i: int = 0, want_ccf: bool = True, size_i: int = 0):
        I = self.info
        proj = np.zeros((I.T, I.T + 2), np.float32)
        ccf = np.zeros((I.H, I.W + 2), np.float32) if want_ccf else None
        _check(lib().tmo_mt_projection_ccf(self.h, self.fp64, pos, psi_idx, defocus_i, size_i, _fp(proj),
                                           _fp(ccf) if want_ccf else None))
        return proj[:, : I.T].copy(), (ccf[:, : I.W].copy() if want_ccf else None)

    def run(self, threads: int = 1, pos_lo: int = -1, pos_hi: int = -1, def_lo: int = -(1 << 30),
            def_hi: int = 1 << 30, psi_lo: int = 0, psi_hi: int = (1 << 30)):
        _check(lib().tmo_mt_run(self.h, self.fp64, threads, pos_lo, pos_hi, def_lo, def_hi, psi_lo, psi_hi))
        return self.results()

    def results(self) -> dict:
        I = self.info
        L = lib()
        res = {}
        for which, name, dt in [(0, "mip", np.float32), (1, "psi", np.float32), (2, "theta", np.float32),
                                (3, "phi", np.float32), (4, "defocus", np.float32), (5, "pixel_size", np.float32),
                                (6, "sum", np.float64), (7, "sumsq", np.float64), (8, "best_index", np.int32)]:
            a = np.zeros(I.numel, dt)
            nccf = L.tmo_mt_results(self.h, self.fp64, which, a.ctypes.data_as(ctypes.c_void_p))
            res[name] = a.reshape(I.H, I.W + 2)[:, : I.W].copy()
        h = np.zeros(512, np.int64)
        res["nccf"] = L.tmo_mt_results(self.h, self.fp64, 9, h.ctypes.data_as(ctypes.c_void_p))
        res["hist"] = h
        return res

    def finalise(self, total_ccfs: int | None = None) -> dict:
        I = self.info
        if total_ccfs is None:
            total_ccfs = I.total_correlation_positions
        arrs = [np.zeros(I.numel, np.float32) for _ in range(4)]
        thr = lib().tmo_mt_finalise(self.h, self.fp64, total_ccfs, *[_fp(a) for a in arrs])
        names = ["mip_scaled", "zscore", "avg", "std"]
        out = {n: a.reshape(I.H, I.W + 2)[:, : I.W].copy() for n, a in zip(names, arrs)}
        out["expected_threshold"] = float(thr)
        return out


# ---------------------------------------------------------------- refine_template
@dataclass
class RTParams:
    """refine_template arguments (refine_template.cpp:264-312) used by the restatement."""
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
    pixel_size_search_range: float = 0.0
    pixel_size_search_step: float = 0.0
    mask_radius: float = 0.0

    def packed(self):
        return np.array([self.pixel_size, self.voltage_kV, self.spherical_aberration_mm, self.amplitude_contrast,
                         self.defocus1, self.defocus2, self.defocus_angle, self.phase_shift_deg, self.angular_step,
                         self.in_plane_angular_step, self.defocus_search_range, self.defocus_search_step,
                         self.padding, self.wanted_threshold, self.min_peak_radius, self.pixel_size_search_range,
                         self.pixel_size_search_step, self.mask_radius], np.float32)


MAP_NAMES = ("mip", "scaled_mip", "psi", "theta", "phi", "defocus", "pixel_size")


class RefineTemplateOracle:
    """The reference refine_template program for one image, on the CPU."""

    def __init__(self, params: RTParams, micrograph: np.ndarray, volume: np.ndarray, fp64: bool = False):
        L = lib()
        m = np.ascontiguousarray(micrograph, np.float32)
        v = np.ascontiguousarray(volume, np.float32)
        self.fp64 = int(fp64)
        self.h = L.tmo_rt_create(_fp(params.packed()), _fp(m), m.shape[1], m.shape[0], _fp(v), v.shape[0], self.fp64)
        if not self.h:
            raise RuntimeError("oracle: " + L.tmo_last_error().decode())
        info = np.zeros(8, np.int64)
        L.tmo_rt_info(self.h, self.fp64, info.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
        self.mx, self.my, self.t, self.P, self.n_curve = [int(x) for x in info[:5]]
        self.params = params

    def __del__(self):
        if getattr(self, "h", None):
            lib().tmo_rt_destroy(self.h, self.fp64)
            self.h = None

    def _get(self, which, n):
        out = np.zeros(n, np.float32)
        lib().tmo_rt_get(self.h, self.fp64, which, _fp(out))
        return out

    def _pad(self):
        return self.mx + (2 if self.mx % 2 == 0 else 1)

    @property
    def whitened_image(self) -> np.ndarray:
        """(my, mx) whitened, normalised input image in real space (refine_template.cpp:526-546)."""
        return self._get(0, self._pad() * self.my).reshape(self.my, self._pad())[:, : self.mx].copy()

    @property
    def template_spectrum(self) -> np.ndarray:
        P = self.P
        return self._get(1, P * P * (P + 2)).view(np.complex64).reshape(P, P, P // 2 + 1)

    @property
    def whitening_curve(self):
        n = self.n_curve
        return self._get(3, n), self._get(2, n)

    @property
    def window_spectrum(self) -> np.ndarray:
        t = self.t
        return self._get(4, t * (t + 2)).view(np.complex64).reshape(t, t // 2 + 1)

    @property
    def projection_filter(self) -> np.ndarray:
        t = self.t
        return self._get(5, t * (t + 2)).view(np.complex64).reshape(t, t // 2 + 1)

    def window(self, peak_x: int, peak_y: int) -> None:
        _check(lib().tmo_rt_window(self.h, self.fp64, int(peak_x), int(peak_y)))

    def window_from_real(self, img: np.ndarray) -> None:
        p = to_padded(np.asarray(img, np.float32))
        _check(lib().tmo_rt_window_from_real(self.h, self.fp64, _fp(p)))

    def score(self, phi, theta, psi, current_defocus=0.0, step=0.0, want_projection=False, pixel_size_factor=1.0):
        """TemplateScore: (x, y, value) of the correlation peak; with want_projection also the
        real-space normalised filtered projection (t, t)."""
        out = np.zeros(3, np.float32)
        t = self.t
        proj = np.zeros((t, t + 2), np.float32) if want_projection else None
        _check(lib().tmo_rt_score(self.h, self.fp64, current_defocus, step, phi, theta, psi, pixel_size_factor, _fp(out),
                                  _fp(proj) if want_projection else None))
        res = (float(out[0]), float(out[1]), float(out[2]))
        return (res, proj[:, :t].copy()) if want_projection else res

    def peak_list(self, scaled_mip: np.ndarray, cap: int = 100000) -> np.ndarray:
        sm = to_padded(np.asarray(scaled_mip, np.float32))
        out = np.zeros((cap, 3), np.float32)
        n = lib().tmo_rt_peak_list(self.h, self.fp64, _fp(sm), _fp(out), cap)
        if n < 0:
            raise RuntimeError("oracle: " + lib().tmo_last_error().decode())
        return out[: min(n, cap)].copy()

    def run(self, maps: dict, shared_local_maps: bool = False, xy_change_threshold: float = 10.0,
            exclude_above_xy_threshold: bool = False, cap: int = 10000):
        """Refines every peak. maps: dense (my, mx) arrays named as MAP_NAMES. Returns
        (list of peak dicts, dict of output maps best_*)."""
        ins = [to_padded(np.asarray(maps[n], np.float32)) for n in MAP_NAMES]
        outs = [np.zeros_like(a) for a in ins]
        in_ptrs = (F32P * 7)(*[_fp(a) for a in ins])
        out_ptrs = (F32P * 7)(*[_fp(a) for a in outs])
        peaks = (RefinedPeak * cap)()
        n = lib().tmo_rt_run(self.h, self.fp64, in_ptrs, int(shared_local_maps), xy_change_threshold,
                             int(exclude_above_xy_threshold), out_ptrs, peaks, cap)
        if n < 0:
            raise RuntimeError("oracle: " + lib().tmo_last_error().decode())
        return peaks_to_dicts(peaks[: min(n, cap)]), {n: a[:, : self.mx].copy() for n, a in zip(MAP_NAMES, outs)}
