"""Host driver mirroring the reference's ``match_template`` flow around the GPU inner loop.

``MatchTemplateSearch`` does what MatchTemplateApp::DoCalculation does for one image
(src/programs/match_template/match_template.cpp:490-1370): factorise/pad/rotate the image and
whiten it, pad and transform the template, build the Euler grid and the in-plane angles, then for
each defocus plane build the CTF x whitening filter and call the inner loop, and finally scale the
maps and compute the z-score. The preparation steps and the search run through the C ABI of
libtm_hip.so (include/tm_prepare.h, include/tm_hip.h); nothing here computes on the CPU beyond
bookkeeping and the final un-rotation/cropping of the maps.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from dataclasses import dataclass

import numpy as np

from . import core, results

F32P = core.F32P


def _lib():
    L = core.lib()
    if not getattr(L, "_tm_prepare_bound", False):
        ip = ctypes.POINTER(ctypes.c_int)
        L.tm_prepare_last_error.restype = ctypes.c_char_p
        L.tm_prepare_sizes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, ip, ip, ip]
        L.tm_prepare_input_image.argtypes = [ctypes.c_int, F32P, ctypes.c_int, ctypes.c_int, ctypes.c_int, F32P, F32P,
                                             F32P]
        L.tm_prepare_template.argtypes = [ctypes.c_int, F32P, ctypes.c_int, ctypes.c_float, F32P, ip]
        L.tm_prepare_template_scaled.argtypes = [ctypes.c_int, F32P, ctypes.c_int, ctypes.c_float, ctypes.c_float, F32P,
                                                 ip]
        L.tm_projection_filter.argtypes = [ctypes.c_int, F32P, F32P, ctypes.c_int] + [ctypes.c_float] * 8 + [F32P]
        L.tm_search_grid.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_float, F32P, ctypes.c_int]
        L.tm_psi_list.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, F32P, ctypes.c_int]
        L._tm_prepare_bound = True
    return L


def _check(rc):
    if rc != 0:
        raise core.TMError(rc, _lib().tm_prepare_last_error().decode())


def _f(a):
    return a.ctypes.data_as(F32P)


PREPARE_EXPORTS = ["tm_prepare_sizes", "tm_prepare_input_image", "tm_prepare_template", "tm_prepare_template_scaled",
                   "tm_projection_filter", "tm_search_grid", "tm_psi_list", "tm_prepare_last_error"]


def prepare_sizes(mx: int, my: int, template_size: int):
    v = [ctypes.c_int(0) for _ in range(5)]
    _check(_lib().tm_prepare_sizes(mx, my, template_size, *[ctypes.byref(x) for x in v]))
    W, H, rot, ncurve, maxpad = (x.value for x in v)
    return W, H, bool(rot), ncurve, maxpad


def prepare_input_image(micrograph: np.ndarray, template_size: int, device: int = 0):
    """Returns (spectrum (H, W/2+1) complex64, curve_x, curve_y, W, H, rotated, max_padding)."""
    m = np.ascontiguousarray(micrograph, np.float32)
    my, mx = m.shape
    W, H, rot, ncurve, maxpad = prepare_sizes(mx, my, template_size)
    spec = np.zeros((H, W // 2 + 1), np.complex64)
    cx = np.zeros(ncurve, np.float32)
    cy = np.zeros(ncurve, np.float32)
    _check(_lib().tm_prepare_input_image(device, _f(m), mx, my, template_size, _f(spec.view(np.float32)), _f(cx),
                                         _f(cy)))
    return spec, cx, cy, W, H, rot, maxpad


def prepare_template(volume: np.ndarray, padding: float, device: int = 0, pixel_size_factor: float = 1.0) -> np.ndarray:
    """(P, P, P/2+1) complex64 template spectrum (match_template.cpp:621-625, 837-845); with
    pixel_size_factor != 1 the padded volume is first resampled by ChangePixelSize (image.cpp:8705-8792)."""
    v = np.ascontiguousarray(volume, np.float32)
    t = v.shape[0]
    P = ctypes.c_int(0)
    L = _lib()
    _check(L.tm_prepare_template(device, _f(v), t, padding, None, ctypes.byref(P)))
    spec = np.zeros((P.value, P.value, P.value // 2 + 1), np.complex64)
    if pixel_size_factor == 1.0:
        _check(L.tm_prepare_template(device, _f(v), t, padding, _f(spec.view(np.float32)), None))
    else:
        _check(L.tm_prepare_template_scaled(device, _f(v), t, padding, pixel_size_factor, _f(spec.view(np.float32)),
                                            None))
    return spec


def projection_filter(T, curve_x, curve_y, voltage_kV, cs_mm, amplitude_contrast, defocus1, defocus2,
                      defocus_angle, pixel_size, phase_shift_deg=0.0) -> np.ndarray:
    out = np.zeros((T, T // 2 + 1), np.complex64)
    _check(_lib().tm_projection_filter(T, _f(curve_x), _f(curve_y), curve_x.size, voltage_kV, cs_mm,
                                       amplitude_contrast, defocus1, defocus2, defocus_angle, pixel_size,
                                       phase_shift_deg, _f(out.view(np.float32))))
    return out


def search_grid(symmetry: str, angular_step: float) -> np.ndarray:
    t, n = symmetry[0].upper().encode(), int(symmetry[1:] or 0)
    cnt = _lib().tm_search_grid(t, n, angular_step, None, 0)
    if cnt < 0:
        _check(cnt)
    out = np.zeros((cnt, 2), np.float32)
    _lib().tm_search_grid(t, n, angular_step, _f(out), cnt)
    return out


def psi_list(psi_start: float, psi_max: float, psi_step: float) -> np.ndarray:
    n = _lib().tm_psi_list(psi_start, psi_max, psi_step, None, 0)
    out = np.zeros(n, np.float32)
    _lib().tm_psi_list(psi_start, psi_max, psi_step, _f(out), n)
    return out


@dataclass
class SearchParams:
    """match_template arguments used by the path (match_template.cpp:346-387)."""
    pixel_size: float = 1.0
    voltage_kV: float = 300.0
    spherical_aberration_mm: float = 2.7
    amplitude_contrast: float = 0.07
    defocus1: float = 15000.0
    defocus2: float = 15000.0
    defocus_angle: float = 0.0
    phase_shift_deg: float = 0.0
    angular_step: float = 10.0
    in_plane_angular_step: float = 10.0
    defocus_search_range: float = 0.0
    defocus_step: float = 0.0
    pixel_size_search_range: float = 0.0
    pixel_size_step: float = 0.0
    padding: float = 1.0
    symmetry: str = "C1"
    first_search_position: int = -1
    last_search_position: int = -1
    batch_size: int = 0
    plane_storage: str = "fp32"  # "fp16"/"bf16": 16-bit inter-pass plane
    column_pass: str = "auto"    # column-transform kernel at image height 4096 (core.COLUMN_PASS)
    orientation_limit: int = 0   # only the first N orientations of the position range (0: all)


class MatchTemplateSearch:
    """Prepared search of one image on one GPU (the local-run path of MatchTemplateApp)."""

    histogram_number_of_points = 512          # match_template.cpp:45-47
    histogram_min, histogram_max = -12.5, 22.5

    def __init__(self, micrograph: np.ndarray, volume: np.ndarray, p: SearchParams, device: int = 0,
                 template_spectrum: np.ndarray | None = None):
        self.p = p
        self.original_shape = micrograph.shape
        T = volume.shape[0]
        self.T = T
        (self.input_spectrum, self.curve_x, self.curve_y, self.W, self.H, self.rotated,
         self.max_padding) = prepare_input_image(micrograph, T, device)
        # a spectrum prepared once (prepare_template) can be shared by the searches of several images
        self.template_spectrum = (prepare_template(volume, p.padding, device) if template_spectrum is None
                                  else template_spectrum)
        self.grid = search_grid(p.symmetry, p.angular_step)
        self.psi = psi_list(0.0, 360.0, p.in_plane_angular_step)
        n = self.grid.shape[0]
        self.first = 0 if p.first_search_position < 0 else p.first_search_position
        self.last = n - 1 if p.last_search_position < 0 else min(p.last_search_position, n - 1)
        self.defocus_step = p.defocus_step if p.defocus_step > 0 else 100.0
        rng = p.defocus_search_range if p.defocus_step > 0 else 0.0
        r = rng / self.defocus_step
        self.n_def_half = int(r + 0.5) if r > 0 else int(r - 0.5)
        # match_template.cpp:755-762: no pixel-size search without a positive step
        self.pixel_size_step = p.pixel_size_step if p.pixel_size_step > 0 else 0.02
        prng = p.pixel_size_search_range if p.pixel_size_step > 0 else 0.0
        r = float(np.float32(prng) / np.float32(self.pixel_size_step))
        self.n_size_half = int(r + 0.5) if r > 0 else int(r - 0.5)
        self.volume = volume
        self.device = device
        sq = math.sqrt(float(self.W * self.H))
        step = (self.histogram_max - self.histogram_min) / float(self.histogram_number_of_points)
        step = float(np.float32(step))
        self.histogram_min_scaled = self.histogram_min / sq
        self.histogram_step_scaled = step / sq
        self.core = core.TemplateMatchingCore(device)
        self.core.init(self.template_spectrum, self.input_spectrum, T, self.grid, self.psi, self.first, self.last,
                       self.histogram_min_scaled, self.histogram_step_scaled, 512, self.max_padding,
                       self.defocus_step, self.pixel_size_step, p.batch_size, p.plane_storage,
                       column_pass=p.column_pass, orientation_limit=p.orientation_limit)

    def set_micrograph(self, micrograph: np.ndarray, on_device: bool = False) -> None:
        """Searches another micrograph of the same size with this handle (tm_set_input): the template,
        grid, work buffers and a captured graph are kept; the maps restart from zero. on_device:
        prepare it on the device (tm_set_micrograph; the micrograph must already be W x H and need
        no rotation); the host copy of the input spectrum is then not kept."""
        if on_device:
            if micrograph.shape != (self.H, self.W) or self.rotated:
                raise ValueError("device preparation needs a micrograph at the search size without rotation")
            self.curve_x, self.curve_y = self.core.set_micrograph(micrograph)
            self.input_spectrum = None
            self.original_shape = micrograph.shape
            return
        spec, cx, cy, W, H, rotated, max_padding = prepare_input_image(micrograph, self.T, self.device)
        if (W, H, rotated) != (self.W, self.H, self.rotated):
            raise ValueError("the micrograph's search size differs from this search's")
        self.input_spectrum, self.curve_x, self.curve_y, self.max_padding = spec, cx, cy, max_padding
        self.original_shape = micrograph.shape
        self.core.set_input(spec)

    def init_params(self, batch_size: int | None = None) -> core.InitParams:
        """The tm_init_params of this search (batch_size: the search's own unless given)."""
        p = self.p
        return core.InitParams(self.template_spectrum.shape[0], self.T, self.W, self.H, self.grid.shape[0],
                               self.psi.size, self.first, self.last, self.histogram_min_scaled,
                               self.histogram_step_scaled, 512, self.max_padding, self.defocus_step,
                               self.pixel_size_step, p.batch_size if batch_size is None else batch_size,
                               core.PLANE_STORAGE[p.plane_storage], p.orientation_limit,
                               core.COLUMN_PASS[p.column_pass], 0)

    def autotune_batch(self, candidates=(16, 32, 64, 128), sample_orientations: int = 1024):
        """Times a sample of this search at each candidate batch depth (tm_autotune_batch) and
        re-initialises the search at the fastest. Returns (batch size, {candidate: us per CCF})."""
        best, us = core.autotune_batch(self.init_params(), self.template_spectrum, self.input_spectrum, self.grid,
                                       self.psi, self.filter_for(0), candidates, sample_orientations, self.device)
        self.p = dataclasses.replace(self.p, batch_size=best)
        self.core.init(self.template_spectrum, self.input_spectrum, self.T, self.grid, self.psi, self.first, self.last,
                       self.histogram_min_scaled, self.histogram_step_scaled, 512, self.max_padding,
                       self.defocus_step, self.pixel_size_step, best, self.p.plane_storage,
                       column_pass=self.p.column_pass, orientation_limit=self.p.orientation_limit)
        return best, us

    def autotune(self, batches=(32, 64, 96), columns=None, sample_orientations: int = 1024):
        """Times a sample of this search at each (column pass, batch depth) (tm_autotune) and
        re-initialises the search at the fastest. columns default: the three column passes at image
        height 4096, the compiled and the narrow column tiles elsewhere. Returns (batch, column pass,
        table of us per CCF)."""
        if columns is None:
            columns = ("pairs", "persistent", "persistent4") if self.H == 4096 else ("auto", "narrow")
        best_b, best_c, table = core.autotune(self.init_params(), self.template_spectrum, self.input_spectrum,
                                              self.grid, self.psi, self.filter_for(0), batches, columns,
                                              sample_orientations, self.device)
        self.p = dataclasses.replace(self.p, batch_size=best_b, column_pass=best_c)
        self.core.init(self.template_spectrum, self.input_spectrum, self.T, self.grid, self.psi, self.first, self.last,
                       self.histogram_min_scaled, self.histogram_step_scaled, 512, self.max_padding,
                       self.defocus_step, self.pixel_size_step, best_b, self.p.plane_storage, column_pass=best_c,
                       orientation_limit=self.p.orientation_limit)
        return best_b, best_c, table

    def filter_for(self, defocus_i: int) -> np.ndarray:
        p = self.p
        d1 = np.float32(p.defocus1) + np.float32(defocus_i) * np.float32(self.defocus_step)
        d2 = np.float32(p.defocus2) + np.float32(defocus_i) * np.float32(self.defocus_step)
        return projection_filter(self.T, self.curve_x, self.curve_y, p.voltage_kV, p.spherical_aberration_mm,
                                 p.amplitude_contrast, float(d1), float(d2), p.defocus_angle, p.pixel_size,
                                 p.phase_shift_deg)

    def size_factor(self, size_i: int) -> float:
        """(pixel_size + size_i * pixel_size_step) / pixel_size in float (match_template.cpp:842)."""
        f32 = np.float32
        return float((f32(self.p.pixel_size) + f32(size_i) * f32(self.pixel_size_step)) / f32(self.p.pixel_size))

    def run(self, progress=None) -> int:
        """The pixel-size and defocus loops of match_template.cpp:837-914: each pixel-size plane's
        template replaces the previous one in the same core, whose maps keep accumulating."""
        n = 0
        for si in range(-self.n_size_half, self.n_size_half + 1):
            if self.n_size_half:
                self.core.set_template(prepare_template(self.volume, self.p.padding, self.device, self.size_factor(si)))
            for di in range(-self.n_def_half, self.n_def_half + 1):
                n += self.core.run_inner_loop(self.filter_for(di), si, di, progress)
        return n

    def worker_record(self, particle_radius_angstroms: float = 0.0, n_ccfs: float | None = None) -> np.ndarray:
        """This search's maps as the record a worker sends to the master (tm_result_pack,
        match_template.cpp:1371-1531): turned back to the micrograph's orientation
        (Rotate2DInPlaceBy90Degrees(false), image.cpp:11127-11138), cut to the valid area and padded to
        the micrograph's size with each map's central average. remove_npix_from_edge is
        myroundint(particle radius / pixel size) in float (match_template.cpp:494); n_ccfs defaults to
        the CCFs this handle computed."""
        r = self.core.results()
        maps = {k: getattr(r, k).astype(np.float32) for k in results.MAP_NAMES}
        if self.rotated:
            maps = {k: np.ascontiguousarray(np.rot90(v, 1)) for k, v in maps.items()}
        q = float(np.float32(particle_radius_angstroms) / np.float32(self.p.pixel_size))
        remove = int(q + 0.5) if q > 0 else int(q - 0.5)
        oy, ox = self.original_shape
        return results.pack(maps, ox, oy, remove, r.hist, float(r.nccf if n_ccfs is None else n_ccfs),
                            math.sqrt(float(self.W * self.H)))

    @property
    def orientations_per_plane(self) -> int:
        n = (self.last - self.first + 1) * self.psi.size
        return min(n, self.p.orientation_limit) if self.p.orientation_limit > 0 else n

    @property
    def total_correlation_positions(self) -> int:
        return self.orientations_per_plane * (2 * self.n_def_half + 1)


def run_concurrent(searches: list, workers: int = 2) -> list:
    """Runs prepared searches (one per image, each with its own GPU stream) on `workers` host threads,
    so the kernels of different images overlap on the GPU where one image alone leaves it partly
    idle. The reference runs one image per process; the per-image results do not depend on the
    schedule (the kernels are deterministic). Returns the correlations computed per search."""
    from concurrent.futures import ThreadPoolExecutor

    if workers <= 1:
        return [s.run() for s in searches]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda s: s.run(), searches))


def run_group(searches: list, group: "core.TemplateMatchingGroup | None" = None, lanes: int = 0) -> list:
    """Runs prepared searches (one per image, any compiled sizes, on one device) with every defocus
    plane of all of them as ONE hipGraph launch (tm_group_run: the images dealt to `lanes` chains,
    0: two, each chain an independent branch of the graph); the
    per-image results equal those of each search's own run(), bit for bit. The searches must have the
    same number of defocus planes and no pixel-size search. group: a TemplateMatchingGroup over the
    searches' cores to reuse (its graph is captured once and replayed). Returns the correlations
    computed per search."""
    if any(m.n_size_half for m in searches):
        raise ValueError("run_group takes searches without a pixel-size search")
    nd = {m.n_def_half for m in searches}
    if len(nd) != 1:
        raise ValueError("run_group takes searches with the same number of defocus planes")
    n_def_half = nd.pop()
    own = group is None
    g = core.TemplateMatchingGroup([m.core for m in searches], lanes) if own else group
    try:
        done = [0] * len(searches)
        for di in range(-n_def_half, n_def_half + 1):
            n = g.run([m.filter_for(di) for m in searches], [0] * len(searches), [di] * len(searches))
            done = [a + b for a, b in zip(done, n)]
        return done
    finally:
        if own:
            g.close()
