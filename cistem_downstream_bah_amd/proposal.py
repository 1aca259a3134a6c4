"""Learned proposals for template matching: a small network (cnn.ProposalNet, hand-written HIP
kernels) reads the micrograph and proposes particle locations and coarse orientation bins; a local
exhaustive refinement scores every orientation of the search grid inside the proposed bins at the
proposed windows (tm_refine_score_batch: TemplateScore of refine_template.cpp:25-80 in batches).
The exhaustive search (match_template) labels the training micrographs and is the judge of recall.
None of this has a counterpart in the reference, which always runs the exhaustive search.

Orientation bins: 60 view directions (a Fibonacci sphere; a grid orientation belongs to the
direction nearest its (phi, theta) view vector) x 10 in-plane bins of 36 degrees = 600 bins.
Locations: one cell of the network's output grid (input_size / grid pixels square) per proposal.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import cnn, peaks as PK, synth
from . import match_template as MT
from .refine import RefineSettings, TemplateRefiner

N_DIRECTIONS = 60
N_PSI_BINS = 10
N_BINS = N_DIRECTIONS * N_PSI_BINS


def fibonacci_directions(n: int = N_DIRECTIONS) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    r = np.sqrt(1.0 - z * z)
    a = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(a), r * np.sin(a), z], axis=1)


def view_vector(phi, theta) -> np.ndarray:
    p, t = np.radians(np.asarray(phi, np.float64)), np.radians(np.asarray(theta, np.float64))
    return np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], axis=-1)


def orientation_bin(phi, theta, psi, dirs: np.ndarray | None = None) -> np.ndarray:
    dirs = fibonacci_directions() if dirs is None else dirs
    d = np.argmax(view_vector(phi, theta) @ dirs.T, axis=-1)
    q = np.floor(np.mod(np.asarray(psi, np.float64), 360.0) / (360.0 / N_PSI_BINS)).astype(np.int64)
    return d * N_PSI_BINS + np.minimum(q, N_PSI_BINS - 1)


class OrientationTable:
    """The search grid's orientations (phi, theta, psi), their bins, the grid orientations of each bin,
    and the coarse grid (twice the angular steps) with its bins."""

    def __init__(self, angular_step: float, in_plane_angular_step: float, symmetry: str = "C1"):
        grid = MT.search_grid(symmetry, angular_step)
        psi = MT.psi_list(0.0, 360.0, in_plane_angular_step)
        self.orients = np.array([(g[0], g[1], s) for g in grid for s in psi], np.float32)
        self.dirs = fibonacci_directions()
        self.bin = orientation_bin(self.orients[:, 0], self.orients[:, 1], self.orients[:, 2], self.dirs)
        self.members = [np.nonzero(self.bin == b)[0] for b in range(N_BINS)]
        cgrid = MT.search_grid(symmetry, 2 * angular_step)
        cpsi = MT.psi_list(0.0, 360.0, 2 * in_plane_angular_step)
        self.coarse = np.array([(g[0], g[1], s) for g in cgrid for s in cpsi], np.float32)
        self.coarse_bin = orientation_bin(self.coarse[:, 0], self.coarse[:, 1], self.coarse[:, 2], self.dirs)


def targets(peaks: list, grid: int, size: int, dirs: np.ndarray | None = None):
    """Location (1 at each peak's cell) and orientation-bin (-1 elsewhere) targets of one micrograph."""
    cell = size // grid
    heat = np.zeros((grid, grid), np.float32)
    bins = -np.ones((grid, grid), np.int32)
    for p in peaks:
        cy, cx = int(p["y"]) // cell, int(p["x"]) // cell
        heat[cy, cx] = 1.0
        bins[cy, cx] = int(orientation_bin(p["phi"], p["theta"], p["psi"], dirs))
    return heat, bins


def propose_cells(heat: np.ndarray, k: int, min_cells: int = 3) -> list:
    """Top-k cells by location logit (heat: G x G), skipping cells within min_cells (Chebyshev) of a
    taken one: [(cell y, cell x)]."""
    G = heat.shape[0]
    order = np.argsort(-heat.ravel(), kind="stable")
    taken = []
    for c in order:
        cy, cx = divmod(int(c), G)
        if all(max(abs(cy - y), abs(cx - x)) >= min_cells for y, x in taken):
            taken.append((cy, cx))
            if len(taken) == k:
                break
    return taken


def propose(logits: np.ndarray, k: int, m: int, min_cells: int = 3):
    """Top-k cells by location logit, skipping cells within min_cells (Chebyshev) of a taken one,
    each with its top-m orientation bins: [((cell y, cell x), bins)]."""
    cells = propose_cells(logits[..., 0], k, min_cells)
    return [((cy, cx), np.argsort(-logits[cy, cx, 1:], kind="stable")[:m]) for cy, cx in cells]


@dataclass
class Labelled:
    micrograph: np.ndarray
    peaks: list                 # exhaustive-search peaks: dicts x, y (pixels), z, phi, theta, psi
    seconds: float = 0.0        # exhaustive search time for this micrograph
    seed: int = 0


@dataclass
class Detection:
    x: float
    y: float
    value: float
    phi: float
    theta: float
    psi: float
    cell: tuple = (0, 0)


@dataclass
class StudyConfig:
    size: int = 512
    template_size: int = 64
    n_atoms: int = 200
    atom_radius: float = 24.0
    copies: tuple = (6, 14)
    snr: tuple = (0.15, 0.35)
    defocus: float = 15000.0
    angular_step: float = 10.0
    in_plane_angular_step: float = 10.0
    threshold: float = 0.0      # z-score of the peaks the proposals must recall; 0: the search's
                                # expected threshold sqrt(2) erfcinv(2 / (pixels x CCFs)), as
                                # match_template reports peaks (about 6.2 here)
    min_peak_radius: int = 10
    template_seed: int = 0


class ProposalStudy:
    """One template, micrographs drawn from seeds, the exhaustive search as labeller and judge."""

    def __init__(self, cfg: StudyConfig = StudyConfig(), device: int = 0):
        self.cfg = cfg
        self.device = device
        rng = np.random.default_rng(cfg.template_seed)
        pos, amp = synth.make_atoms(rng, cfg.n_atoms, cfg.atom_radius)
        self.template = (pos, amp, synth.render_volume(cfg.template_size, pos, amp))
        self.volume = self.template[2]
        self.params = MT.SearchParams(angular_step=cfg.angular_step, in_plane_angular_step=cfg.in_plane_angular_step,
                                      padding=1.0, defocus1=cfg.defocus, defocus2=cfg.defocus)
        self.tpl_spectrum = MT.prepare_template(self.volume, 1.0, device)
        tab = OrientationTable(cfg.angular_step, cfg.in_plane_angular_step)
        self.orients, self.dirs, self.members = tab.orients, tab.dirs, tab.members
        self.coarse, self.coarse_bin = tab.coarse, tab.coarse_bin
        self._refiner = None  # one refinement handle, its image replaced per micrograph
        self._exhaustive_search = None  # one search handle, its input replaced per micrograph

    # ---------------------------------------------------------------- data
    def micrograph(self, seed: int) -> np.ndarray:
        c = self.cfg
        rng = np.random.default_rng(10_000 + seed)
        n = int(rng.integers(c.copies[0], c.copies[1] + 1))
        snr = float(rng.uniform(*c.snr))
        s = synth.make_synthetic(seed, c.size, c.size, c.template_size, c.n_atoms, c.atom_radius, n, snr, c.defocus,
                                 min_dist=c.template_size, template=self.template)
        return s.micrograph

    def _search(self, micrograph: np.ndarray):
        return MT.MatchTemplateSearch(micrograph, self.volume, self.params, self.device,
                                      template_spectrum=self.tpl_spectrum)

    def _peaks(self, search, n: int) -> list:
        """The search's peaks above the threshold, found on the device (tm_find_peaks; the
        micrographs here need no padding or rotation, so the search geometry is the image's)."""
        thr = self.cfg.threshold if self.cfg.threshold > 0 else PK.expected_threshold(search.W, search.H, n)
        pk = search.core.find_peaks(n, thr, self.cfg.template_size // 4 + 1, float(self.cfg.min_peak_radius))
        return [dict(x=p["x"], y=p["y"], z=p["peak_height"], phi=p["phi"], theta=p["theta"], psi=p["psi"])
                for p in pk]

    def exhaustive(self, micrograph: np.ndarray):
        """The full search: its peaks above the threshold (with their best angles) and the time
        taken (preparation of the image, every orientation, z-score and peak search). One handle
        serves every micrograph (tm_set_input), replaying its plane from a captured hipGraph."""
        t0 = time.perf_counter()
        search = self._exhaustive_search
        if search is None:
            search = self._exhaustive_search = self._search(micrograph)
            search.core.set_graph(True)
        else:
            search.set_micrograph(micrograph, on_device=True)
        t1 = time.perf_counter()
        n = search.run()
        t2 = time.perf_counter()
        pk = self._peaks(search, n)
        t3 = time.perf_counter()
        self.last_exhaustive_times = {"image_preparation": t1 - t0, "search": t2 - t1, "z_score_and_peaks": t3 - t2}
        return pk, t3 - t0

    def labelled(self, seeds, workers: int = 4, log=None) -> list:
        """Micrographs of the seeds labelled by the exhaustive search, `workers` searches at a time
        on their own streams (seconds: wall time per micrograph at that concurrency)."""
        seeds = [int(s) for s in seeds]
        out = []
        for c0 in range(0, len(seeds), workers):
            chunk = seeds[c0:c0 + workers]
            ms = [self.micrograph(s) for s in chunk]
            t0 = time.perf_counter()
            searches = [self._search(m) for m in ms]
            ns = MT.run_concurrent(searches, workers)
            pks = [self._peaks(se, n) for se, n in zip(searches, ns)]
            dt = (time.perf_counter() - t0) / len(chunk)
            for se in searches:
                se.core.close()
            out += [Labelled(m, pk, dt, s) for m, pk, s in zip(ms, pks, chunk)]
            if log and (len(out) % 200 < workers or len(out) == len(seeds)):
                log(f"labelled {len(out)} / {len(seeds)} micrographs")
        return out

    def targets(self, peaks: list, grid: int):
        return targets(peaks, grid, self.cfg.size, self.dirs)

    # ---------------------------------------------------------------- training
    def train(self, net: cnn.ProposalNet, data: list, epochs: int, batch: int = 16, lr: float = 1e-3,
              pos_weight: float = 30.0, seed: int = 0, log=None) -> list:
        rng = np.random.default_rng(seed)
        tg = [self.targets(d.peaks, net.grid) for d in data]
        hist = []
        for ep in range(epochs):
            order = rng.permutation(len(data))
            tot, nb = np.zeros(2), 0
            for b0 in range(0, len(order) - batch + 1, batch):
                idx = order[b0:b0 + batch]
                net.forward(np.stack([data[i].micrograph for i in idx]), want_logits=False)
                (lh, lb), _ = net.backward(np.stack([tg[i][0] for i in idx]), np.stack([tg[i][1] for i in idx]),
                                           pos_weight)
                net.adam_step(lr=lr)
                tot += (lh, lb)
                nb += 1
            hist.append((tot / max(nb, 1)).tolist())
            if log:
                log(f"epoch {ep}: location loss {hist[-1][0]:.4f}, orientation loss {hist[-1][1]:.4f}")
        return hist

    # ---------------------------------------------------------------- inference
    def propose(self, logits: np.ndarray, k: int, m: int, min_cells: int = 3):
        return propose(logits, k, m, min_cells)

    def _windows(self, proposals: list, grid: int):
        cell = self.cfg.size // grid
        half = self.cfg.size // 2
        return [(cx * cell + cell // 2 - half, cy * cell + cell // 2 - half) for (cy, cx), _ in proposals]

    def coarse_bins(self, refiner: TemplateRefiner, proposals: list, grid: int, m: int) -> list:
        """Bins ranked by a coarse local search: every orientation of the coarse grid (twice the
        search's angular steps) scored at each proposal's window; the bins of the best-scoring coarse
        orientations, best first, up to m per proposal."""
        wins = self._windows(proposals, grid)
        nc = self.coarse.shape[0]
        xy = np.repeat(np.array(wins, np.int32), nc, axis=0)
        ang = np.tile(self.coarse, (len(wins), 1))
        res = refiner.score_batch(xy, ang)
        out = []
        for q in range(len(wins)):
            v = res[q * nc:(q + 1) * nc, 2]
            order = np.argsort(-v, kind="stable")
            bins = []
            for j in order:
                b = int(self.coarse_bin[j])
                if b not in bins:
                    bins.append(b)
                    if len(bins) == m:
                        break
            out.append(np.array(bins, np.int64))
        return out, int(xy.shape[0])

    def refine(self, refiner: TemplateRefiner, proposals: list, grid: int) -> list:
        """Local exhaustive refinement: every grid orientation of each proposal's bins, scored at the
        window centred on the proposal's cell; the best per proposal is its detection."""
        half = self.cfg.size // 2
        wins = self._windows(proposals, grid)
        xy, ang, owner = [], [], []
        for q, (_, bins) in enumerate(proposals):
            idx = np.concatenate([self.members[b] for b in bins]) if len(bins) else np.zeros(0, np.int64)
            xy.append(np.tile(wins[q], (idx.size, 1)))
            ang.append(self.orients[idx])
            owner.append(np.full(idx.size, q))
        xy, ang, owner = np.concatenate(xy), np.concatenate(ang), np.concatenate(owner)
        res = refiner.score_batch(xy, ang)
        dets = []
        for q, ((cy, cx), _) in enumerate(proposals):
            sel = np.nonzero(owner == q)[0]
            if sel.size == 0:
                continue
            j = sel[np.argmax(res[sel, 2])]
            dets.append(Detection(float(xy[j, 0] + res[j, 0] + half), float(xy[j, 1] + res[j, 1] + half),
                                  float(res[j, 2]), *map(float, ang[j]), (cy, cx)))
        return dets, int(xy.shape[0])

    def detect(self, net: cnn.ProposalNet, micrograph: np.ndarray, k: int, m: int, proposals=None,
               bins: str = "network"):
        """Proposals + local refinement for one micrograph: (detections, TemplateScore calls, seconds).
        Locations: the network's top-k cells (or the given proposals). Orientation bins per location:
        "network" (its top-m bins), "coarse" (top-m bins of a coarse local search) or "union"."""
        t0 = time.perf_counter()
        if proposals is None:
            # the location logits, then the orientation logits of the chosen cells only
            net.forward(micrograph[None], want_logits=False)
            cells = propose_cells(net.heat()[0], k)
            G = net.grid
            lg = net.cell_logits([cy * G + cx for cy, cx in cells]) if cells else np.zeros((0, 1 + N_BINS))
            proposals = [(c, np.argsort(-lg[q, 1:], kind="stable")[:m]) for q, c in enumerate(cells)]
        t1 = time.perf_counter()
        refiner = self._refiner
        if refiner is None:
            refiner = self._refiner = TemplateRefiner(micrograph, self.volume,
                                                      RefineSettings(defocus1=self.cfg.defocus, defocus2=self.cfg.defocus,
                                                                     padding=1.0, max_in_flight=2048),
                                                      template_spectrum=self.tpl_spectrum, device=self.device)
        else:
            refiner.set_micrograph(micrograph)  # prepared on the device
        t2 = time.perf_counter()
        n0 = 0
        if bins in ("coarse", "union"):
            cb, n0 = self.coarse_bins(refiner, proposals, net.grid, m)
            if bins == "coarse":
                proposals = [(c, b) for (c, _), b in zip(proposals, cb)]
            else:
                proposals = [(c, np.unique(np.concatenate([nb, b]))) for (c, nb), b in zip(proposals, cb)]
        dets, n = self.refine(refiner, proposals, net.grid)
        t3 = time.perf_counter()
        self.last_times = {"network_and_proposals": t1 - t0, "image_preparation": t2 - t1, "scores": t3 - t2}
        return dets, n + n0, t3 - t0

    def random_proposals(self, rng, grid: int, k: int, m: int):
        cells = rng.choice(grid * grid, size=k, replace=False)
        return [((int(c) // grid, int(c) % grid), rng.choice(N_BINS, size=m, replace=False)) for c in cells]


def match(dets: list, peaks: list, radius: float = 2.0) -> list:
    """Per exhaustive peak: (found, same orientation), found if a detection lies within radius
    pixels (the nearest one is compared for the orientation)."""
    out = []
    for p in peaks:
        best = None
        for d in dets:
            r2 = (d.x - p["x"]) ** 2 + (d.y - p["y"]) ** 2
            if r2 <= radius * radius and (best is None or r2 < best[0]):
                best = (r2, d)
        if best is None:
            out.append((False, False))
        else:
            d = best[1]
            out.append((True, abs(d.phi - p["phi"]) < 1e-3 and abs(d.theta - p["theta"]) < 1e-3
                        and abs(d.psi - p["psi"]) < 1e-3))
    return out


def recall(dets: list, peaks: list, radius: float = 2.0):
    """Fraction of exhaustive peaks with a detection within radius pixels, and the fraction of those
    whose detection has the peak's orientation."""
    if not peaks:
        return 1.0, 1.0, 0
    m = match(dets, peaks, radius)
    hit = sum(f for f, _ in m)
    same = sum(s for f, s in m if f)
    return hit / len(peaks), (same / hit if hit else 0.0), len(peaks)
