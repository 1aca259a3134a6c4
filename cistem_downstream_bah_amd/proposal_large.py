"""Learned proposals at the scale of the 4096^2 search: 4096^2 micrographs of 300 copies of one 256^3
template at SNR 0.05, judged by the truncated exhaustive search of the benchmark (the first 200,000
orientations of the 2.5 / 1.5 deg C1 grid in reference order, 11 defocus planes from -1000 to +1000 A,
z-score peaks at or above 7.5). None of this has a counterpart in the reference, which always runs
the exhaustive search (match_template.cpp:996-1169).

* network input: the micrograph binned 8 x 8 (block mean) to 512^2; output grid 64^2 cells of 64
  pixels of the micrograph, a location logit and the 600 orientation bins of proposal.py per cell
* labels: the exhaustive search run on the micrograph binned 4 x 4 (1024^2, 4 A per pixel) with the
  template binned alike (64^3, padded to 128^3), over the positions of the 5 deg grid up to theta 50 deg
  (the judge's first 200,000 orientations reach theta 42.5 deg) at 5 deg in-plane, one defocus plane;
  its peaks above its own expected threshold, positions scaled back to the micrograph. A full-scale
  search per training micrograph would cost 160 s; this one costs well under a second
* proposals: the K cells with the highest location logits (at least two cells apart), each with its
  m best orientation bins among those the judge's orientations reach
* local exhaustive refinement: TemplateScore (refine_template.cpp:25-80, tm_refine_score_batch) of
  every judge orientation in the proposed bins, in the 256 x 256 window centred on the cell, at the
  micrograph's defocus; the best score per cell is its detection
* the judge: the truncated exhaustive search itself, its peaks at z >= 7.5; recall is the fraction of
  those peaks with a detection within r pixels
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import cnn, peaks as PK, proposal as PR, synth
from . import match_template as MT
from .refine import RefineSettings, TemplateRefiner


@dataclass
class LargeConfig:
    size: int = 4096
    template_size: int = 256
    n_atoms: int = 3000
    atom_radius: float = 100.0
    copies: int = 300
    snr: float = 0.05
    defocus: float = 15000.0
    min_dist: float = 200.0
    template_seed: int = 1000          # the 4096^2 benchmark's template
    grid: int = 64                     # network output cells per side
    net_bin: int = 8                   # network input = size / net_bin
    label_bin: int = 4
    label_step: float = 5.0
    label_psi_step: float = 5.0
    label_theta_max: float = 50.0
    label_threshold: float = 0.0       # 0: the label search's expected threshold
    judge_step: float = 2.5
    judge_psi_step: float = 1.5
    judge_orientations: int = 200_000
    judge_defocus_range: float = 1000.0
    judge_defocus_step: float = 200.0
    judge_threshold: float = 7.5


def bin_image(a: np.ndarray, k: int) -> np.ndarray:
    h, w = a.shape
    return a.reshape(h // k, k, w // k, k).mean(axis=(1, 3), dtype=np.float64).astype(np.float32)


def bin_volume(v: np.ndarray, k: int) -> np.ndarray:
    n = v.shape[0] // k
    return v.reshape(n, k, n, k, n, k).mean(axis=(1, 3, 5), dtype=np.float64).astype(np.float32)


def rotation_quaternions(phi, theta, psi) -> np.ndarray:
    """Unit quaternions (w, x, y, z) of synth.euler_matrix's rotations Rz(phi) Ry(theta) Rz(psi)
    (degrees), shape (n, 4): the rotation angle between two orientations is 2 acos |q1 . q2|."""
    a, b, c = (np.radians(np.asarray(v, np.float64)).ravel() / 2.0 for v in (phi, theta, psi))
    w1, x1, y1, z1 = np.cos(a) * np.cos(b), -np.sin(a) * np.sin(b), np.cos(a) * np.sin(b), np.sin(a) * np.cos(b)
    cc, sc = np.cos(c), np.sin(c)
    return np.stack([w1 * cc - z1 * sc, x1 * cc + y1 * sc, y1 * cc - x1 * sc, w1 * sc + z1 * cc], axis=1)


def default_net() -> cnn.NetSpec:
    """Fourteen 3x3 convolutions (256^2 x 32, 128^2 x 64 and 64^2 x 128..160 channels; receptive field
    163 input pixels, 1304 micrograph pixels), 1.57 M parameters."""
    return cnn.NetSpec(input_size=512, channels=[32, 32, 64, 64, 64, 128, 128, 128, 128, 128, 128, 128, 160, 160],
                       strides=[2, 1, 2, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1], n_bins=PR.N_BINS)


class JudgeOrientations:
    """The judge's orientations (the first judge_orientations of the C1 grid in reference order), their
    orientation bins and rotation quaternions, and the judge orientations near given angles."""

    def __init__(self, cfg: LargeConfig):
        c = cfg
        grid = MT.search_grid("C1", c.judge_step)
        psi = MT.psi_list(0.0, 360.0, c.judge_psi_step)
        npos = int(math.ceil(c.judge_orientations / psi.size))
        o = np.array([(g[0], g[1], s) for g in grid[:npos] for s in psi], np.float32)[: c.judge_orientations]
        self.orients = o
        self.judge_last_position = npos - 1
        self.n_psi = psi.size
        self.dirs = PR.fibonacci_directions()
        self.bin = PR.orientation_bin(o[:, 0], o[:, 1], o[:, 2], self.dirs)
        self.members = [np.nonzero(self.bin == b)[0] for b in range(PR.N_BINS)]
        self.reachable = np.array([m.size > 0 for m in self.members])
        self.quats = rotation_quaternions(o[:, 0], o[:, 1], o[:, 2])
        self.pos_views = PR.view_vector(grid[:npos, 0], grid[:npos, 1])  # R e_z of each position

    def orientation_set(self, angles: np.ndarray, radius_deg: float) -> np.ndarray:
        """Indices of the judge orientations within radius_deg (rotation angle) of any of the given
        (phi, theta, psi) rows. A rotation by alpha moves every direction by at most alpha, so only the
        positions whose view direction lies within radius_deg of a given one's are examined."""
        angles = np.asarray(angles, np.float64).reshape(-1, 3)
        qc = rotation_quaternions(angles[:, 0], angles[:, 1], angles[:, 2])
        vc = PR.view_vector(angles[:, 0], angles[:, 1]).reshape(-1, 3)
        pos = np.nonzero((self.pos_views @ vc.T).max(axis=1) >= math.cos(math.radians(radius_deg)) - 1e-12)[0]
        if pos.size == 0:
            return np.zeros(0, np.int64)
        idx = (pos[:, None] * self.n_psi + np.arange(self.n_psi)[None, :]).ravel()
        idx = idx[idx < self.orients.shape[0]]
        keep = np.abs(self.quats[idx] @ qc.T).max(axis=1) >= math.cos(math.radians(radius_deg) / 2.0)
        return idx[keep]

    def judge_index(self, peak: dict) -> int:
        """Index of a peak's orientation among the judge's orientations (-1 if absent)."""
        o = self.orients
        hit = np.nonzero((o[:, 0] == np.float32(peak["phi"])) & (o[:, 1] == np.float32(peak["theta"]))
                         & (o[:, 2] == np.float32(peak["psi"])))[0]
        return int(hit[0]) if hit.size else -1


class LargeStudy(JudgeOrientations):
    """One template; micrographs from seeds; the binned search labels, the exhaustive search judges."""

    def __init__(self, cfg: LargeConfig = LargeConfig(), device: int = 0):
        super().__init__(cfg)
        self.cfg = c = cfg
        self.device = device
        rng = np.random.default_rng(c.template_seed)
        pos, amp = synth.make_atoms(rng, c.n_atoms, c.atom_radius)
        self.template = (pos, amp, synth.render_volume(c.template_size, pos, amp))
        self.volume = self.template[2]
        self.judge_spectrum = MT.prepare_template(self.volume, 2.0, device)
        self.label_volume = bin_volume(self.volume, c.label_bin)
        self.label_spectrum = MT.prepare_template(self.label_volume, 2.0, device)
        lgrid = MT.search_grid("C1", c.label_step)
        self.label_last_position = int(np.nonzero(lgrid[:, 1] <= c.label_theta_max + 1e-3)[0].max())
        self._label_search = None
        self._judge_search = None
        self._refiner = None

    # ------------------------------------------------------------------ data
    def micrograph(self, seed: int, timing: dict | None = None) -> synth.Synthetic:
        c = self.cfg
        return synth.make_synthetic_gpu(seed, c.size, c.size, c.template_size, c.n_atoms, c.atom_radius, c.copies,
                                        c.snr, c.defocus, min_dist=c.min_dist, template=self.template, timing=timing)

    def net_input(self, micrograph: np.ndarray) -> np.ndarray:
        return bin_image(micrograph, self.cfg.net_bin)

    def binned_search(self, micrograph: np.ndarray):
        """The exhaustive search of the micrograph binned label_bin x label_bin with the binned template
        (positions up to label_theta_max at label_step, psi at label_psi_step, one defocus plane):
        (search handle, CCFs). One handle serves every micrograph."""
        c = self.cfg
        b = bin_image(micrograph, c.label_bin)
        se = self._label_search
        if se is None:
            sp = MT.SearchParams(pixel_size=float(c.label_bin), angular_step=c.label_step,
                                 in_plane_angular_step=c.label_psi_step, padding=2.0, defocus1=c.defocus,
                                 defocus2=c.defocus, first_search_position=0,
                                 last_search_position=self.label_last_position)
            se = self._label_search = MT.MatchTemplateSearch(b, self.label_volume, sp, self.device,
                                                             template_spectrum=self.label_spectrum)
        else:
            se.set_micrograph(b, on_device=True)
        return se, se.run()

    def label(self, micrograph: np.ndarray) -> list:
        """Peaks of the binned search above its expected threshold (or label_threshold), in
        micrograph coordinates, with their best angles."""
        c = self.cfg
        se, n = self.binned_search(micrograph)
        thr = c.label_threshold if c.label_threshold > 0 else PK.expected_threshold(se.W, se.H, n)
        pk = se.core.find_peaks(n, thr, se.T // 4 + 1, 10.0)
        off = (c.label_bin - 1) / 2.0
        return [dict(x=p["x"] * c.label_bin + off, y=p["y"] * c.label_bin + off, z=p["peak_height"], phi=p["phi"],
                     theta=p["theta"], psi=p["psi"]) for p in pk]

    def targets(self, peaks: list):
        return PR.targets(peaks, self.cfg.grid, self.cfg.size, self.dirs)

    # ------------------------------------------------------------------ training
    def train(self, net: cnn.ProposalNet, inputs: list, labels: list, epochs: int, batch: int = 16, lr: float = 1e-3,
              pos_weight: float = 30.0, seed: int = 0, augment: bool = True, log=None) -> list:
        """Adam on the location and orientation-bin losses. augment: each sample is turned by one of the
        eight symmetries of the square (drawn per sample and epoch); a turned sample keeps its location
        targets (turned alike) and drops its orientation targets, which a mirror or turn of the image
        would change."""
        rng = np.random.default_rng(seed)
        tg = [self.targets(p) for p in labels]
        hist = []

        def turn(a, t):
            a = np.rot90(a, t % 4)
            return np.ascontiguousarray(a[:, ::-1] if t >= 4 else a)

        for ep in range(epochs):
            order = rng.permutation(len(inputs))
            tot, nb = np.zeros(2), 0
            for b0 in range(0, len(order) - batch + 1, batch):
                idx = order[b0:b0 + batch]
                ts = rng.integers(0, 8, idx.size) if augment else np.zeros(idx.size, np.int64)
                net.forward(np.stack([turn(inputs[i], t) for i, t in zip(idx, ts)]), want_logits=False)
                heat = np.stack([turn(tg[i][0], t) for i, t in zip(idx, ts)])
                bins = np.stack([tg[i][1] if t == 0 else np.full_like(tg[i][1], -1) for i, t in zip(idx, ts)])
                (lh, lb), _ = net.backward(heat, bins, pos_weight)
                net.adam_step(lr=lr)
                tot += (lh, lb)
                nb += 1
            hist.append((tot / max(nb, 1)).tolist())
            if log:
                log(f"epoch {ep}: location loss {hist[-1][0]:.4f}, orientation loss {hist[-1][1]:.4f}")
        return hist

    # ------------------------------------------------------------------ the judge
    def judge(self, micrograph: np.ndarray, log=None):
        """The truncated exhaustive search from the raw micrograph to its z >= 7.5 peaks:
        (peaks, seconds, CCFs)."""
        c = self.cfg
        se = self._judge_search
        t0 = time.perf_counter()
        if se is None:
            sp = MT.SearchParams(angular_step=c.judge_step, in_plane_angular_step=c.judge_psi_step, padding=2.0,
                                 defocus1=c.defocus, defocus2=c.defocus, defocus_search_range=c.judge_defocus_range,
                                 defocus_step=c.judge_defocus_step, first_search_position=0,
                                 last_search_position=self.judge_last_position,
                                 orientation_limit=c.judge_orientations)
            se = self._judge_search = MT.MatchTemplateSearch(micrograph, self.volume, sp, self.device,
                                                             template_spectrum=self.judge_spectrum)
            t0 = time.perf_counter()  # the first handle's one-off set-up is not part of a micrograph's time
        se.set_micrograph(micrograph, on_device=True)
        n = 0
        for d in range(-se.n_def_half, se.n_def_half + 1):
            n += se.core.run_inner_loop(se.filter_for(d), 0, d)
            if log:
                log(f"judge plane {d:+d}: {n} CCFs, {time.perf_counter() - t0:.1f}s")
        pk = se.core.find_peaks(n, c.judge_threshold, se.T // 4 + 1, 10.0)
        dt = time.perf_counter() - t0
        return [dict(x=p["x"], y=p["y"], z=p["peak_height"], phi=p["phi"], theta=p["theta"], psi=p["psi"])
                for p in pk], dt, n

    # ------------------------------------------------------------------ the learned path
    def detect(self, net: cnn.ProposalNet, micrograph: np.ndarray, k: int, m: int, min_cells: int = 2,
               cells: list | None = None):
        """Proposals + local exhaustive refinement for one micrograph: (detections, TemplateScores,
        seconds, {stage: seconds}). cells: score these cells [(cell y, cell x)] instead of the
        network's top k (their orientation bins still come from the network)."""
        c = self.cfg
        t0 = time.perf_counter()
        net.forward(self.net_input(micrograph)[None], want_logits=False)
        if cells is None:
            cells = PR.propose_cells(net.heat()[0], k, min_cells)
        G = net.grid
        lg = net.cell_logits([cy * G + cx for cy, cx in cells]) if cells else np.zeros((0, 1 + PR.N_BINS))
        bins = []
        for q in range(len(cells)):
            lo = np.where(self.reachable, lg[q, 1:], -np.inf)
            bins.append(np.argsort(-lo, kind="stable")[:m])
        t1 = time.perf_counter()
        if self._refiner is None:
            self._refiner = TemplateRefiner(micrograph, self.volume,
                                            RefineSettings(defocus1=c.defocus, defocus2=c.defocus, padding=2.0,
                                                           max_in_flight=256),
                                            template_spectrum=self.judge_spectrum, device=self.device)
            t1 = time.perf_counter()  # one-off set-up of the handle
        else:
            self._refiner.set_micrograph(micrograph)  # prepared on the device
        t2 = time.perf_counter()
        cell_px = c.size // G
        half = c.size // 2
        xy, ang, owner = [], [], []
        for q, ((cy, cx), bq) in enumerate(zip(cells, bins)):
            idx = np.concatenate([self.members[b] for b in bq])
            xy.append(np.tile((cx * cell_px + cell_px // 2 - half, cy * cell_px + cell_px // 2 - half), (idx.size, 1)))
            ang.append(self.orients[idx])
            owner.append(np.full(idx.size, q))
        dets = []
        n = 0
        if xy:
            xy, ang, owner = np.concatenate(xy), np.concatenate(ang), np.concatenate(owner)
            res = self._refiner.score_batch(xy, ang)
            n = int(xy.shape[0])
            for q, cell in enumerate(cells):
                sel = np.nonzero(owner == q)[0]
                j = sel[np.argmax(res[sel, 2])]
                dets.append(PR.Detection(float(xy[j, 0] + res[j, 0] + half), float(xy[j, 1] + res[j, 1] + half),
                                         float(res[j, 2]), *map(float, ang[j]), cell))
        t3 = time.perf_counter()
        return dets, n, t3 - t0, {"network_and_proposals": t1 - t0, "image_preparation": t2 - t1, "scores": t3 - t2}

    def _refine_windows(self, micrograph: np.ndarray, centres: list, orient_sets: list, t_start: float,
                        times: dict):
        """Scores every (window, orientation) pair: window q centred at centres[q] (micrograph pixels),
        orientations orient_sets[q] (indices into the judge's orientations); the best per window is its
        detection."""
        c = self.cfg
        t1 = time.perf_counter()
        if self._refiner is None:
            self._refiner = TemplateRefiner(micrograph, self.volume,
                                            RefineSettings(defocus1=c.defocus, defocus2=c.defocus, padding=2.0,
                                                           max_in_flight=256),
                                            template_spectrum=self.judge_spectrum, device=self.device)
            t1 = time.perf_counter()
        else:
            self._refiner.set_micrograph(micrograph)
        t2 = time.perf_counter()
        half = c.size // 2
        xy, ang, owner = [], [], []
        for q, ((x, y), idx) in enumerate(zip(centres, orient_sets)):
            xy.append(np.tile((int(x) - half, int(y) - half), (idx.size, 1)))
            ang.append(self.orients[idx])
            owner.append(np.full(idx.size, q))
        dets, n = [], 0
        if xy:
            xy, ang, owner = np.concatenate(xy), np.concatenate(ang), np.concatenate(owner)
            n = int(xy.shape[0])
            if n:
                res = self._refiner.score_batch(xy, ang)
                order = np.argsort(owner, kind="stable")
                bounds = np.searchsorted(owner[order], np.arange(len(centres) + 1))
                for q in range(len(centres)):
                    sel = order[bounds[q]:bounds[q + 1]]
                    if sel.size == 0:
                        continue
                    j = sel[np.argmax(res[sel, 2])]
                    dets.append(PR.Detection(float(xy[j, 0] + res[j, 0] + half), float(xy[j, 1] + res[j, 1] + half),
                                             float(res[j, 2]), *map(float, ang[j]), (q,)))
        t3 = time.perf_counter()
        times["image_preparation"] = times.get("image_preparation", 0.0) + t2 - t1
        times["scores"] = t3 - t2
        return dets, n, t3 - t_start

    def detect_coarse(self, net: cnn.ProposalNet | None, micrograph: np.ndarray, k: int, radius_deg: float = 7.5,
                      locations: str = "network", rng=None, min_cells: int = 2, neighbourhood: int = 1,
                      k_network: int = 0):
        """Locations, then a coarse orientation from the binned search, then local exhaustive refinement
        at full resolution: (detections, TemplateScores, seconds, {stage: seconds}).

        locations "network": the network's top k cells, each placed at the binned search's z-score
        maximum inside the cell; "random": k random cells placed the same way (ablation of the network);
        "binned_peaks": the binned search's own k best peaks, no network; "union": those k peaks and the
        network's top k_network cells that lie further than 10 binned pixels from all of them. Coarse
        orientations: the binned search's best angles at the location's binned pixel and its neighbours
        (a (2n+1)^2 block); the refinement scores every judge orientation within radius_deg (rotation
        angle) of any of them, in the template-sized window centred on the location."""
        c = self.cfg
        lb = c.label_bin
        t0 = time.perf_counter()
        times = {}
        cells = None
        if locations in ("network", "union"):
            net.forward(self.net_input(micrograph)[None], want_logits=False)
            cells = PR.propose_cells(net.heat()[0], k if locations == "network" else k_network, min_cells)
        elif locations == "random":
            G = c.grid
            cells = [(int(q) // G, int(q) % G) for q in rng.choice(G * G, size=k, replace=False)]
        t1 = time.perf_counter()
        times["network_and_proposals"] = t1 - t0
        se, n = self.binned_search(micrograph)
        z = se.core.finalize(n)["zscore"]
        r = se.core.results()
        Hb, Wb = z.shape
        pix = []
        if locations in ("binned_peaks", "union"):
            pk = se.core.find_peaks(n, -1e30, se.T // 4 + 1, 10.0, capacity=k)
            pix = [(int(p["x"]), int(p["y"])) for p in pk]
        if cells is not None:
            cb = c.size // c.grid // lb  # binned pixels per cell
            taken = np.array(pix, np.float64).reshape(-1, 2)
            for cy, cx in cells:
                y0, x0 = cy * cb, cx * cb
                box = z[y0:y0 + cb, x0:x0 + cb]
                by, bx = np.unravel_index(int(np.argmax(box)), box.shape)
                q = (x0 + int(bx), y0 + int(by))
                if locations == "union" and taken.size and (((taken - q) ** 2).sum(1) <= 100.0).any():
                    continue
                pix.append(q)
        t2 = time.perf_counter()
        times["binned_search"] = t2 - t1
        centres, sets = [], []
        for bx, by in pix:
            ys = slice(max(by - neighbourhood, 0), min(by + neighbourhood + 1, Hb))
            xs = slice(max(bx - neighbourhood, 0), min(bx + neighbourhood + 1, Wb))
            ang = np.unique(np.stack([r.phi[ys, xs].ravel(), r.theta[ys, xs].ravel(), r.psi[ys, xs].ravel()], 1), axis=0)
            centres.append((round(bx * lb + (lb - 1) / 2.0), round(by * lb + (lb - 1) / 2.0)))
            sets.append(self.orientation_set(ang, radius_deg))
        t3 = time.perf_counter()
        times["orientation_sets"] = t3 - t2
        self.last_centres, self.last_sets = centres, sets
        dets, nscores, dt = self._refine_windows(micrograph, centres, sets, t0, times)
        return dets, nscores, dt, times

    def proposal_quality(self, peaks: list, reach: float = 96.0) -> list:
        """For the last detect_coarse call, per judge peak: (located, covered): a location lies within
        reach pixels on both axes, and the nearest such location's orientation set holds the peak's
        orientation."""
        out = []
        for p in peaks:
            best = None
            for q, (x, y) in enumerate(self.last_centres):
                d = max(abs(x - p["x"]), abs(y - p["y"]))
                if d <= reach and (best is None or d < best[0]):
                    best = (d, q)
            if best is None:
                out.append((False, False))
            else:
                out.append((True, self.judge_index(p) in set(self.last_sets[best[1]].tolist())))
        return out
