"""Seeded synthetic inputs for the template-matching workloads.

* Template: an isotropic-Gaussian pseudo-atom volume (T^3 float32), atoms uniform in a ball
  centred on the box centre (T/2, T/2, T/2) — the reference's box-centre convention
  (src/core/image.cpp:5019-5034).
* Micrograph: copies of the template projected at uniform-random SO(3) orientations and
  non-overlapping positions, multiplied in Fourier space by a CTF (the -sin(chi) model of
  src/core/ctf.cpp:445-541), plus white Gaussian noise at a stated SNR
  (SNR = var(signal image) / sigma_noise^2 over the full frame).

Orientation convention. The search extracts the Fourier slice at R(phi, theta, psi) * (kx, ky, 0)
(src/core/image.cpp:2450-2607 with the ZYZ matrix of src/core/matrix.cpp:490-515), which is the
transform of p(u) = sum_z V(R (u, z)). A Gaussian atom at a therefore projects to a 2-D Gaussian
at (R^T a)_xy, with integral along z equal to amplitude * sigma * sqrt(2 pi). Inserted copies use
exactly that, so the ground-truth angles are directly comparable with the search's best angles.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

PIf = np.float32(3.14159265359)


def euler_matrix(phi: float, theta: float, psi: float) -> np.ndarray:
    """ZYZ rotation matrix m[row][col] from degrees (src/core/matrix.cpp:490-515)."""
    d2r = np.pi / 180.0
    cp, sp = np.cos(phi * d2r), np.sin(phi * d2r)
    ct, st = np.cos(theta * d2r), np.sin(theta * d2r)
    cs, ss = np.cos(psi * d2r), np.sin(psi * d2r)
    m = np.empty((3, 3))
    m[0, 0] = cp * ct * cs - sp * ss
    m[1, 0] = sp * ct * cs + cp * ss
    m[2, 0] = -st * cs
    m[0, 1] = -cp * ct * ss - sp * cs
    m[1, 1] = -sp * ct * ss + cp * cs
    m[2, 1] = st * ss
    m[0, 2] = st * cp
    m[1, 2] = st * sp
    m[2, 2] = ct
    return m


def random_euler(rng: np.random.Generator):
    """Uniform SO(3) sample returned as ZYZ (phi, theta, psi) degrees."""
    phi = rng.uniform(0.0, 360.0)
    psi = rng.uniform(0.0, 360.0)
    theta = np.degrees(np.arccos(rng.uniform(-1.0, 1.0)))
    return float(phi), float(theta), float(psi)


def make_atoms(rng: np.random.Generator, n_atoms: int, radius: float):
    """Positions uniform in a ball (relative to the box centre) and amplitudes U(0.5, 1)."""
    d = rng.standard_normal((n_atoms, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.uniform(0.0, 1.0, n_atoms) ** (1.0 / 3.0)
    pos = d * r[:, None]
    amp = rng.uniform(0.5, 1.0, n_atoms)
    return pos, amp


def render_volume(T: int, pos: np.ndarray, amp: np.ndarray, sigma: float = 1.5) -> np.ndarray:
    """(z, y, x) float32 volume; atom coordinates (x, y, z) are relative to the box centre T/2."""
    vol = np.zeros((T, T, T), np.float64)
    c = T // 2
    half = int(np.ceil(4 * sigma))
    offs = np.arange(-half, half + 1)
    for (x, y, z), a in zip(pos, amp):
        ix, iy, iz = int(np.floor(c + x)), int(np.floor(c + y)), int(np.floor(c + z))
        xs, ys, zs = ix + offs, iy + offs, iz + offs
        gx = np.exp(-((xs - (c + x)) ** 2) / (2 * sigma * sigma))
        gy = np.exp(-((ys - (c + y)) ** 2) / (2 * sigma * sigma))
        gz = np.exp(-((zs - (c + z)) ** 2) / (2 * sigma * sigma))
        mx, my, mz = (xs >= 0) & (xs < T), (ys >= 0) & (ys < T), (zs >= 0) & (zs < T)
        blk = a * gz[mz][:, None, None] * gy[my][None, :, None] * gx[mx][None, None, :]
        vol[zs[mz][0]:zs[mz][-1] + 1, ys[my][0]:ys[my][-1] + 1, xs[mx][0]:xs[mx][-1] + 1] += blk
    return vol.astype(np.float32)


def ctf_2d(ny: int, nx: int, pixel_size: float, kv: float, cs_mm: float, amp_contrast: float,
           defocus_angstrom: float, phase_shift_rad: float = 0.0) -> np.ndarray:
    """-sin(chi) on the rfft grid (ny, nx//2+1), the model of src/core/ctf.cpp:93-150, 445-541."""
    lam = 12.2639 / np.sqrt(1000.0 * kv + 0.97845e-6 * (1000.0 * kv) ** 2) / pixel_size
    cs = cs_mm * 1e7 / pixel_size
    df = defocus_angstrom / pixel_size
    ac = np.arctan(amp_contrast / np.sqrt(1.0 - amp_contrast ** 2))
    fy = np.fft.fftfreq(ny)[:, None]
    fx = np.arange(nx // 2 + 1)[None, :] / nx
    g2 = fx * fx + fy * fy
    chi = np.pi * lam * g2 * (df - 0.5 * lam * lam * g2 * cs) + phase_shift_rad + ac
    return (-np.sin(chi)).astype(np.float32)


def place_particles(rng: np.random.Generator, n: int, W: int, H: int, margin: int, min_dist: float):
    """Random sequential placement with a minimum centre distance (relaxed by 5% when stuck)."""
    pts = []
    acc = np.empty((n, 2))
    d = float(min_dist)
    tries = 0
    while len(pts) < n:
        x = rng.uniform(margin, W - margin)
        y = rng.uniform(margin, H - margin)
        k = len(pts)
        if k == 0 or bool(np.all((x - acc[:k, 0]) ** 2 + (y - acc[:k, 1]) ** 2 >= d * d)):
            pts.append((x, y))
            acc[k] = (x, y)
            tries = 0
        else:
            tries += 1
            if tries > 2000:
                d *= 0.95
                tries = 0
    return np.array(pts), d


@dataclass
class Particle:
    x: float
    y: float
    phi: float
    theta: float
    psi: float


@dataclass
class Synthetic:
    micrograph: np.ndarray          # (H, W) float32
    volume: np.ndarray              # (T, T, T) float32, z-y-x
    particles: list = field(default_factory=list)
    noise_sigma: float = 0.0
    signal_var: float = 0.0
    min_dist: float = 0.0
    defocus: float = 0.0


def project_copy(pos: np.ndarray, amp: np.ndarray, m: np.ndarray, cx: float, cy: float, W: int, H: int,
                 box: int, sigma: float, out: np.ndarray) -> None:
    """Adds the projection of the atom model at rotation m, centred at (cx, cy), into out (H, W)."""
    proj = pos @ m  # rows: (R^T a)^T = a^T R
    px = cx + proj[:, 0]
    py = cy + proj[:, 1]
    x0 = max(0, int(np.floor(cx)) - box // 2)
    y0 = max(0, int(np.floor(cy)) - box // 2)
    x1 = min(W, int(np.floor(cx)) + box // 2 + 1)
    y1 = min(H, int(np.floor(cy)) + box // 2 + 1)
    xs = np.arange(x0, x1, dtype=np.float64)
    ys = np.arange(y0, y1, dtype=np.float64)
    s2 = 2.0 * sigma * sigma
    gx = np.exp(-((xs[None, :] - px[:, None]) ** 2) / s2)
    gy = np.exp(-((ys[None, :] - py[:, None]) ** 2) / s2)
    w = amp * sigma * np.sqrt(2.0 * np.pi)
    out[y0:y1, x0:x1] += (gy.T * w[None, :]) @ gx


def make_synthetic(seed: int, W: int, H: int, T: int, n_atoms: int, atom_radius: float, n_copies: int,
                   snr: float, defocus: float, pixel_size: float = 1.0, kv: float = 300.0, cs_mm: float = 2.7,
                   amp_contrast: float = 0.07, sigma: float = 1.5, min_dist: float | None = None,
                   template: tuple | None = None) -> Synthetic:
    """template = (atom positions, amplitudes, rendered volume) to reuse a model; else one is drawn
    from the seed first."""
    rng = np.random.default_rng(seed)
    if template is None:
        pos, amp = make_atoms(rng, n_atoms, atom_radius)
        vol = render_volume(T, pos, amp, sigma)
    else:
        pos, amp, vol = template
    if min_dist is None:
        min_dist = T
    centres, used = place_particles(rng, n_copies, W, H, T // 2, min_dist)
    img = np.zeros((H, W), np.float64)
    parts = []
    box = int(2 * (atom_radius + 4 * sigma)) + 2
    for (cx, cy) in centres:
        phi, theta, psi = random_euler(rng)
        m = euler_matrix(phi, theta, psi)
        project_copy(pos, amp, m, cx, cy, W, H, box, sigma, img)
        parts.append(Particle(float(cx), float(cy), phi, theta, psi))
    spec = np.fft.rfft2(img)
    spec *= ctf_2d(H, W, pixel_size, kv, cs_mm, amp_contrast, defocus)
    signal = np.fft.irfft2(spec, s=(H, W))
    var = float(signal.var())
    noise_sigma = float(np.sqrt(var / snr)) if snr > 0 else 0.0
    noisy = signal + noise_sigma * rng.standard_normal((H, W))
    return Synthetic(noisy.astype(np.float32), vol, parts, noise_sigma, var, used, defocus)


_CTF_CACHE: dict = {}


def make_synthetic_gpu(seed: int, W: int, H: int, T: int, n_atoms: int, atom_radius: float, n_copies: int,
                       snr: float, defocus: float, pixel_size: float = 1.0, kv: float = 300.0, cs_mm: float = 2.7,
                       amp_contrast: float = 0.07, sigma: float = 1.5, min_dist: float | None = None,
                       template: tuple | None = None, timing: dict | None = None) -> Synthetic:
    """make_synthetic with the copies' projections, the CTF and the noise scaling computed on the GPU in
    float64 (torch): the same random draws in the same order (template, positions, orientations, then
    the noise from the numpy generator), so the micrograph equals make_synthetic's up to float64
    rounding. For the workloads that need hundreds of 4096^2 micrographs. Falls back to make_synthetic
    when torch sees no GPU (timing["host_fallback"] then says why); torch must be imported before
    libtm_hip.so is loaded, so that both use torch's HIP runtime. timing: a dict that receives the
    seconds of each phase (the device is synchronised after each phase)."""
    try:
        import torch
        if not torch.cuda.is_available():
            raise ImportError("torch sees no GPU")
    except ImportError as e:
        if timing is not None:
            timing["host_fallback"] = str(e)
        return make_synthetic(seed, W, H, T, n_atoms, atom_radius, n_copies, snr, defocus, pixel_size, kv, cs_mm,
                              amp_contrast, sigma, min_dist, template)
    dev = torch.device("cuda")
    clock = [time.perf_counter()]

    def phase(name):
        if timing is not None:
            torch.cuda.synchronize()
            now = time.perf_counter()
            timing[name] = timing.get(name, 0.0) + now - clock[0]
            clock[0] = now

    rng = np.random.default_rng(seed)
    if template is None:
        pos, amp = make_atoms(rng, n_atoms, atom_radius)
        vol = render_volume(T, pos, amp, sigma)
    else:
        pos, amp, vol = template
    if min_dist is None:
        min_dist = T
    centres, used = place_particles(rng, n_copies, W, H, T // 2, min_dist)
    phase("placement")
    box = int(2 * (atom_radius + 4 * sigma)) + 2
    img = torch.zeros((H, W), dtype=torch.float64, device=dev)
    s2 = 2.0 * sigma * sigma
    w = torch.tensor(amp * sigma * np.sqrt(2.0 * np.pi), dtype=torch.float64, device=dev)
    eulers = [random_euler(rng) for _ in range(len(centres))]  # the host generator's draw order
    parts = [Particle(float(cx), float(cy), *e) for (cx, cy), e in zip(centres, eulers)]
    lo = [(int(np.floor(cx)) - box // 2, int(np.floor(cy)) - box // 2) for cx, cy in centres]
    inside = all(x0 >= 0 and y0 >= 0 and x0 + box + 1 <= W and y0 + box + 1 <= H for x0, y0 in lo)
    chunk = 32 if inside else 1
    nb = box + 1
    for c0 in range(0, len(centres), chunk):
        cs = range(c0, min(c0 + chunk, len(centres)))
        proj = np.stack([pos @ euler_matrix(*eulers[i]) for i in cs])  # (copies, atoms, 3)
        px = torch.tensor(np.array([centres[i][0] for i in cs])[:, None] + proj[:, :, 0], dtype=torch.float64, device=dev)
        py = torch.tensor(np.array([centres[i][1] for i in cs])[:, None] + proj[:, :, 1], dtype=torch.float64, device=dev)
        if inside:  # every box lies inside the image: one batched product per chunk of copies
            x0 = torch.tensor([lo[i][0] for i in cs], dtype=torch.float64, device=dev)
            y0 = torch.tensor([lo[i][1] for i in cs], dtype=torch.float64, device=dev)
            r = torch.arange(nb, dtype=torch.float64, device=dev)
            xs, ys = x0[:, None] + r[None, :], y0[:, None] + r[None, :]
        else:
            i = cs[0]
            cx, cy = centres[i]
            x0, y0 = max(0, lo[i][0]), max(0, lo[i][1])
            x1 = min(W, int(np.floor(cx)) + box // 2 + 1)
            y1 = min(H, int(np.floor(cy)) + box // 2 + 1)
            xs = torch.arange(x0, x1, dtype=torch.float64, device=dev)[None]
            ys = torch.arange(y0, y1, dtype=torch.float64, device=dev)[None]
        gx = torch.exp(-((xs[:, None, :] - px[:, :, None]) ** 2) / s2)  # (copies, atoms, nx)
        gy = torch.exp(-((ys[:, None, :] - py[:, :, None]) ** 2) / s2)
        tiles = torch.bmm((gy * w[None, :, None]).transpose(1, 2).contiguous(), gx.contiguous())
        for k, i in enumerate(cs):  # added copy by copy, in the host generator's order
            xa, ya = max(0, lo[i][0]), max(0, lo[i][1])
            img[ya:ya + tiles.shape[1], xa:xa + tiles.shape[2]] += tiles[k]
    phase("projections")
    key = (H, W, pixel_size, kv, cs_mm, amp_contrast, defocus)
    ctf = _CTF_CACHE.get(key)
    if ctf is None:
        ctf = torch.tensor(ctf_2d(*key), dtype=torch.float64, device=dev)
        _CTF_CACHE.clear()  # one micrograph shape at a time
        _CTF_CACHE[key] = ctf
    signal = torch.fft.irfft2(torch.fft.rfft2(img) * ctf, s=(H, W))
    var = float(signal.var(unbiased=False))
    noise_sigma = float(np.sqrt(var / snr)) if snr > 0 else 0.0
    phase("ctf")
    noise = rng.standard_normal((H, W))
    phase("noise_draw")
    noise = torch.from_numpy(noise).to(dev)
    noisy = (signal + noise_sigma * noise).to(torch.float32).cpu().numpy()
    phase("noise_add_and_copy")
    return Synthetic(noisy, vol, parts, noise_sigma, var, used, defocus)


# Workload presets (sizes and counts of the benchmark configurations).
def config0(seed: int = 0) -> Synthetic:
    """512^2 micrograph, 64^3 template (200 atoms, r=24), 10 copies, SNR 0.1, defocus 15000 A."""
    return make_synthetic(seed, 512, 512, 64, 200, 24.0, 10, 0.1, 15000.0, min_dist=64)


def config1(seed: int = 1000) -> Synthetic:
    """4096^2 micrograph, 256^3 template (3000 atoms, r=100), 300 copies, SNR 0.05, defocus 15000 A."""
    return make_synthetic(seed, 4096, 4096, 256, 3000, 100.0, 300, 0.05, 15000.0, min_dist=200)


# Multi-micrograph workload: original micrograph sizes (x, y) and, after the reference's size policy,
# 2048^2, 2048 x 3072 (rotated), 4096^2, 4096 x 2916, 2048 x 1536, 4096^2, 4096^2, 2592^2.
CONFIG2_SIZES = [(2048, 2048), (3072, 2048), (4096, 4096), (4096, 2880), (2048, 1536), (3840, 3840), (4096, 4096),
                 (2560, 2560)]


def config2(seed: int = 2000, sizes=None) -> list:
    """One 256^3 template (3000 atoms, r=100, as the 4096^2 workload) and one micrograph per size with
    U(50, 300) copies, SNR U(0.05, 0.1) and defocus U(8000, 25000) A (recorded in .defocus)."""
    rng = np.random.default_rng(seed)
    pos, amp = make_atoms(rng, 3000, 100.0)
    tpl = (pos, amp, render_volume(256, pos, amp, 1.5))
    out = []
    for i, (w, h) in enumerate(sizes or CONFIG2_SIZES):
        n = int(rng.integers(50, 301))
        snr = float(rng.uniform(0.05, 0.1))
        df = float(rng.uniform(8000.0, 25000.0))
        out.append(make_synthetic(seed + 1 + i, w, h, 256, 3000, 100.0, n, snr, df, min_dist=200, template=tpl))
    return out
