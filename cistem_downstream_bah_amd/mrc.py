"""MRC mode-2 (float32) reader and writer.

Mirrors the subset of the reference's MRC handling the template-matching path reads and
writes: a 1024-byte header (nx, ny, nz, mode, ..., nsymbt extended-header length) followed by
little-endian float32 data, x fastest (src/core/mrc_header.h:58-150, src/core/mrc_file.cpp).
Only mode 2 is supported; other modes raise ValueError.
"""
from __future__ import annotations

import numpy as np

HEADER_BYTES = 1024


def read_mrc(path: str) -> np.ndarray:
    """Returns the volume as (nz, ny, nx) float32 ((ny, nx) when nz == 1)."""
    with open(path, "rb") as f:
        raw = f.read()
    hdr = np.frombuffer(raw[:HEADER_BYTES], dtype="<i4")
    nx, ny, nz, mode = (int(v) for v in hdr[:4])
    if mode != 2:
        raise ValueError(f"{path}: MRC mode {mode} not supported (mode 2 float32 only)")
    nsymbt = int(hdr[23])
    start = HEADER_BYTES + nsymbt
    count = nx * ny * nz
    data = np.frombuffer(raw[start:start + 4 * count], dtype="<f4").astype(np.float32)
    if data.size != count:
        raise ValueError(f"{path}: truncated MRC data ({data.size} of {count} values)")
    vol = data.reshape(nz, ny, nx)
    return vol[0] if nz == 1 else vol


def write_mrc(path: str, data: np.ndarray, pixel_size: float = 1.0) -> None:
    """Writes a (ny, nx) or (nz, ny, nx) array as MRC mode 2 with a minimal header."""
    a = np.asarray(data, dtype="<f4")
    if a.ndim == 2:
        a = a[None]
    nz, ny, nx = a.shape
    hdr = np.zeros(256, dtype="<i4")
    hdr[0:3] = (nx, ny, nz)
    hdr[3] = 2
    hdr[7:10] = (nx, ny, nz)                       # mx, my, mz
    cell = np.array([nx * pixel_size, ny * pixel_size, nz * pixel_size], "<f4")
    hdr[10:13] = cell.view("<i4")
    hdr[13:16] = np.array([90.0, 90.0, 90.0], "<f4").view("<i4")
    hdr[16:19] = (1, 2, 3)
    hdr[19:22] = np.array([a.min(), a.max(), a.mean()], "<f4").view("<i4")
    hdr[52] = int.from_bytes(b"MAP ", "little")
    hdr[53] = 0x00004444                            # little-endian machine stamp
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.ascontiguousarray(a).tobytes())
