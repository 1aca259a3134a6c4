"""Comparison helpers shared by the GPU parity tests (the HIP path against the CPU oracle).

Bars (match_template.cpp:996-1169 is the reference semantics):
  * MIP: |gpu - oracle| <= 1e-4 * |oracle| + 1e-6 * max |oracle| per pixel. The absolute term is
    the FP32 rounding of a correlation plane (single planes agree to ~3e-7 of their maximum); it
    matters only where a maximum over few orientations lies near zero
  * best orientation index: identical wherever the oracle's best value beats its runner-up by more
    than the FP32 noise of the two FFT implementations (checked by recomputing the oracle's CCF at
    the GPU's choice); angle maps identical wherever the index is
  * FP64 sum / sum-of-squares: relative 1e-5 (values differ only by FFT rounding)
  * histogram: totals identical; cumulative counts agree to 1e-5 of the total (values on a bin
    edge may fall on either side)
"""
import numpy as np


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def decode_index(info, gi):
    """Global orientation index -> (size_i, defocus_i, pos, psi_idx), the inverse of
    (plane * n_pos + pos) * n_psi + psi_idx with plane = (size_i + n_size_half) * (2 n_def_half + 1)
    + defocus_i + n_def_half (pos over the whole grid)."""
    plane, rem = divmod(int(gi), info.n_pos * info.n_psi)
    size_plane, def_plane = divmod(plane, 2 * info.n_def_half + 1)
    return (size_plane - getattr(info, "n_size_half", 0), def_plane - info.n_def_half, rem // info.n_psi,
            rem % info.n_psi)


def near_tie_check(mt, ref, got, max_orientations=400):
    """Where the best index differs, the oracle's own CCF at the GPU's orientation must equal the
    oracle's MIP m within FP32 noise, |v - m| <= 1e-5 |m| + 1e-6 max|m|: the two orientations tie to
    within the rounding of two FFT implementations, so either choice is a correct argmax. At most
    max_orientations of the GPU's differing choices are recomputed (the most frequent first)."""
    I = mt.info
    bad = np.argwhere(got.best_index != ref["best_index"])
    if bad.size == 0:
        return 0
    g = got.best_index[bad[:, 0], bad[:, 1]]
    uniq, counts = np.unique(g, return_counts=True)
    order = np.argsort(-counts, kind="stable")[:max_orientations]
    scale = float(np.abs(ref["mip"]).max())
    checked = 0
    worst, at = 0.0, None
    for gi in uniq[order]:
        si, d, pos, k = decode_index(I, gi)
        _, ccf = mt.projection_and_ccf(pos, k, d, size_i=si)
        sel = bad[g == gi]
        v = ccf[sel[:, 0], sel[:, 1]]
        m = ref["mip"][sel[:, 0], sel[:, 1]]
        gap = np.abs(v - m) / (np.abs(m) + 0.1 * scale)  # <= 1e-5 <=> the bar above
        i = int(np.argmax(gap))
        if gap[i] > worst:
            worst, at = float(gap[i]), (tuple(sel[i]), float(m[i]), float(v[i]))
        checked += len(sel)
    print(f"near-tie check: {checked} of {len(bad)} differing pixels over {len(order)} of {len(uniq)} "
          f"orientations, worst gap / (|m| + 0.1 max|m|) {worst:.2e} at {at}; max|m| {scale:.4e}", flush=True)
    assert worst <= 1e-5
    return checked


def compare_search(mt, ref, got, label, max_tie_orientations=400):
    mip_r, mip_g = ref["mip"], got.mip
    scale = float(np.abs(mip_r).max())
    r = np.abs(mip_g - mip_r) / (np.abs(mip_r) + 1e-2 * scale)  # <= 1e-4 <=> the bar above
    same = got.best_index == ref["best_index"]
    print(f"{label}: MIP max err / (|oracle| + 0.01 max|oracle|) {r.max():.2e}; best index identical at "
          f"{same.mean() * 100:.4f}% of pixels", flush=True)
    if r.max() > 1e-4:
        iy, ix = np.unravel_index(np.argmax(r), r.shape)
        print(f"{label}: worst MIP pixel ({iy}, {ix}): oracle {mip_r[iy, ix]:.6e} gpu {mip_g[iy, ix]:.6e} "
              f"index oracle {ref['best_index'][iy, ix]} gpu {got.best_index[iy, ix]}; max |oracle MIP| {scale:.6e}",
              flush=True)
    assert r.max() <= 1e-4
    near_tie_check(mt, ref, got, max_tie_orientations)
    srel = np.abs(got.sum - ref["sum"]).max() / np.abs(ref["sum"]).max()
    qrel = np.abs(got.sumsq - ref["sumsq"]).max() / np.abs(ref["sumsq"]).max()
    print(f"{label}: sum rel {srel:.2e} sumsq rel {qrel:.2e}", flush=True)
    assert srel <= 1e-5 and qrel <= 1e-5
    assert got.hist.sum() == ref["hist"].sum()
    # values within FP32 rounding of a bin edge may land in the neighbouring bin; each such value
    # moves the cumulative count at exactly one edge, so the cumulative histograms must agree to
    # within the expected number of edge values (<< 1e-5 of the total)
    cd = np.abs(np.cumsum(got.hist) - np.cumsum(ref["hist"])).max()
    print(f"{label}: histogram total {got.hist.sum()} max cumulative diff {cd}", flush=True)
    assert cd <= 1e-5 * got.hist.sum()
    # angle and defocus maps follow from the index
    m = same
    for name in ("psi", "theta", "phi", "defocus", "pixel_size"):
        assert np.array_equal(getattr(got, name)[m], ref[name][m]), name
