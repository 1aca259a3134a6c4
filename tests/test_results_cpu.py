"""CPU: worker/master result records (include/tm_results.h; match_template.cpp:1371-1531,
1617-1760, 2015-2103) against the oracle (oracle/results_oracle.py), and the oracle's pins.

* known answers: a centred crop (out smaller than the map), a centred pad whose fill is the disc
  average (a 6 x 6 ramp with radius 4.2: every pixel but the far corner, mean 18), the valid-area
  cut of remove_npix_from_edge, a constant map packing to that constant, padding floats 0, the
  histogram's int64 bits carried through float pairs
* product vs oracle: pack, add and finalize bit-identical on random maps (even and odd widths, output
  larger and smaller than the maps, histogram counts above 2^32); sd and z within 2 ulp (the
  reference squares the mean with powf, the oracle rounds the exact square)
* merge: two workers' records over disjoint position ranges of the oracle's CPU search (config 0,
  positions 100-101 and 102-103), merged by the master, equal the record of one search over 100-103
  inside the valid area (MIP and every angle and defocus map identical, sums to float rounding,
  histogram and CCF count exact); outside it the MIP holds the larger worker fill; the master's
  z-score equals the local path's closed form (oracle finalise) inside the valid area
* errors: records of different sizes, a wrong size, NULL arguments return errors
"""
import ctypes

import numpy as np
import pytest

from cistem_downstream_bah_amd import core, results as R, synth
from oracle import oracle as O
from oracle import results_oracle as RO

NAMES = R.MAP_NAMES


def _maps(rng, ny, nx):
    return {k: rng.standard_normal((ny, nx)).astype(np.float32) for k in NAMES}


def test_known_answers():
    ramp = np.arange(36, dtype=np.float32).reshape(6, 6)
    maps = {k: ramp + i for i, k in enumerate(NAMES)}
    hist = np.array([1, 2 ** 40 + 3, 0, 7], np.int64)
    # output smaller than the map: centred crop, rows and columns 1..4
    u = R.unpack(R.pack(maps, 4, 4, 0, hist, 12.0, 6.0))
    assert (u["x"], u["y"], u["n"], u["bins"], u["n_ccfs"], u["sqrt_input_pixels"]) == (4, 4, 24, 4, 12.0, 6.0)
    assert np.array_equal(u["mip"], ramp[1:5, 1:5]) and np.array_equal(u["sumsq"], ramp[1:5, 1:5] + 7)
    assert np.array_equal(u["histogram"], hist)
    # output larger than the map: centred pad by one, fill = mean over the disc of radius 0.35 * 12 =
    # 4.2 about (3, 3), i.e. every pixel but (0, 0): 630 / 35 = 18
    u = R.unpack(R.pack(maps, 8, 8, 0, hist, 1.0, 1.0))
    assert np.array_equal(u["mip"][1:7, 1:7], ramp)
    border = np.ones((8, 8), bool)
    border[1:7, 1:7] = False
    assert np.all(u["mip"][border] == 18.0) and np.all(u["psi"][border] == 19.0)
    # valid-area cut: 8 x 8 map, remove 2 -> the centred 6 x 6 kept, the rest the disc average
    big = np.arange(64, dtype=np.float32).reshape(8, 8)
    u = R.unpack(R.pack({k: big for k in NAMES}, 8, 8, 2, hist, 1.0, 1.0))
    assert np.array_equal(u["mip"][1:7, 1:7], big[1:7, 1:7])
    fill = u["mip"][0, 0]
    assert np.all(u["mip"][border] == fill) and 27.0 < fill < 36.0
    # constant maps pack to the constant; padding floats of each row are 0
    rec = R.pack({k: np.full((5, 7), 2.5, np.float32) for k in NAMES}, 9, 5, 1, hist, 1.0, 1.0)
    u = R.unpack(rec)
    assert all(np.all(u[k] == 2.5) for k in NAMES)
    n = int(rec[2])
    rows = rec[R.META:R.META + n].reshape(5, 10)
    assert np.all(rows[:, 9] == 0.0)


@pytest.mark.parametrize("shape,out,remove", [((48, 64), (64, 48), 5), ((50, 63), (70, 44), 3), ((50, 50), (50, 50), 0),
                                              ((33, 40), (41, 31), 0)])
def test_product_matches_oracle(shape, out, remove):
    rng = np.random.default_rng(sum(shape) + remove)
    hists = [rng.integers(0, 2 ** 36, 512).astype(np.int64) for _ in range(3)]
    recs = []
    for w in range(3):
        maps = _maps(rng, *shape)
        maps["mip"] += 0.3 * w
        got = R.pack(maps, out[0], out[1], remove, hists[w], 100.0 + w, 41.5)
        ref = RO.pack([maps[k] for k in NAMES], out[0], out[1], remove, hists[w], 100.0 + w, 41.5)
        assert got.size == ref.size == R.result_size(out[0], out[1], 512)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        recs.append(got)
    agg = R.Aggregate()
    col, tot = None, np.float32(0)
    for r in recs:
        agg.add(r)
        col, tot = RO.add(col, r, tot)
        assert np.array_equal(agg.collated.view(np.uint32), col.view(np.uint32))
    assert agg.total_ccfs.value == tot == np.float32(303.0)
    got, ref = agg.finalize(), RO.finalize(col, tot)
    for k in ("mip_scaled", "average", "psi", "theta", "phi", "defocus", "pixel_size", "histogram"):
        assert np.array_equal(got[k], ref[k]), k
    for k in ("std_dev", "zscore"):
        np.testing.assert_array_max_ulp(got[k], ref[k], maxulp=2)
    assert np.array_equal(got["histogram"], hists[0] + hists[1] + hists[2])


@pytest.fixture(scope="module")
def searches():
    s = synth.config0(0)
    p = dict(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0)
    out = {}
    for name, lo, hi in (("a", 100, 101), ("b", 102, 103), ("all", 100, 103)):
        mt = O.MatchTemplateOracle(O.MTParams(**p, first_search_position=lo, last_search_position=hi), s.micrograph,
                                   s.volume)
        out[name] = (mt, mt.run(threads=4))
    return out


def _record(mt, res, remove):
    I = mt.info
    maps = {k: res[k].astype(np.float32) for k in NAMES}
    return R.pack(maps, I.W, I.H, remove, res["hist"], float(res["nccf"]), I.sqrt_input_pixels)


def test_merge_of_two_workers_equals_one_search(searches):
    remove = 16
    (ma, ra), (mb, rb), (mf, rf) = searches["a"], searches["b"], searches["all"]
    assert not mf.info.rotated and ra["nccf"] + rb["nccf"] == rf["nccf"] > 0
    agg = R.Aggregate()
    rec_a, rec_b = _record(ma, ra, remove), _record(mb, rb, remove)
    agg.add(rec_a)
    agg.add(rec_b)
    merged, single = R.unpack(agg.collated), R.unpack(_record(mf, rf, remove))
    inner = (slice(remove // 2, -remove // 2), slice(remove // 2, -remove // 2))
    for k in ("mip", "psi", "theta", "phi", "defocus", "pixel_size"):
        assert np.array_equal(merged[k][inner], single[k][inner]), k
    for k in ("sum", "sumsq"):
        np.testing.assert_allclose(merged[k][inner], single[k][inner], rtol=2e-6, atol=1e-6 * np.abs(single[k]).max())
    assert np.array_equal(merged["histogram"], rf["hist"])
    assert agg.total_ccfs.value == rf["nccf"]
    fa, fb = R.unpack(rec_a)["mip"][0, 0], R.unpack(rec_b)["mip"][0, 0]
    assert merged["mip"][0, 0] == max(fa, fb)
    # the master's z-score against the local path's closed form over the same CCFs
    fin = agg.finalize()
    loc = mf.finalise(total_ccfs=int(rf["nccf"]))
    z, zl = fin["zscore"][inner], loc["zscore"][inner]
    np.testing.assert_allclose(z, zl, rtol=1e-3, atol=1e-3 * np.abs(zl).max())
    np.testing.assert_allclose(fin["mip_scaled"][inner], loc["mip_scaled"][inner], rtol=1e-6)


def test_errors_are_returned():
    L = R._lib()
    rng = np.random.default_rng(1)
    r1 = R.pack(_maps(rng, 8, 8), 8, 8, 0, np.zeros(4, np.int64), 1.0, 8.0)
    r2 = R.pack(_maps(rng, 8, 8), 10, 8, 0, np.zeros(4, np.int64), 1.0, 8.0)
    agg = R.Aggregate()
    agg.add(r1)
    with pytest.raises(core.TMError):
        agg.add(r2)                     # different map sizes
    with pytest.raises(core.TMError):
        R.pack(_maps(rng, 8, 8), 8, 8, 9, np.zeros(4, np.int64), 1.0, 8.0)   # no valid area left
    bad = r1.copy()
    bad[2] += 2                         # header inconsistent with x, y
    tot = ctypes.c_float(0)
    assert L.tm_result_add(agg.collated.ctypes.data_as(R.F32P), bad.ctypes.data_as(R.F32P), bad.size,
                           ctypes.byref(tot), 0) != 0 and L.tm_result_last_error()
    assert L.tm_result_add(None, None, 0, None, 1) != 0
    assert L.tm_result_size(0, 4, 512) == -1
