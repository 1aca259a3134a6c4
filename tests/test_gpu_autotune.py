"""GPU: the launch-shape autotuner (tm_autotune, tm_autotune_batch; no counterpart in the reference).

* the batch depth changes no result: config 0 searched at batch 1, 7, 64 and 256 gives bit-identical
  MIP, best index, FP64 sum and sum of squares, and histogram (each pixel's state is updated in
  orientation order whatever the batch, including a ragged last batch)
* the tuner times every candidate, returns the fastest, and re-initialises the search at it; the
  tuned search then gives the same bit-identical results
* invalid candidates are rejected with an error, not clamped
* at 4096^2 (config-1 micrograph and template, position 3, 241 psi) the three column passes (one
  block per column pair with 4-column Z tiles; the persistent pass with 2- and 4-column tiles) at
  batch 64 and 41 give bit-identical MIP, index, sums and histogram; tm_autotune over (column pass,
  batch) times every pair, picks the fastest and the search re-initialised at it gives the same maps;
  a column pass that does not apply (2-column tiles with 16-bit planes) is reported as such
* with FP16 and BF16 planes the persistent pass (4-column tiles, the default) and the column-pair pass
  give bit-identical results
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import core, synth
from cistem_downstream_bah_amd import match_template as MT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0():
    return synth.config0(0)


def _search(s, batch):
    return MT.MatchTemplateSearch(s.micrograph, s.volume,
                                  MT.SearchParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                                                  first_search_position=0, last_search_position=59,
                                                  batch_size=batch))


def _results(search):
    search.run()
    r = search.core.results()
    return r


def test_batch_depth_changes_no_result(c0):
    ref = None
    for b in (64, 1, 7, 256):
        print(f"case: config0 positions 0-59, batch {b}", flush=True)
        r = _results(_search(c0, b))
        if ref is None:
            ref = r
            continue
        assert np.array_equal(r.mip, ref.mip)
        assert np.array_equal(r.best_index, ref.best_index)
        assert np.array_equal(r.sum, ref.sum)
        assert np.array_equal(r.sumsq, ref.sumsq)
        assert np.array_equal(r.hist, ref.hist)
        assert r.nccf == ref.nccf


def test_autotune_picks_the_fastest(c0):
    print("case: autotune config0, candidates 8 16 32 64 128", flush=True)
    base = _results(_search(c0, 64))
    search = _search(c0, 64)
    cands = (8, 16, 32, 64, 128)
    best, us = search.autotune_batch(cands, sample_orientations=512)
    print(f"autotune: {best} from {us}", flush=True)
    assert set(us) == set(cands) and all(v > 0 for v in us.values())
    assert best == min(cands, key=lambda c: (us[c], cands.index(c)))
    assert search.p.batch_size == best
    r = _results(search)
    assert np.array_equal(r.mip, base.mip) and np.array_equal(r.best_index, base.best_index)
    assert np.array_equal(r.sum, base.sum) and np.array_equal(r.hist, base.hist)


def test_autotune_rejects_bad_candidates(c0):
    search = _search(c0, 64)
    with pytest.raises(core.TMError):
        search.autotune_batch((64, 300))
    with pytest.raises(core.TMError):
        search.autotune_batch((0,))


@pytest.fixture(scope="module")
def c1():
    return synth.config1(1000)


def _search4096(s, batch, column_pass, plane_storage="fp32"):
    return MT.MatchTemplateSearch(s.micrograph, s.volume,
                                  MT.SearchParams(angular_step=2.5, in_plane_angular_step=1.5, padding=2.0,
                                                  first_search_position=3, last_search_position=3,
                                                  batch_size=batch, column_pass=column_pass,
                                                  plane_storage=plane_storage))


def test_column_passes_change_no_result(c1):
    ref = None
    for cp, b in (("persistent", 64), ("pairs", 64), ("persistent4", 64), ("persistent", 41), ("pairs", 41)):
        print(f"case: 4096^2 position 3, column pass {cp}, batch {b}", flush=True)
        se = _search4096(c1, b, cp)
        shape = se.core.launch_shape()
        assert shape["column_pass"] == cp and shape["batch"] == b
        assert shape["z_tile"] == (2 if cp == "persistent" else 4)
        r = _results(se)
        se.core.close()
        if ref is None:
            ref = r
            continue
        for f in ("mip", "best_index", "sum", "sumsq", "hist"):
            assert np.array_equal(getattr(r, f), getattr(ref, f)), (cp, b, f)


def test_autotune_over_column_passes(c1):
    print("case: autotune 4096^2 over column passes x batch 32, 64", flush=True)
    search = _search4096(c1, 64, "auto")
    base = _results(search)
    best_b, best_c, table = search.autotune((32, 64), ("pairs", "persistent", "persistent4"), sample_orientations=192)
    print(f"autotune: {best_c} / {best_b} from {table}", flush=True)
    assert all(v is not None and v > 0 for v in table.values()) and len(table) == 6
    assert table[(best_c, best_b)] == min(table.values())
    shape = search.core.launch_shape()
    assert (shape["column_pass"], shape["batch"]) == (best_c, best_b)
    r = _results(search)
    assert np.array_equal(r.mip, base.mip) and np.array_equal(r.best_index, base.best_index)
    assert np.array_equal(r.sum, base.sum) and np.array_equal(r.hist, base.hist)
    search.core.close()
    fp16 = _search4096(c1, 64, "auto", "fp16")
    assert fp16.core.launch_shape()["column_pass"] == "persistent4"
    _, c16, t16 = fp16.autotune((64,), ("persistent", "pairs", "persistent4"), sample_orientations=128)
    assert c16 in ("pairs", "persistent4") and t16[("persistent", 64)] is None
    assert t16[("pairs", 64)] > 0 and t16[("persistent4", 64)] > 0
    with pytest.raises(core.TMError):
        _search4096(c1, 64, "persistent", "fp16")
    fp16.core.close()


@pytest.mark.parametrize("st", ["fp16", "bf16"])
def test_16bit_column_passes_change_no_result(c1, st):
    """The persistent pass writes the 16-bit plane exactly as the column-pair pass does."""
    ref = None
    for cp in ("pairs", "persistent4"):
        print(f"case: 4096^2 position 3, {st}, column pass {cp}", flush=True)
        se = _search4096(c1, 64, cp, st)
        assert se.core.launch_shape()["column_pass"] == cp
        r = _results(se)
        se.core.close()
        if ref is None:
            ref = r
            continue
        for f in ("mip", "best_index", "sum", "sumsq", "hist"):
            assert np.array_equal(getattr(r, f), getattr(ref, f)), (cp, f)


@pytest.mark.parametrize("W,H,st", [(512, 512, "fp32"), (384, 324, "fp32"), (2048, 1536, "fp32"), (2048, 2048, "fp16"),
                                    (2048, 3072, "fp16"), (4096, 2916, "bf16")])
def test_narrow_column_tiles_change_no_result(W, H, st):
    """At heights other than 4096 the column pass's tiles of the compiled width and of half that width
    (TM_COLUMNS_NARROW) give bit-identical MIP, index, sums and histogram wherever both keep the same
    row pass (2-column tiles read by the direct-load row pass at 2048 x 1536 FP32 and 4-column tiles at
    2048^2 FP16; one-column tiles at 2916 and 3072 with 16-bit planes, whose row pass gathers in both
    cases); the autotuner times both. Where narrow tiles would switch the row pass to the gathering
    kernel (one-column tiles with FP32 planes at width 2048) they are refused."""
    s = synth.make_synthetic(8800 + H, W, H, 64, 200, 24.0, 6, 0.1, 15000.0, min_dist=64)
    tpl = MT.prepare_template(s.volume, 1.0)
    ref = None
    for cp in ("auto", "narrow"):
        print(f"case: {W}x{H} {st}, column tiles {cp}", flush=True)
        se = MT.MatchTemplateSearch(s.micrograph, s.volume,
                                    MT.SearchParams(angular_step=10.0, in_plane_angular_step=15.0, padding=1.0,
                                                    first_search_position=40, last_search_position=41,
                                                    batch_size=16, column_pass=cp, plane_storage=st),
                                    template_spectrum=tpl)
        shape = se.core.launch_shape()
        assert shape["column_pass"] == cp
        if ref is None:
            wide = shape["z_tile"]
        else:
            assert shape["z_tile"] == wide // 2
        r = _results(se)
        if ref is None:
            ref = r
            best_b, best_c, table = se.autotune((16, 32), sample_orientations=64)
            print(f"autotune {W}x{H}: {best_c} / {best_b} from {table}", flush=True)
            assert set(table) == {("auto", 16), ("auto", 32), ("narrow", 16), ("narrow", 32)}
            assert all(v is not None and v > 0 for v in table.values())
            assert table[(best_c, best_b)] == min(table.values())
        else:
            for f in ("mip", "best_index", "sum", "sumsq", "hist"):
                assert np.array_equal(getattr(r, f), getattr(ref, f)), (W, H, st, f)
        se.core.close()
    with pytest.raises(core.TMError):
        _ = MT.MatchTemplateSearch(s.micrograph, s.volume,
                                   MT.SearchParams(angular_step=10.0, in_plane_angular_step=15.0, padding=1.0,
                                                   first_search_position=40, last_search_position=40,
                                                   column_pass="pairs"),
                                   template_spectrum=tpl)
    if (W, H) == (2048, 3072):
        with pytest.raises(core.TMError):
            _ = MT.MatchTemplateSearch(s.micrograph, s.volume,
                                       MT.SearchParams(angular_step=10.0, in_plane_angular_step=15.0, padding=1.0,
                                                       first_search_position=40, last_search_position=40,
                                                       column_pass="narrow"),
                                       template_spectrum=tpl)
