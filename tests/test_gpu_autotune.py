"""GPU: the batch-depth autotuner (tm_autotune_batch; no counterpart in the reference).

* the batch depth changes no result: config 0 searched at batch 1, 7, 64 and 256 gives bit-identical
  MIP, best index, FP64 sum and sum of squares, and histogram (each pixel's state is updated in
  orientation order whatever the batch, including a ragged last batch)
* the tuner times every candidate, returns the fastest, and re-initialises the search at it; the
  tuned search then gives the same bit-identical results
* invalid candidates are rejected with an error, not clamped
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
