"""GPU: images searched concurrently (one stream and handle each, several host threads) give the
same per-image results, bit for bit, as the same searches run one after another: the kernels are
deterministic (fixed per-pixel accumulation order, integer histogram counts)."""
import numpy as np
import pytest

from cistem_downstream_bah_amd import match_template as MT
from cistem_downstream_bah_amd import synth

pytestmark = pytest.mark.gpu


def test_concurrent_equals_sequential():
    sizes = [(512, 512), (384, 324), (260, 260)]
    mics = [synth.make_synthetic(9000 + i, w, h, 64, 200, 24.0, 3, 0.1, 15000.0, min_dist=64)
            for i, (w, h) in enumerate(sizes)]
    tpl = MT.prepare_template(mics[0].volume, 1.0)
    searches = []
    for i, s in enumerate(mics):
        sp = MT.SearchParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                             first_search_position=5 * i, last_search_position=5 * i + 7, batch_size=16)
        # all three use the first micrograph's template model, so the shared spectrum applies
        searches.append(MT.MatchTemplateSearch(s.micrograph, mics[0].volume, sp, template_spectrum=tpl))
    n_seq = MT.run_concurrent(searches, 1)
    seq = [m.core.results() for m in searches]
    for m in searches:
        m.core.reset()
    n_con = MT.run_concurrent(searches, 3)
    con = [m.core.results() for m in searches]
    assert n_seq == n_con == [8 * m.psi.size for m in searches]
    for a, b, (w, h) in zip(seq, con, [(m.W, m.H) for m in searches]):
        for name in ("mip", "best_index", "sum", "sumsq", "hist"):
            assert np.array_equal(getattr(a, name), getattr(b, name)), (w, h, name)
    for m in searches:
        m.core.close()
