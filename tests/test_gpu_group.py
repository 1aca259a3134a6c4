"""GPU: searches of images of different sizes run as one hipGraph launch per plane (tm_group_run: the
images dealt to chains, each chain an independent branch of the graph) give each image the maps, sums
and histogram of its own tm_run_inner_loop calls, bit for bit, with one, two and three chains: every
image's kernels run on its own buffers in its own order. The second plane replays the captured graph
with a new plane offset; a handle re-initialised after the capture (its buffers moved) makes the group
capture again."""
import numpy as np
import pytest

from cistem_downstream_bah_amd import core
from cistem_downstream_bah_amd import match_template as MT
from cistem_downstream_bah_amd import synth

pytestmark = pytest.mark.gpu

NAMES = ("mip", "best_index", "psi", "theta", "phi", "defocus", "sum", "sumsq", "hist", "nccf")


def _searches():
    sizes = [(512, 512), (384, 324), (260, 260)]
    mics = [synth.make_synthetic(9100 + i, w, h, 64, 200, 24.0, 3, 0.1, 15000.0, min_dist=64)
            for i, (w, h) in enumerate(sizes)]
    tpl = MT.prepare_template(mics[0].volume, 1.0)
    out = []
    for i, s in enumerate(mics):
        sp = MT.SearchParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0, defocus_search_range=200.0,
                             defocus_step=200.0, first_search_position=4 * i, last_search_position=4 * i + 5 + i,
                             batch_size=16 + 8 * i)
        out.append(MT.MatchTemplateSearch(s.micrograph, mics[0].volume, sp, template_spectrum=tpl))
    return out


def _same(a, b, tag):
    for name in NAMES:
        assert np.array_equal(np.asarray(getattr(a, name)), np.asarray(getattr(b, name))), (tag, name)


@pytest.mark.parametrize("lanes", [1, 2, 3])
def test_group_equals_separate_runs(lanes):
    searches = _searches()
    print(f"case: group of 3 images (512^2, 384x324, 270^2), 3 defocus planes, {lanes} chains", flush=True)
    n_sep = [m.run() for m in searches]
    sep = [m.core.results() for m in searches]
    for m in searches:
        m.core.reset()
    g = core.TemplateMatchingGroup([m.core for m in searches], lanes)
    n_grp = MT.run_group(searches, g)
    assert n_grp == n_sep == [m.total_correlation_positions for m in searches]
    for m, a in zip(searches, sep):
        _same(a, m.core.results(), (m.W, m.H))
    # replays after a reset, with eager work in between, give the same maps again
    for m in searches:
        m.core.reset()
        m.core.debug_ccf(m.filter_for(0), m.first, 0)
    MT.run_group(searches, g)
    for m, a in zip(searches, sep):
        _same(a, m.core.results(), ("replay", m.W, m.H))
    # a re-initialised handle (new buffers) is captured again
    m = searches[1]
    m.core.init(m.template_spectrum, m.input_spectrum, m.T, m.grid, m.psi, m.first, m.last, m.histogram_min_scaled,
                m.histogram_step_scaled, 512, m.max_padding, m.defocus_step, m.pixel_size_step, m.p.batch_size)
    for q in searches:
        q.core.reset()
    MT.run_group(searches, g)
    for q, a in zip(searches, sep):
        _same(a, q.core.results(), ("re-init", q.W, q.H))
    g.close()
    for q in searches:
        q.core.close()


def test_group_rejects_bad_arguments():
    searches = _searches()[:2]
    cores = [m.core for m in searches]
    with pytest.raises(core.TMError):
        core.TemplateMatchingGroup([cores[0], cores[0]])
    with pytest.raises(core.TMError):
        core.TemplateMatchingGroup(cores, -1)
    g = core.TemplateMatchingGroup(cores)
    with pytest.raises(ValueError):
        g.run([searches[0].filter_for(0)])
    cores[0].set_timing(True)
    with pytest.raises(core.TMError):
        g.run([m.filter_for(0) for m in searches])
    cores[0].set_timing(False)
    g.close()
    for c in cores:
        c.close()
