"""GPU: checkpoint and resume of a search, truncation to the first N orientations, and the
allocation-failure path (tm_run_partial, tm_export_state / tm_import_state, tm_init_params
orientation_limit and device_memory_limit).

* a two-plane search interrupted inside its first plane (after three tm_run_partial calls, one of
  which ends inside a batch), exported, and resumed by a new handle from the exported state gives
  the uninterrupted search's MIP, best index, angle and defocus maps, FP64 sums, histogram and CCF
  count bit for bit: every pixel sees the orientations in the same order
* a handle limited to its first 200 orientations equals tm_run_partial(0, 200) of the full range
* tm_init under a device-memory limit below its working set returns TM_ERR_HIP and leaves the
  handle uninitialised; the same handle then initialises without the limit and searches normally
* a state from another search is refused
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import core, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prepared():
    s = synth.config0(0)
    mt = O.MatchTemplateOracle(O.MTParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                                          defocus_search_range=100.0, defocus_step=100.0,
                                          first_search_position=100, last_search_position=109),
                               s.micrograph, s.volume)
    return mt


def _handle(mt, **kw):
    I = mt.info
    g = core.TemplateMatchingCore(0)
    g.init(mt.template_spectrum, mt.input_spectrum, I.T, mt.grid, mt.psi, I.first_pos, I.last_pos,
           I.histogram_min_scaled, I.histogram_step_scaled, 512, I.max_padding, I.defocus_step, I.pixel_size_step,
           batch_size=16, **kw)
    return g


def _same(a, b):
    for f in ("mip", "psi", "theta", "phi", "defocus", "sum", "sumsq", "best_index", "hist"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.dtype == y.dtype and np.array_equal(x.view(np.uint8), y.view(np.uint8)), f
    assert a.nccf == b.nccf


def test_resume_from_exported_state_is_bit_identical(prepared):
    mt = prepared
    print("case: config0 positions 100-109, two planes, interrupted at orientation 137 of the first", flush=True)
    f0, f1 = mt.projection_filter(-1), mt.projection_filter(0)
    full = _handle(mt)
    first = _handle(mt)
    second = _handle(mt)
    try:
        n = full.run_inner_loop(f0, 0, -1) + full.run_inner_loop(f1, 0, 0)
        total = full.n_orientations
        assert total == 370 and n == 740
        assert first.run_partial(f0, 0, -1, 0, 37) == 37
        assert first.run_partial(f0, 0, -1, 37, 70) == 70      # ends inside a batch of 16
        assert first.run_partial(f0, 0, -1, 107, 30) == 30
        state = first.export_state()
        first.close()
        second.import_state(state)
        with pytest.raises(core.TMError):
            second.run_partial(f0, 0, -1, 100, 10)             # not where the plane stopped
        assert second.run_partial(f0, 0, -1, 137, total - 137) == total - 137
        assert second.run_partial(f1, 0, 0, 0, total) == total
        _same(second.results(), full.results())
    finally:
        for g in (full, first, second):
            g.close()


def test_orientation_limit_equals_a_partial_run(prepared):
    mt = prepared
    f = mt.projection_filter(0)
    cut = _handle(mt, orientation_limit=200)
    part = _handle(mt)
    try:
        assert cut.n_orientations == 200
        assert cut.run_inner_loop(f, 0, 0) == 200
        assert part.run_partial(f, 0, 0, 0, 200) == 200
        _same(cut.results(), part.results())
    finally:
        cut.close()
        part.close()


def test_device_memory_limit_fails_cleanly_then_recovers(prepared):
    mt = prepared
    I = mt.info
    g = core.TemplateMatchingCore(0)
    try:
        with pytest.raises(core.TMError) as e:
            g.init(mt.template_spectrum, mt.input_spectrum, I.T, mt.grid, mt.psi, I.first_pos, I.last_pos,
                   I.histogram_min_scaled, I.histogram_step_scaled, 512, I.max_padding, I.defocus_step,
                   I.pixel_size_step, batch_size=16, device_memory_limit=4 << 20)
        print(f"limited init: {e.value}", flush=True)
        assert e.value.code == -2 and "out of memory" in str(e.value)
        with pytest.raises(core.TMError) as e2:
            g.run_inner_loop(mt.projection_filter(0), 0, 0)
        assert e2.value.code == -4                              # TM_ERR_STATE: not initialised
        g.init(mt.template_spectrum, mt.input_spectrum, I.T, mt.grid, mt.psi, I.first_pos, I.last_pos,
               I.histogram_min_scaled, I.histogram_step_scaled, 512, I.max_padding, I.defocus_step,
               I.pixel_size_step, batch_size=16)
        assert g.run_inner_loop(mt.projection_filter(0), 0, 0) == 370
        ref = _handle(mt)
        try:
            ref.run_inner_loop(mt.projection_filter(0), 0, 0)
            _same(g.results(), ref.results())
        finally:
            ref.close()
    finally:
        g.close()


def test_state_of_another_search_is_refused(prepared):
    mt = prepared
    a = _handle(mt)
    b = _handle(mt, orientation_limit=100)
    try:
        a.run_partial(mt.projection_filter(0), 0, 0, 0, 16)
        with pytest.raises(core.TMError):
            b.import_state(a.export_state())
        with pytest.raises(core.TMError):
            b.import_state(b"not a state")
    finally:
        a.close()
        b.close()
