"""GPU: graph mode of the search (tm_set_graph; no counterpart in the reference). One plane's
launches are captured into a hipGraph and replayed for every later plane, with the plane's index
offset read from device memory.

* five planes (four defocus values, one repeated) through a graph-mode handle equal, plane by plane,
  a launch-by-launch handle bit for bit: MIP, best index, defocus map, FP64 sum and sum of squares,
  histogram; eager work on the graph handle's stream (a debug correlation) and a synchronisation sit
  between the replays
* a second tm_init drops the graph; the handle captures again and stays equal
* tm_set_input: after searching one micrograph, a graph-mode handle searches another of the same
  size exactly as a handle made for that micrograph
* tm_set_micrograph: the input preparation on the device gives the host path's whitening curve
  (1e-5) and a search within 1e-5 of its MIP and sums, best index identical on >= 99% of pixels
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import synth
from cistem_downstream_bah_amd import match_template as MT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0():
    return synth.config0(0)


def _search(s, graph):
    se = MT.MatchTemplateSearch(s.micrograph, s.volume,
                                MT.SearchParams(angular_step=10.0, in_plane_angular_step=10.0, padding=1.0,
                                                defocus_search_range=200.0, defocus_step=100.0,
                                                first_search_position=0, last_search_position=59, batch_size=64))
    se.core.set_graph(graph)
    return se


def _same(a, b):
    ra, rb = a.core.results(), b.core.results()
    for f in ("mip", "best_index", "defocus", "sum", "sumsq", "hist"):
        assert np.array_equal(getattr(ra, f), getattr(rb, f)), f
    assert ra.nccf == rb.nccf


def test_graph_replays_equal_eager(c0):
    eager, graph = _search(c0, False), _search(c0, True)
    for k, d in enumerate((-1, 0, 2, 0, 1)):
        print(f"case: config0 positions 0-59, plane {k} defocus {d}, graph vs eager", flush=True)
        f = eager.filter_for(d)
        assert eager.core.run_inner_loop(f, 0, d) == graph.core.run_inner_loop(f, 0, d)
        _same(eager, graph)
        graph.core.debug_ccf(f, 3, 2)   # eager launches on the graph handle's stream between replays
    print("case: re-init drops the graph", flush=True)
    for se in (eager, graph):
        se.core.init(se.template_spectrum, se.input_spectrum, se.T, se.grid, se.psi, se.first, se.last,
                     se.histogram_min_scaled, se.histogram_step_scaled, 512, se.max_padding, se.defocus_step,
                     se.pixel_size_step, 64, "fp32")
    for d in (0, -2):
        f = eager.filter_for(d)
        eager.core.run_inner_loop(f, 0, d)
        graph.core.run_inner_loop(f, 0, d)
        _same(eager, graph)


def test_set_input_equals_a_fresh_search(c0):
    """tm_set_input: a handle (in graph mode) that searched one micrograph searches another of the
    same size exactly as a handle made for it."""
    print("case: set_micrograph + graph vs a fresh handle", flush=True)
    other = synth.config0(7).micrograph
    reused, fresh = _search(c0, True), None
    try:
        f0 = reused.filter_for(0)
        reused.core.run_inner_loop(f0, 0, 0)
        reused.set_micrograph(other)
        fresh = MT.MatchTemplateSearch(other, c0.volume, reused.p)
        for d in (0, 1):
            reused.core.run_inner_loop(reused.filter_for(d), 0, d)
            fresh.core.run_inner_loop(fresh.filter_for(d), 0, d)
            assert np.array_equal(reused.filter_for(d), fresh.filter_for(d))
        _same(reused, fresh)
    finally:
        reused.core.close()
        if fresh is not None:
            fresh.core.close()


def test_device_preparation_search(c0):
    """tm_set_micrograph: the input preparation on the device gives the host preparation's curve and
    a search within the parity bars of a handle prepared on the host."""
    print("case: set_micrograph on the device vs host preparation", flush=True)
    other = synth.config0(7).micrograph
    dev, host = _search(c0, False), None
    try:
        dev.set_micrograph(other, on_device=True)
        host = _search(synth.Synthetic(other, c0.volume, []), False)
        assert np.array_equal(dev.curve_x, host.curve_x)
        e_curve = float(np.abs(dev.curve_y - host.curve_y).max() / np.abs(host.curve_y).max())
        print(f"curve {e_curve:.2e}", flush=True)
        assert e_curve <= 1e-5
        for d in (0, 1):
            dev.core.run_inner_loop(dev.filter_for(d), 0, d)
            host.core.run_inner_loop(host.filter_for(d), 0, d)
        ra, rb = dev.core.results(), host.core.results()
        e_mip = float(np.abs(ra.mip - rb.mip).max() / np.abs(rb.mip).max())
        same = float(np.mean(ra.best_index == rb.best_index))
        e_sum = float(np.abs(ra.sum - rb.sum).max() / np.abs(rb.sum).max())
        print(f"MIP {e_mip:.2e}, index identical {same:.4f}, sum {e_sum:.2e}", flush=True)
        assert e_mip <= 1e-5 and same >= 0.99 and e_sum <= 1e-5
        assert np.array_equal(ra.hist.sum(), rb.hist.sum())
    finally:
        dev.core.close()
        if host is not None:
            host.core.close()
