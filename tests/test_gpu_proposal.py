"""GPU: learned proposals + local exhaustive refinement (cistem_downstream_bah_amd/proposal.py).

* tm_refine_score_batch equals tm_refine_score request by request (bit-identical), with more
  windows and requests than one batch holds (3 and 7 windows over 4 slots); after tm_refine_set_image a handle scores exactly as a
  handle made for that micrograph
* the one-kernel TemplateScore (k_score_fused, t = 64) equals the multi-kernel path bit for bit over
  300 requests in two windows, and a single score equals its batched twin
* with the exhaustive search's own peaks as proposals and every orientation bin, the local
  refinement finds every peak (recall 1 within 2 pixels) and mostly its orientation: the window
  correlation and the search's full-image correlation agree on position
* a short training run on micrographs labelled by the exhaustive search lowers the loss, and its
  proposed locations recall at least 30 points more of the held-out peaks than random locations
  with the same scoring budget (the gain of the network over the method without it)
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import cnn, proposal as PR
from cistem_downstream_bah_amd.refine import RefineSettings, TemplateRefiner

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def study():
    return PR.ProposalStudy(PR.StudyConfig())


def test_score_batch_equals_single_scores(study):
    print("case: score_batch vs score, 3 windows x 5 orientations, max_in_flight 4", flush=True)
    m = study.micrograph(3)
    r = TemplateRefiner(m, study.volume, RefineSettings(padding=1.0, max_in_flight=4),
                        template_spectrum=study.tpl_spectrum)
    try:
        rng = np.random.default_rng(0)
        wins = [(-100, 40), (12, -7), (200, 190)]
        xy, ang = [], []
        for k in range(15):
            xy.append(wins[k % 3])
            ang.append(study.orients[rng.integers(len(study.orients))])
        got = r.score_batch(np.array(xy), np.array(ang))
        for k in range(15):
            ref = r.score(xy[k][0], xy[k][1], *ang[k])
            assert tuple(got[k]) == tuple(np.float32(v) for v in ref), (k, got[k], ref)
    finally:
        r.close()


def test_score_batch_more_windows_than_in_flight(study):
    """Seven distinct windows with four slots: batches still in flight must drain before their
    window slots are refilled. Fused path (t = 64, padding 1) against one score per request."""
    print("case: score_batch vs score, 7 windows x 3 orientations, max_in_flight 4", flush=True)
    m = study.micrograph(6)
    r = TemplateRefiner(m, study.volume, RefineSettings(padding=1.0, max_in_flight=4),
                        template_spectrum=study.tpl_spectrum)
    try:
        rng = np.random.default_rng(9)
        wins = [(-150, 40), (12, -7), (200, 190), (-60, -180), (90, 33), (-210, 120), (0, 0)]
        xy = np.array([w for w in wins for _ in range(3)])
        ang = study.orients[rng.integers(len(study.orients), size=len(xy))]
        got = r.score_batch(xy, ang)
        for k in range(len(xy)):
            ref = r.score(int(xy[k][0]), int(xy[k][1]), *ang[k])
            assert tuple(got[k]) == tuple(np.float32(v) for v in ref), (k, got[k], ref)
    finally:
        r.close()


def test_fused_scores_equal_the_multi_kernel_path(study):
    print("case: k_score_fused vs the multi-kernel TemplateScore, 2 windows x 300 orientations", flush=True)
    m = study.micrograph(4)
    r = TemplateRefiner(m, study.volume, RefineSettings(padding=1.0, max_in_flight=256),
                        template_spectrum=study.tpl_spectrum)
    try:
        rng = np.random.default_rng(2)
        xy = np.array([(-37, 88), (140, -3)] * 150)
        ang = study.orients[rng.integers(len(study.orients), size=300)]
        fused = r.score_batch(xy, ang)
        r.set_fused(False)
        plain = r.score_batch(xy, ang)
        r.set_fused(True)
        assert np.array_equal(fused.view(np.uint32), plain.view(np.uint32))
        # a single score (one item in the batch) through both paths
        one = r.score(-37, 88, *ang[0])
        assert tuple(np.float32(v) for v in one) == tuple(fused[0])
    finally:
        r.close()


def test_set_image_equals_a_fresh_handle(study):
    print("case: tm_refine_set_image vs a new handle", flush=True)
    a, b = study.micrograph(7), study.micrograph(8)
    settings = RefineSettings(padding=1.0)
    reused = TemplateRefiner(a, study.volume, settings, template_spectrum=study.tpl_spectrum)
    fresh = TemplateRefiner(b, study.volume, settings, template_spectrum=study.tpl_spectrum)
    try:
        reused.set_image(b)
        xy = np.array([(-60, 20), (31, -90), (5, 5)] * 4)
        ang = study.orients[np.arange(12) * 997 % len(study.orients)]
        assert np.array_equal(reused.score_batch(xy, ang), fresh.score_batch(xy, ang))
    finally:
        reused.close()
        fresh.close()


def test_true_cells_with_all_bins_recall_every_peak(study):
    print("case: refinement at the exhaustive peaks' cells, all bins", flush=True)
    m = study.micrograph(5)
    peaks, dt = study.exhaustive(m)
    print(f"{len(peaks)} peaks above the expected threshold in {dt * 1e3:.1f} ms", flush=True)
    assert len(peaks) >= 3
    grid = 64
    cell = study.cfg.size // grid
    props = [((int(p["y"]) // cell, int(p["x"]) // cell), np.arange(PR.N_BINS)) for p in peaks]
    net = cnn.ProposalNet(cnn.NetSpec(), max_batch=1)
    try:
        dets, n, t = study.detect(net, m, 0, 0, proposals=props)
    finally:
        net.close()
    rec, same, _ = PR.recall(dets, peaks)
    print(f"recall {rec:.3f}, same orientation {same:.3f}, {n} TemplateScores in {t * 1e3:.1f} ms", flush=True)
    assert rec == 1.0 and same >= 0.6


def test_trained_locations_beat_random_locations(study):
    print("case: label 96 micrographs, train 6 epochs, network vs random locations", flush=True)
    data = study.labelled(range(100, 196))
    print(f"labelled: {sum(len(d.peaks) for d in data)} peaks, "
          f"{np.mean([d.seconds for d in data]) * 1e3:.1f} ms per exhaustive search (4 in flight)", flush=True)
    net = cnn.ProposalNet(cnn.NetSpec(), max_batch=16, seed=1)
    try:
        hist = study.train(net, data, epochs=6, batch=16, lr=1e-3, log=print)
        assert hist[-1][0] < hist[0][0]
        rng = np.random.default_rng(3)
        hits = {"network": 0, "random": 0}
        total = 0
        for seed in (90001, 90002, 90003, 90004):
            m = study.micrograph(seed)
            peaks, _ = study.exhaustive(m)
            total += len(peaks)
            for kind in hits:
                props = study.random_proposals(rng, net.grid, 16, 20) if kind == "random" else None
                dets, n, t = study.detect(net, m, 16, 20, proposals=props, bins="coarse")
                assert len(dets) == 16 and n > 0
                hits[kind] += sum(f for f, _ in PR.match(dets, peaks))
        print(f"recall over {total} peaks: network {hits['network'] / total:.3f}, "
              f"random {hits['random'] / total:.3f}", flush=True)
        assert total > 0 and hits["network"] >= hits["random"] + 0.3 * total
    finally:
        net.close()
