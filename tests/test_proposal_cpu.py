"""CPU: the host logic of the learned proposals (cistem_downstream_bah_amd/proposal.py), no GPU.

* view directions are unit vectors spread over the sphere; every one of the 600 bins holds grid
  orientations of the 10/10 degree search (and of the coarse 20/20 degree grid)
* orientation bins: psi 360 and 0 share a bin, psi bins are 36 degrees wide, an orientation's bin
  direction is the nearest of the 60 directions (brute force)
* proposals: top-k cells by location logit, none closer than min_cells to a better one, each with
  its top-m bins by orientation logit
* targets: one location label per peak at its cell, the bin of its orientation there
* match / recall: nearest detection within the radius, orientation compared exactly
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import proposal as PR


@pytest.fixture(scope="module")
def table():
    return PR.OrientationTable(10.0, 10.0)


def test_directions_and_bins_cover_the_grid(table):
    d = table.dirs
    assert d.shape == (PR.N_DIRECTIONS, 3) and np.allclose(np.linalg.norm(d, axis=1), 1.0)
    dots = d @ d.T
    np.fill_diagonal(dots, -1)
    assert np.degrees(np.arccos(dots.max(axis=1))).max() < 30.0   # no direction isolated
    counts = np.array([m.size for m in table.members])
    assert counts.sum() == table.orients.shape[0] == 15466 and counts.min() >= 1
    assert set(np.unique(table.coarse_bin)) <= set(range(PR.N_BINS))


def test_orientation_bin_rules(table):
    dirs = table.dirs
    assert PR.orientation_bin(30.0, 40.0, 360.0, dirs) == PR.orientation_bin(30.0, 40.0, 0.0, dirs)
    b = [int(PR.orientation_bin(30.0, 40.0, p, dirs)) % PR.N_PSI_BINS for p in (0.0, 35.9, 36.0, 71.9, 72.0, 359.9)]
    assert b == [0, 0, 1, 1, 2, 9]
    rng = np.random.default_rng(0)
    phi, theta = rng.uniform(0, 360, 200), rng.uniform(0, 180, 200)
    got = PR.orientation_bin(phi, theta, np.zeros(200), dirs) // PR.N_PSI_BINS
    v = PR.view_vector(phi, theta)
    brute = [int(np.argmin(np.linalg.norm(dirs - v[i], axis=1))) for i in range(200)]
    assert list(got) == brute


def test_propose_suppresses_neighbours_and_ranks_bins():
    G, nb = 16, 12
    logits = np.full((G, G, 1 + nb), -5.0, np.float32)
    for (y, x, v) in [(3, 3, 9.0), (3, 4, 8.5), (10, 12, 7.0), (5, 5, 6.0), (14, 1, 5.0)]:
        logits[y, x, 0] = v
    logits[3, 3, 1 + 7] = 4.0
    logits[3, 3, 1 + 2] = 3.0
    props = PR.propose(logits, k=3, m=2, min_cells=3)
    cells = [c for c, _ in props]
    assert cells == [(3, 3), (10, 12), (14, 1)]    # (3, 4) and (5, 5) lie within 3 cells of (3, 3)
    assert list(props[0][1]) == [7, 2]
    for i, (a, _) in enumerate(props):
        for b, _ in props[:i]:
            assert max(abs(a[0] - b[0]), abs(a[1] - b[1])) >= 3


def test_targets_mark_peak_cells(table):
    peaks = [dict(x=100.0, y=37.0, phi=10.0, theta=20.0, psi=30.0), dict(x=5.0, y=500.0, phi=200.0, theta=100.0, psi=300.0)]
    heat, bins = PR.targets(peaks, 64, 512, table.dirs)
    assert heat.sum() == 2 and heat[37 // 8, 100 // 8] == 1 and heat[500 // 8, 5 // 8] == 1
    assert bins[37 // 8, 100 // 8] == PR.orientation_bin(10.0, 20.0, 30.0, table.dirs)
    assert (bins >= 0).sum() == 2


def test_match_and_recall():
    peaks = [dict(x=10.0, y=10.0, phi=1.0, theta=2.0, psi=3.0), dict(x=100.0, y=50.0, phi=4.0, theta=5.0, psi=6.0),
             dict(x=300.0, y=300.0, phi=7.0, theta=8.0, psi=9.0)]
    dets = [PR.Detection(11.0, 11.0, 0.5, 1.0, 2.0, 3.0), PR.Detection(102.5, 50.0, 0.4, 4.0, 5.0, 0.0),
            PR.Detection(100.0, 51.0, 0.3, 4.0, 5.0, 6.0)]
    m = PR.match(dets, peaks, 2.0)
    assert m == [(True, True), (True, True), (False, False)]   # the nearer of two candidates is compared
    rec, same, n = PR.recall(dets, peaks, 2.0)
    assert (rec, same, n) == (pytest.approx(2 / 3), 1.0, 3)
    assert PR.recall(dets, peaks, 0.5)[0] == 0.0
