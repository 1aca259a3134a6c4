"""CPU: the host logic of the 4096^2 proposal study (cistem_downstream_bah_amd/proposal_large.py).

* rotation_quaternions is synth.euler_matrix's rotation: the matrix of each quaternion (the textbook
  quaternion-to-matrix formula) equals euler_matrix at random and special angles, and the rotation angle
  2 acos |q1 . q2| equals the angle of M1^T M2 from its trace
* the judge's orientations are the first 200,000 of the 2.5 / 1.5 deg C1 grid in the search's loop
  order (position-major, then psi), the last one being psi index 210 of position 829
* orientation_set equals a brute-force scan over all 200,000 judge orientations by matrix angle, at
  the pole, mid-range and the judge's theta edge; a radius of 0 finds an orientation of the grid
* the block means of bin_image / bin_volume against explicit loops
* the dihedral augmentation turns the network input and the location targets alike
"""
import math

import numpy as np
import pytest

from cistem_downstream_bah_amd import match_template as MT
from cistem_downstream_bah_amd import proposal_large as PL
from cistem_downstream_bah_amd import synth


def quat_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def matrix_angle(a, b):
    c = (np.trace(a.T @ b) - 1.0) / 2.0
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def test_quaternions_are_euler_matrix():
    rng = np.random.default_rng(0)
    ang = np.concatenate([rng.uniform([0, 0, 0], [360, 180, 360], (50, 3)),
                          [[0, 0, 0], [90, 0, 0], [0, 90, 0], [0, 0, 90], [10, 180, 20], [359, 42.5, 1.5]]])
    q = PL.rotation_quaternions(ang[:, 0], ang[:, 1], ang[:, 2])
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
    mats = [synth.euler_matrix(*a) for a in ang]
    for qi, m in zip(q, mats):
        assert np.abs(quat_matrix(qi) - m).max() < 1e-12
    for i in range(0, len(ang) - 1):
        d_q = math.degrees(2 * math.acos(min(1.0, abs(float(q[i] @ q[i + 1])))))
        assert abs(d_q - matrix_angle(mats[i], mats[i + 1])) < 1e-6


@pytest.fixture(scope="module")
def judge():
    return PL.JudgeOrientations(PL.LargeConfig())


def test_judge_orientations_are_the_search_prefix(judge):
    grid = MT.search_grid("C1", 2.5)
    psi = MT.psi_list(0.0, 360.0, 1.5)
    assert psi.size == 241 and judge.orients.shape == (200_000, 3)
    assert judge.judge_last_position == 829
    assert tuple(judge.orients[-1]) == (grid[829, 0], grid[829, 1], psi[210])
    assert tuple(judge.orients[241 * 5 + 7]) == (grid[5, 0], grid[5, 1], psi[7])
    assert judge.orients[:, 1].max() == np.float32(42.5)


@pytest.mark.parametrize("angles,radius", [([[0.0, 0.0, 33.0]], 7.5), ([[120.0, 20.0, 200.0], [123.0, 22.0, 205.0]], 7.5),
                                           ([[200.0, 44.0, 10.0]], 6.0), ([[17.0, 30.0, 359.0]], 10.0)])
def test_orientation_set_equals_brute_force(judge, angles, radius):
    angles = np.array(angles)
    got = judge.orientation_set(angles, radius)
    # brute force over every judge orientation, by matrix angle (slow path: vectorised matrices)
    o = judge.orients.astype(np.float64)
    d2r = np.pi / 180
    cp, sp = np.cos(o[:, 0] * d2r), np.sin(o[:, 0] * d2r)
    ct, st = np.cos(o[:, 1] * d2r), np.sin(o[:, 1] * d2r)
    cs, ss = np.cos(o[:, 2] * d2r), np.sin(o[:, 2] * d2r)
    M = np.empty((o.shape[0], 3, 3))
    M[:, 0, 0] = cp * ct * cs - sp * ss
    M[:, 1, 0] = sp * ct * cs + cp * ss
    M[:, 2, 0] = -st * cs
    M[:, 0, 1] = -cp * ct * ss - sp * cs
    M[:, 1, 1] = -sp * ct * ss + cp * cs
    M[:, 2, 1] = st * ss
    M[:, 0, 2] = st * cp
    M[:, 1, 2] = st * sp
    M[:, 2, 2] = ct
    assert np.abs(M[12345] - synth.euler_matrix(*o[12345])).max() < 1e-12
    want = np.zeros(o.shape[0], bool)
    for a in angles:
        m = synth.euler_matrix(*a)
        c = (np.einsum("ij,nij->n", m, M) - 1.0) / 2.0
        want |= np.degrees(np.arccos(np.clip(c, -1, 1))) <= radius
    want_idx = np.nonzero(want)[0]
    # the two tests differ only for orientations at the radius to rounding
    diff = np.setxor1d(got, want_idx)
    for i in diff:
        d = min(matrix_angle(synth.euler_matrix(*a), M[i]) for a in angles)
        assert abs(d - radius) < 1e-6, (i, d)
    assert want_idx.size >= 10


def test_orientation_set_radius_zero_hits_the_grid(judge):
    i = 77_777
    got = judge.orientation_set(judge.orients[i:i + 1].astype(np.float64), 1e-4)
    assert i in got.tolist() and got.size <= 2
    assert judge.judge_index({"phi": judge.orients[i, 0], "theta": judge.orients[i, 1], "psi": judge.orients[i, 2]}) == i


def test_binning():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((24, 16)).astype(np.float32)
    b = PL.bin_image(a, 4)
    want = np.array([[a[4 * y:4 * y + 4, 4 * x:4 * x + 4].astype(np.float64).mean() for x in range(4)] for y in range(6)])
    assert b.shape == (6, 4) and np.abs(b - want).max() < 1e-6
    v = rng.standard_normal((8, 8, 8)).astype(np.float32)
    bv = PL.bin_volume(v, 2)
    assert bv.shape == (4, 4, 4)
    assert abs(bv[1, 2, 3] - v[2:4, 4:6, 6:8].astype(np.float64).mean()) < 1e-6


def test_augmentation_turns_inputs_and_targets_alike():
    """A bright spot in the input and the target cell under it stay together through every turn."""
    G, S = 8, 64

    class Net:
        def __init__(self):
            self.seen = []

        def forward(self, x, want_logits=False):
            self.seen.append(x.copy())

        def backward(self, heat, bins, pos_weight):
            self.seen.append((heat.copy(), bins.copy()))
            return (0.0, 0.0), None

        def adam_step(self, lr):
            pass

    st = PL.LargeStudy.__new__(PL.LargeStudy)
    st.cfg = PL.LargeConfig(size=S * 8, grid=G)
    st.dirs = None
    img = np.zeros((S, S), np.float32)
    img[10, 50] = 1.0  # cell (10 // 8, 50 // 8) = (1, 6) of the 8 x 8 grid
    peaks = [dict(x=50 * 8 + 3, y=10 * 8 + 3, phi=0.0, theta=0.0, psi=0.0)]
    net = Net()
    st.train(net, [img] * 16, [peaks] * 16, epochs=1, batch=16, augment=True)
    x, (heat, bins) = net.seen
    for k in range(16):
        yy, xx = np.unravel_index(np.argmax(x[k]), x[k].shape)
        assert heat[k, yy // 8, xx // 8] == 1.0 and heat[k].sum() == 1.0
        assert (bins[k] >= 0).sum() in (0, 1)
