"""GPU: 16-bit storage of the inter-pass plane (FP16, BF16) at image sizes other than 4096^2, against
the FP32 search of the same inputs (itself pinned to the oracle by test_gpu_sizes / test_gpu_parity).

The sizes cover every row- and column-pass form: 512^2 and 384 x 324 (generic column pass, gathering
row pass), 2048 x 1536 (4-column tiles: the direct-load row pass reads the 16-bit plane), 2048 x 3072
and 2592^2 (2-column tiles: the gathering row pass reads the 16-bit plane). Two grid positions, 25
in-plane angles (0 to 360 inclusive in 15 deg steps), one defocus plane; bars as in test_gpu_precision (the plane's unit roundoff times the
correlation's size):
  * MIP: FP16 max <= 2e-3, BF16 max <= 2e-2 of max |MIP_fp32|
  * z-score: FP16 max <= 5e-2, BF16 max <= 0.5
  * best orientation equal to FP32 at >= 95% of pixels (the rest are near-ties at 16-bit rounding)
  * histogram: cumulative counts within 1e-3 (FP16) and 1e-2 (BF16) of all values of the FP32 run's
"""
import numpy as np
import pytest

from cistem_downstream_bah_amd import synth
from cistem_downstream_bah_amd import match_template as MT

pytestmark = pytest.mark.gpu

MIP_BAR = {"fp16": 2e-3, "bf16": 2e-2}
Z_BAR = {"fp16": 5e-2, "bf16": 0.5}
HIST_BAR = {"fp16": 1e-3, "bf16": 1e-2}


def _run(s, st):
    se = MT.MatchTemplateSearch(s.micrograph, s.volume,
                                MT.SearchParams(angular_step=10.0, in_plane_angular_step=15.0, padding=1.0,
                                                first_search_position=40, last_search_position=41,
                                                plane_storage=st))
    n = se.run()
    r = se.core.results()
    z = se.core.finalize(n)["zscore"]
    shape = se.core.launch_shape()
    se.core.close()
    return se, n, r, z, shape


@pytest.mark.parametrize("W,H", [(512, 512), (384, 324), (2048, 1536), (2048, 3072), (2592, 2592)])
def test_16bit_planes_at_other_sizes(W, H):
    s = synth.make_synthetic(9000 + H, W, H, 64, 200, 24.0, 8, 0.1, 15000.0, min_dist=64)
    print(f"case: 16-bit planes at {W} x {H}", flush=True)
    se, n32, r32, z32, shape32 = _run(s, "fp32")
    assert (se.W, se.H) == (W, H)
    scale = np.abs(r32.mip).max()
    for st in ("fp16", "bf16"):
        _, n, r, z, shape = _run(s, st)
        assert n == n32 == 2 * se.psi.size and shape["z_tile"] == shape32["z_tile"]
        d = np.abs(r.mip.astype(np.float64) - r32.mip) / scale
        dz = np.abs(z.astype(np.float64) - z32)
        same = (r.best_index == r32.best_index).mean()
        print(f"{W}x{H} {st} (z tile {shape['z_tile']}): MIP err max {d.max():.3e} rms "
              f"{np.sqrt((d ** 2).mean()):.3e}; z err max {dz.max():.3e}; best index equal {same * 100:.2f}%",
              flush=True)
        assert d.max() <= MIP_BAR[st] and dz.max() <= Z_BAR[st]
        assert same >= 0.95
        # the histogram: its cumulative counts move by the values that 16-bit rounding carries across
        # a bin edge (bins are 0.068 z wide), a small fraction of every count
        dc = np.abs(np.cumsum(r.hist) - np.cumsum(r32.hist)).max() / float(W * H * n)
        print(f"{W}x{H} {st}: histogram cumulative difference {dc:.2e} of the values", flush=True)
        assert dc <= HIST_BAR[st]
