"""GPU: the C++ match_template driver (cistem_downstream_bah_amd/match_template_hip) end to end on the
config-0 inputs written as MRC files: its maps equal the C-ABI path's, its peak list (found on the
device) equals the host peak rule applied to its own output maps, it contains every inserted
particle above the threshold, its histogram file holds the search's counts with the survival and
expected-survival columns of tm_histogram_table and the master's expected threshold, and its worker
record (--worker-record, include/tm_results.h) agrees with the Python mirror's record of the same
search."""
import os
import subprocess

import numpy as np
import pytest

from cistem_downstream_bah_amd import match_template as MT
from cistem_downstream_bah_amd import mrc, peaks as PK, results as R, synth

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cistem_downstream_bah_amd",
                      "match_template_hip")


def test_driver_end_to_end(tmp_path):
    s = synth.config0(0)
    img, vol = str(tmp_path / "img.mrc"), str(tmp_path / "vol.mrc")
    mrc.write_mrc(img, s.micrograph)
    mrc.write_mrc(vol, s.volume)
    out = str(tmp_path / "run")
    r = subprocess.run([DRIVER, "--image", img, "--template", vol, "--out", out, "--angular-step", "10",
                        "--in-plane-step", "10", "--min-peak-radius", "10", "--particle-radius", "16",
                        "--worker-record", out + "_record.f32"], capture_output=True, text=True,
                       timeout=240)
    print(r.stderr, flush=True)
    assert r.returncode == 0
    z = mrc.read_mrc(out + "_scaled_mip.mrc")
    # same search through the Python mirror of the driver
    search = MT.MatchTemplateSearch(s.micrograph, s.volume, MT.SearchParams(angular_step=10.0,
                                                                            in_plane_angular_step=10.0))
    n = search.run()
    ref = search.core.finalize(n)
    assert np.abs(z - ref["zscore"]).max() <= 1e-5 * max(1.0, np.abs(ref["zscore"]).max())
    peaks = np.loadtxt(out + "_peaks.txt", ndmin=2)
    thr = float(open(out + "_histogram.txt").readline().split("=")[1])
    detectable = 0
    for pt in s.particles:
        x, y = int(round(pt.x)), int(round(pt.y))
        if z[y - 2:y + 3, x - 2:x + 3].max() < thr:
            continue  # this copy's correlation peak is below the expected threshold at SNR 0.1
        detectable += 1
        assert any(abs(px - pt.x) <= 2 and abs(py - pt.y) <= 2 for px, py in peaks[:, :2]), (pt.x, pt.y)
    print(f"{len(peaks)} peaks, {detectable} of {len(s.particles)} inserted copies above {thr}", flush=True)
    # the driver found its peaks on the device (no padding or rotation here): the host rule on its
    # own output maps gives the same list
    maps = {k: mrc.read_mrc(out + f"_{k}.mrc") for k in ("psi", "theta", "phi", "defocus")}
    host = PK.find_peaks(z, thr, 64, 10, maps=maps)
    assert len(host) == len(peaks)
    for row, h in zip(peaks, host):
        assert (row[0], row[1]) == (h["x"], h["y"])
        assert abs(row[7] - h["peak_height"]) <= 1e-5 * abs(h["peak_height"])
        assert np.allclose(row[2:6], [h["psi"], h["theta"], h["phi"], h["defocus"]], rtol=0, atol=1e-5)
    assert detectable >= 8
    hist = np.loadtxt(out + "_histogram.txt")
    assert hist.shape == (512, 4) and hist[:, 1].sum() == n * 512 * 512
    # the file's columns are the histogram table of its own counts (match_template.cpp:1333-1364) and
    # its threshold line the master's formula (test_threshold_cpu pins both)
    assert thr == round(R.expected_threshold(512, 512, n), 2)
    counts = hist[:, 1].astype(np.int64)
    assert np.array_equal(counts, search.core.results().hist)
    snr, surv, expected = R.histogram_table(counts, -12.5, float(np.float32(35.0 / 512)), 512.0, n)
    assert np.allclose(hist[:, 0], snr, rtol=0, atol=1e-6)
    assert np.array_equal(hist[:, 2], surv)
    assert np.allclose(hist[:, 3], expected, rtol=1e-6, atol=1e-6)
    # the driver's worker record (--worker-record) against the Python mirror's for the same search
    u = R.unpack(np.fromfile(out + "_record.f32", np.float32))
    py = R.unpack(search.worker_record(16.0))
    assert (u["x"], u["y"], u["n"], u["bins"], u["n_ccfs"]) == (py["x"], py["y"], py["n"], py["bins"], py["n_ccfs"])
    assert u["n_ccfs"] == n and u["sqrt_input_pixels"] == np.float32(512.0)
    assert np.abs(u["mip"] - py["mip"]).max() <= 1e-5 * np.abs(py["mip"]).max()
    assert (u["psi"] == py["psi"]).mean() > 0.99 and (u["theta"] == py["theta"]).mean() > 0.99
    assert u["histogram"].sum() == py["histogram"].sum() == n * 512 * 512


def test_job_arguments_and_interactive_answers_give_the_named_run(tmp_path):
    """The reference's argument forms (--job-args: the 42 job arguments; --interactive: its 32
    questions on stdin) run the same search as the named options: identical maps and histogram file."""
    s = synth.config0(0)
    img, vol = str(tmp_path / "img.mrc"), str(tmp_path / "vol.mrc")
    mrc.write_mrc(img, s.micrograph)
    mrc.write_mrc(vol, s.volume)
    named = str(tmp_path / "named")
    print("case: driver named options, config 0", flush=True)
    r = subprocess.run([DRIVER, "--image", img, "--template", vol, "--out", named, "--angular-step", "10",
                        "--in-plane-step", "10", "--defocus1", "15000"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    kinds = ["mip", "scaled_mip", "psi", "theta", "phi", "defocus", "pixel_size", "avg", "std"]

    def names(p):
        return [f"{p}_{k}.mrc" for k in kinds]

    ja = str(tmp_path / "ja")
    m = names(ja)
    job = [img, vol, "1.0", "300", "2.7", "0.07", "15000", "15000", "0", "300", "8", "10", "20", "0", "0", "0", "0",
           "1.0", "No", "0", "0", m[0], m[2], m[3], m[4], m[5], m[6], m[1], m[7], "C1", "10", ja + "_histogram.txt",
           "-1", "-1", "0", "0", m[8], "/dev/null", "/dev/null", "10", "Yes", "1"]
    print("case: driver --job-args, config 0", flush=True)
    r = subprocess.run([DRIVER, "--job-args", *job], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    it = str(tmp_path / "it")
    m = names(it)
    answers = [img, vol, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], it + "_histogram.txt", "1.0", "300",
               "2.7", "0.07", "15000", "15000", "0", "0", "8", "10", "10", "0", "0", "0", "0", "1", "0", "C1", "Yes", "1"]
    print("case: driver --interactive, config 0", flush=True)
    r = subprocess.run([DRIVER, "--interactive"], input="\n".join(answers) + "\n", capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    for k in kinds:
        a = mrc.read_mrc(f"{named}_{k}.mrc")
        for other in (ja, it):
            assert np.array_equal(a, mrc.read_mrc(f"{other}_{k}.mrc")), (other, k)
    h = open(named + "_histogram.txt").read()
    assert h == open(ja + "_histogram.txt").read() == open(it + "_histogram.txt").read()


@pytest.mark.parametrize("mx,my,copies", [(3072, 2048, 300), (470, 360, 5)])
def test_driver_unrotates_and_crops_to_the_micrograph(tmp_path, mx, my, copies):
    """The driver on micrographs the size policy rotates (3072 x 2048 -> 2048 x 3072) or pads
    (470 x 360 -> 512 x 384, match_template.cpp:527-598): its output maps come back in the
    micrograph's own size and orientation (Rotate2DInPlaceBy90Degrees(false) and Resize,
    image.cpp:11104-11141, 8825-8838). The ground truth is the insertion itself: the z-score map is
    higher at the inserted copies' true positions than where an un-rotation in the wrong sense
    (180 degrees off) or a crop off by the padding would put them (there it is the MIP of noise,
    z ~ 4), and the peaks the driver lists lie on copies. Both micrographs hold copies over about a
    tenth of their area at SNR 1 (noise of the frame's signal variance), where copies of this template
    reach z 6-8 on the 10 deg grid; every copy above the threshold is a listed peak, and at most a few
    listed peaks (noise, or the offset peaks of touching copies) lie off the copies."""
    s = synth.make_synthetic(6100 + mx, mx, my, 64, 200, 24.0, copies, 1.0, 15000.0, min_dist=64)
    W, H, rotated, _, _ = MT.prepare_sizes(mx, my, 64)
    assert rotated == (mx == 3072) and (W, H) != (mx, my)
    img, vol = str(tmp_path / "img.mrc"), str(tmp_path / "vol.mrc")
    mrc.write_mrc(img, s.micrograph)
    mrc.write_mrc(vol, s.volume)
    out = str(tmp_path / "run")
    print(f"case: driver on {mx}x{my} (search size {W}x{H}, rotated {rotated})", flush=True)
    r = subprocess.run([DRIVER, "--image", img, "--template", vol, "--out", out, "--angular-step", "10",
                        "--in-plane-step", "10", "--min-peak-radius", "10"], capture_output=True, text=True,
                       timeout=240)
    print(r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    z = mrc.read_mrc(out + "_scaled_mip.mrc")
    assert z.shape == (my, mx)
    for name in ("psi", "theta", "phi", "defocus"):
        assert mrc.read_mrc(out + f"_{name}.mrc").shape == (my, mx)

    def z_near(x, y):
        x, y = int(round(x)), int(round(y))
        if not (2 <= x < mx - 2 and 2 <= y < my - 2):
            return np.nan
        return float(z[y - 2:y + 3, x - 2:x + 3].max())

    truth = np.array([(p.x, p.y) for p in s.particles])
    at_truth = np.array([z_near(x, y) for x, y in truth])
    wrong = {"rotated 180 deg": [(mx - 1 - x, my - 1 - y) for x, y in truth],
             "shifted by the padding": [(x + (W - mx) // 2, y + (H - my) // 2) for x, y in truth]}
    print(f"z at the true positions: median {np.median(at_truth):.2f}, range {at_truth.min():.2f} - "
          f"{at_truth.max():.2f}", flush=True)
    for name, pts in wrong.items():
        if rotated != (name == "rotated 180 deg"):
            continue
        at_wrong = np.array([z_near(x, y) for x, y in pts])
        print(f"z {name}: median {np.nanmedian(at_wrong):.2f}", flush=True)
        assert np.nanmedian(at_wrong) <= np.median(at_truth) - 1.0
    peaks = np.loadtxt(out + "_peaks.txt", ndmin=2)
    print(f"{len(peaks)} peaks above {open(out + '_histogram.txt').readline().strip()}", flush=True)
    assert len(peaks) >= 1
    on_copy = [bool((np.abs(truth - (px, py)).max(axis=1) <= 2).any()) for px, py in peaks[:, :2]]
    print(f"{sum(on_copy)} of {len(peaks)} peaks within 2 px of a copy", flush=True)
    # the expected threshold admits a few noise peaks per search, and among 300 copies some peaks of
    # touching copies sit a few pixels off
    assert len(on_copy) - sum(on_copy) <= max(3, 0.1 * len(on_copy)), peaks[:, :2]
    # every copy whose z at its true position reaches the threshold is a listed peak
    thr = float(open(out + "_histogram.txt").readline().split("=")[1])
    for (x, y), zt in zip(truth, at_truth):
        if zt >= thr:
            assert (np.abs(peaks[:, :2] - (x, y)).max(axis=1) <= 2).any(), (x, y)
