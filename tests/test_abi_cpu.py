"""CPU: the C-ABI library loads and exports every symbol include/tm_hip.h declares; host logic
that needs no GPU (header/export agreement, MRC round trip, synthetic generator conventions)."""
import ctypes
import os
import re

import numpy as np

from cistem_downstream_bah_amd import cnn, core, mrc, refine, results, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "tm_hip.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(tm_\w+)\(", header, re.M))
    assert declared == set(core.EXPORTS)
    L = ctypes.CDLL(core.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name


def test_refine_exports_and_structs_match_header():
    header = open(os.path.join(ROOT, "include", "tm_refine.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(tm_\w+)\(", header, re.M))
    prep = open(os.path.join(ROOT, "include", "tm_prepare.h")).read()
    assert "tm_prepare_refine_image" in set(re.findall(r"^(?:int|const char\*)\s+(tm_\w+)\(", prep, re.M))
    assert declared | {"tm_prepare_refine_image"} == set(refine.REFINE_EXPORTS)
    L = ctypes.CDLL(core.LIB_PATH)
    for name in refine.REFINE_EXPORTS:
        assert hasattr(L, name), name
    for struct, cls in (("tm_refine_params", refine.RefineParams), ("tm_refined_peak", refine.RefinedPeak)):
        body = header[header.index("typedef struct " + struct):header.index("} " + struct + ";")]
        fields = re.findall(r"^\s+(?:int|double|float|long long)\s+([\w, ]+);", body, re.M)
        names = [f.strip() for group in fields for f in group.split(",")]
        assert names == [f[0] for f in cls._fields_], struct
    assert ctypes.sizeof(refine.RefinedPeak) == 104


def test_cnn_exports_and_config_match_header():
    header = open(os.path.join(ROOT, "include", "tm_cnn.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(tm_\w+)\(", header, re.M))
    assert declared == set(cnn.CNN_EXPORTS)
    L = ctypes.CDLL(core.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    body = header[header.index("typedef struct tm_cnn_config"):header.index("} tm_cnn_config;")]
    fields = re.findall(r"^\s+int\s+(\w+)", body, re.M)
    assert fields == [f[0] for f in cnn.CnnConfig._fields_]
    assert ctypes.sizeof(cnn.CnnConfig) == 4 * (4 + 2 * cnn.MAX_LAYERS)
    Lc = cnn._lib()
    assert Lc.tm_cnn_destroy(None) == 0
    assert Lc.tm_cnn_forward(None, None, 1, None) != 0 and Lc.tm_cnn_last_error()


def test_result_record_exports_match_header():
    header = open(os.path.join(ROOT, "include", "tm_results.h")).read()
    declared = set(re.findall(r"^(?:int|long long|double|const char\*)\s+(tm_\w+)\(", header, re.M))
    assert declared == set(results.RESULT_EXPORTS)
    L = ctypes.CDLL(core.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert "#define TM_RESULT_META 6" in header and "#define TM_RESULT_MAPS 8" in header
    assert (results.META, results.MAPS) == (6, 8)


def test_refine_errors_are_returned_not_fatal():
    L = refine._lib()
    assert L.tm_refine_destroy(None) == 0
    n = ctypes.c_int(0)
    rc = L.tm_refine_run(None, None, None, None, 0, ctypes.byref(n))
    assert rc != 0 and L.tm_refine_last_error()


def test_init_params_struct_matches_header():
    header = open(os.path.join(ROOT, "include", "tm_hip.h")).read()
    body = header[header.index("typedef struct tm_init_params"):header.index("} tm_init_params;")]
    fields = re.findall(r"^\s+(?:int|double|float|long long)\s+([\w, ]+);", body, re.M)
    names = [f.strip() for group in fields for f in group.split(",")]
    assert names == [f[0] for f in core.InitParams._fields_]


def test_errors_are_returned_not_fatal():
    L = core.lib()
    assert L.tm_destroy(None) == 0
    rc = L.tm_run_inner_loop(None, None, 0, 0, core.PROGRESS_FN(), None, None)
    assert rc != 0 and L.tm_last_error()


def test_mrc_round_trip(tmp_path):
    a = np.random.default_rng(0).standard_normal((3, 5, 7)).astype(np.float32)
    p = str(tmp_path / "a.mrc")
    mrc.write_mrc(p, a, 1.5)
    assert np.array_equal(mrc.read_mrc(p), a)


def test_synthetic_projection_convention():
    """A single atom at a (relative to the box centre) projects to (R^T a)_xy."""
    pos = np.array([[5.0, -3.0, 7.0]])
    amp = np.array([1.0])
    m = synth.euler_matrix(30.0, 60.0, 100.0)
    img = np.zeros((64, 64))
    synth.project_copy(pos, amp, m, 32.0, 32.0, 64, 64, 40, 1.5, img)
    iy, ix = np.unravel_index(np.argmax(img), img.shape)
    want = m.T @ pos[0]
    assert abs(ix - (32 + want[0])) <= 0.5 and abs(iy - (32 + want[1])) <= 0.5
    # integral of the projected atom: amplitude * sigma sqrt(2 pi) (along z) * 2 pi sigma^2 (in-plane)
    assert abs(img.sum() - 1.5 * np.sqrt(2 * np.pi) * 2 * np.pi * 1.5 ** 2) < 1e-3


def test_new_entry_points_reject_bad_arguments():
    """Entry points without a reference counterpart return an error code (never crash) on NULL
    handles and invalid arguments; none of these calls reaches a GPU."""
    L = core.lib()
    n = ctypes.c_int(0)
    assert L.tm_set_graph(None, 1) != 0
    assert L.tm_set_input(None, None) != 0
    assert L.tm_set_micrograph(None, None, None, None) != 0
    assert L.tm_find_peaks(None, 1, 0.0, 0, 10.0, 1.0, None, 0, ctypes.byref(n)) != 0
    assert L.tm_run_partial(None, None, 0, 0, 0, 1, None) != 0
    size = ctypes.c_longlong(0)
    assert L.tm_state_size(None, ctypes.byref(size)) != 0
    assert L.tm_export_state(None, None, 0) != 0
    assert L.tm_import_state(None, None, 0) != 0
    assert L.tm_autotune_batch(0, None, None, None, None, None, None, None, 0, 0, None, None) != 0
    R = refine._lib()
    assert R.tm_refine_set_micrograph(None, None) != 0
    assert R.tm_refine_set_image(None, None, None, None, 0) != 0
    assert R.tm_refine_score_batch(None, 0, None, None, 0.0, None) != 0
    assert R.tm_refine_get_image(None, None, None, None) != 0
    C = cnn._lib()
    h = ctypes.c_void_p()
    assert C.tm_cnn_create(0, None, ctypes.byref(h)) != 0
    bad = cnn.CnnConfig(512, 0, (ctypes.c_int * cnn.MAX_LAYERS)(), (ctypes.c_int * cnn.MAX_LAYERS)(), 600, 1)
    assert C.tm_cnn_create(0, ctypes.byref(bad), ctypes.byref(h)) != 0 and not h.value
    assert C.tm_cnn_heat(None, None) != 0
    assert C.tm_cnn_cell_logits(None, None, 0, None) != 0
    assert C.tm_cnn_adam_step(None, 1e-3, 0.9, 0.999, 1e-8) != 0
    assert C.tm_cnn_forward_device(None, None, 1) != 0
    assert C.tm_cnn_backward_device(None, None, None, 0, 1.0) != 0
    assert C.tm_cnn_loss(None, None) != 0 and C.tm_cnn_synchronize(None) != 0
    assert C.tm_cnn_last_error()
