"""The C-ABI library on a machine without a GPU: it loads, exports every
entry point include/hamstr_hip.h declares, runs its host-only set-up
(MESHBLOCK::preprocess) bit-identically to the oracle, and refuses device
work with HAMSTR_ERR_NODEV instead of falling back to the CPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from hamstr_oop_amd import hip, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "hamstr_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|long|const char \*)\s*(hamstr_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(hip.EXPORTS) == declared


@pytest.mark.parametrize("mk", [lambda: meshgen.circle_ogrid(24, 10), lambda: meshgen.naca_ogrid(40, 12)])
def test_preprocess_matches_oracle(oracle, mk):
    mesh = mk()
    f1, n1, r1 = hip.preprocess(mesh.x, mesh.faces)
    f2, n2, r2 = oracle.preprocess(mesh.x, mesh.faces)
    assert np.array_equal(f1, f2) and np.array_equal(n1, n2) and np.array_equal(r1, r2)


def test_preprocess_matches_golden():
    d = np.load(os.path.join(ROOT, "tests", "golden", "case_naca16x8_o3.npz"))
    f, nv, rm = hip.preprocess(d["x"], d["faces_in"])
    assert np.array_equal(f, d["faces"]) and np.array_equal(nv, d["normvec"]) and np.array_equal(rm, d["refmtx"])


def test_no_device_is_an_error_not_a_fallback():
    try:
        s = hip.Solver(meshgen.circle_ogrid(8, 4), 0.3, 0.0, 1, 10.0)
    except hip.HamstrError as e:
        assert e.code == hip.ERR_NODEV
        return
    s.close()
    pytest.skip("a GPU is present")


def test_bad_arguments_are_rejected():
    with pytest.raises(hip.HamstrError) as e:
        hip.preprocess(np.zeros((0, 2)), np.zeros((0, 6), np.int32))
    assert e.value.code == hip.ERR_ARG
    L = hip.lib()
    # null contexts and handles are errors, not crashes (no device needed)
    assert L.hamstr_set_retire(None, 1, 0) == hip.ERR_ARG
    assert L.hamstr_set_graph(None, 1) == hip.ERR_ARG
    assert L.hamstr_case_status(None, None, None) == hip.ERR_ARG
    assert L.hamstr_mlp_forward(None, 1, None, None) == hip.ERR_ARG
    assert L.hamstr_warm_start(None, None, None, None, None) == hip.ERR_ARG
    assert L.hamstr_write_solution(None, b"x", 0) == hip.ERR_ARG
    assert L.hamstr_set_face_staging(None, 1) == hip.ERR_ARG
    assert L.hamstr_face_staging(None, None) == hip.ERR_ARG
    assert L.hamstr_read_solution(None, b"x", None) == hip.ERR_ARG
    assert L.hamstr_set_scheme(None, 1) == hip.ERR_ARG
    assert L.hamstr_get_scheme(None) == -1
    assert L.hamstr_update_solution(None, 0, 6, ctypes.c_double(1.0)) == hip.ERR_ARG


def test_driver_built():
    assert os.path.exists(os.path.join(ROOT, "hamstr_oop_amd", "hamstr_mi355x"))
