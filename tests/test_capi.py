"""CPU-side checks of the C ABI: the library loads, exports every function that
include/nrs_hip.h declares, and its host-only setup helpers agree with the
oracle.  No device compute is issued here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nrs_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char|const char)\s*\*?\s*(nrs_\w+)\s*\(",
                       src, flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared_functions()
    for must in ["nrs_ax", "nrs_ogs_setup", "nrs_ogs_gather_scatter", "nrs_elliptic_solve",
                 "nrs_elliptic_operator", "nrs_coarsen", "nrs_prolongate", "nrs_mask"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from nekrs_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers every declared symbol
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_product_has_no_oracle_import():
    pkg = os.path.join(ROOT, "nekrs_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text and "nrs_oracle" not in text, f


@pytest.mark.parametrize("mp,eps", [("none", 0.0), ("sine", 0.0), ("kershaw", 0.3)])
def test_box_mesh_matches_oracle(orc, mp, eps):
    import nekrs_amd
    m = nekrs_amd.BoxMesh(3, 2, 4, 5, mp, eps)
    o = orc.BoxMesh(3, 2, 4, 5, mp, eps)
    assert np.array_equal(m.gid, o.gid)
    assert np.array_equal(m.EToB, o.EToB)
    for a, b in [(m.x, o.x), (m.y, o.y), (m.z, o.z)]:
        assert np.max(np.abs(a - b)) < 1e-15


@pytest.mark.parametrize("N", range(1, 10))
def test_gll_matches_oracle(orc, N):
    import nekrs_amd
    z, w, D = nekrs_amd.gll(N)
    zo, wo = orc.gll(N)
    assert np.max(np.abs(z - zo)) < 1e-15 and np.max(np.abs(w - wo)) < 1e-15
    assert np.max(np.abs(D - orc.dmatrix(N))) < 1e-13


def test_error_codes_without_device():
    import nekrs_amd
    from nekrs_amd import _lib
    lib = _lib.load()
    x = np.zeros(8)
    rc = lib.nrs_box_mesh(0, 1, 1, 1, 0, 0.0, x.ctypes.data, x.ctypes.data, x.ctypes.data, None, None)
    assert rc == _lib.NRS_ERR_INVALID
    assert b"bad sizes" in lib.nrs_last_error()
    rc = lib.nrs_box_mesh(1, 1, 1, 1, 7, 0.0, x.ctypes.data, x.ctypes.data, x.ctypes.data, None, None)
    assert rc == _lib.NRS_ERR_INVALID
    with pytest.raises(nekrs_amd.NrsError):
        nekrs_amd.box_mesh(1, 1, 0, 3)
    rc = lib.nrs_ogs_setup(4, None, C.byref(C.c_void_p()))
    assert rc == _lib.NRS_ERR_INVALID
    rc = lib.nrs_elliptic_setup(None, None, C.byref(C.c_void_p()))
    assert rc == _lib.NRS_ERR_INVALID


def test_default_opts():
    import nekrs_amd
    o = nekrs_amd.default_opts()
    assert o.solver == 0 and o.preconditioner == 1 and o.tol == 1e-8 and o.relative == 1
    assert o.chebyMinFactor == 0.1 and o.chebyMaxFactor == 1.1
