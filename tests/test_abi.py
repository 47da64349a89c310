"""The C ABI library loads and exports every entry point include/qcbflow.h
declares; the Python binding knows all of them.  No compute calls (no GPU)."""
import ctypes
import os
import re

import qcbflow
from conftest import ROOT


def declared():
    with open(os.path.join(ROOT, "include", "qcbflow.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(qcbf_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("qcbf_create", "qcbf_rk3_step", "qcbf_flow_measure", "qcbf_load_ildg",
                 "qcbf_save_ildg", "qcbf_measure", "qcbf_su3_exp"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(qcbflow.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(qcbflow.SIGNATURES) == declared()
    qcbflow.load()  # sets restype/argtypes for all of them


def test_library_is_gfx950_code_object():
    with open(qcbflow.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"gfx950" in blob


def test_create_refuses_volumes_of_2_28_sites():
    """qcbf_create and qcbf_batch_create refuse 2^28 sites or more before any
    HIP call (no GPU needed): the FP32-mixed kernels address a field's word
    arrays by 32-bit byte offsets of 16 bytes per site."""
    lib = qcbflow.load()
    h = ctypes.c_void_p()
    for prec in (qcbflow.FP64, qcbflow.FP32_MIXED):
        dims = (ctypes.c_int * 4)(128, 128, 128, 128)
        assert lib.qcbf_create(dims, prec, ctypes.byref(h)) == qcbflow.QCBF_ERR_ARG
        assert not h.value
        both = (ctypes.c_int * 8)(4, 4, 4, 4, 128, 128, 128, 128)
        assert lib.qcbf_batch_create(2, both, prec, ctypes.byref(h)) == qcbflow.QCBF_ERR_ARG
        assert not h.value
