"""The C ABI library loads and exports every symbol include/fcap.h declares (CPU only, no compute)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fcap.h")
LIB = os.path.join(ROOT, "freecappuccino_amd", "libfcap.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_entry_points():
    syms = declared_symbols()
    for s in ("fc_iccg", "fc_dpcg", "fc_bicgstab", "fc_laplacian", "fc_csr_build", "fc_mesh_geometry",
              "fc_spmv", "fc_sys_solve"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libfcap.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_front_end_binds_every_symbol():
    from freecappuccino_amd import fcap
    lib = fcap.library()
    for s in declared_symbols():
        fn = getattr(lib, s)
        assert fn.restype is not None or s == "fc_last_error"


def test_no_device_is_reported_not_raised_by_the_library():
    """Without a GPU, fc_init returns a status instead of aborting (skipped where a GPU is present)."""
    import pytest
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    from freecappuccino_amd import fcap
    h = ctypes.c_void_p()
    rc = fcap.library().fc_init(ctypes.byref(h), 0)
    assert rc == -7
