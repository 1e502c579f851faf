"""The C ABI (CPU): libcmbl.so loads without a GPU, exports every function
declared in include/cmbl.h, the ctypes binding covers them all, and argument
validation fails cleanly before any device work."""
import ctypes
import re

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "cmbl.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cmbl_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["cmbl_plan_create", "cmbl_lenseflow", "cmbl_lenseflow_vjp", "cmbl_gradf_logpdf",
              "cmbl_phi_prepare", "cmbl_rfft2", "cmbl_irfft2", "cmbl_dot_f64"]:
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    from cmblensing_amd import _lib
    lib = _lib.load(require_device=False)
    for n in declared():
        assert hasattr(lib, n), n
    assert set(declared()) <= set(_lib.exported_symbols())


def test_validation_errors_without_device(pkg):
    from cmblensing_amd import _lib
    lib = _lib.load(require_device=False)
    assert lib.cmbl_version().decode().startswith("cmbl")
    h = ctypes.c_void_p()
    assert lib.cmbl_plan_create(ctypes.byref(h), 100, 2.0, 0, 4) != 0      # N not 2^m
    assert "power of two" in lib.cmbl_last_error().decode()
    assert lib.cmbl_plan_create(ctypes.byref(h), 256, 2.0, 7, 4) != 0      # bad dtype
    assert lib.cmbl_plan_create(ctypes.byref(h), 256, -1.0, 0, 4) != 0     # bad pixel size
    assert lib.cmbl_plan_create(None, 256, 2.0, 0, 4) != 0
    assert lib.cmbl_lenseflow(None, None, None, None, 1, 2, 7, 1, 0, None) != 0
    assert lib.cmbl_gradf_final(None, None, None, None, None, 1, 2, None) != 0
    assert "NULL argument" in lib.cmbl_last_error().decode()
    assert lib.cmbl_plan_destroy(None) == 0


def test_product_path_refuses_without_device(pkg):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    from cmblensing_amd import _lib
    with pytest.raises(_lib.CmblError):
        _lib.load(require_device=True)
    f = pkg.FlatS0(scalar_matrix=torch.zeros(16, 16, dtype=torch.float64))
    with pytest.raises(_lib.CmblError):
        pkg.lense_flow(f, pkg.FlatS0(scalar_matrix=torch.zeros(16, 16, dtype=torch.float64)), n=2)
