"""The C-ABI libraries load without a GPU and export every function their
headers declare. No compute calls."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from mocc_amd.native import LIB_DIR

HEADERS = {
    "mocgpu.h": ("libmocgpu.so", "mocgpu_"),
    "mocc_host.h": ("libmocc_host.so", "mocc_host_"),
}


def declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports(header):
    lib, prefix = HEADERS[header]
    names = declared(header, prefix)
    assert len(names) >= 10
    L = C.CDLL(os.path.join(LIB_DIR, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_h5_exports():
    """libmocc_h5 (loaded after the HDF5 C library it links) exports every
    function mocc_h5.h declares."""
    from mocc_amd import output
    try:
        L = output.h5lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    names = declared("mocc_h5.h", "mocc_h5_")
    assert len(names) == 6
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bad_arguments_are_status_codes():
    from mocc_amd.native import MocgpuOptions, gpu
    o = MocgpuOptions()
    gpu().mocgpu_default_options(C.byref(o))
    assert o.n_inner >= 1
    h = C.c_void_p()
    assert gpu().mocgpu_create(None, C.byref(o), C.byref(h)) == -1
    assert b"null" in gpu().mocgpu_last_error(None)


def test_no_device_is_a_status_code(pins3x3):
    """Without a usable device, creating a handle fails with a status code
    and a message (no crash and no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from mocc_amd.native import MocgpuOptions, gpu
    o = pins3x3.options()
    h = C.c_void_p()
    rc = gpu().mocgpu_create(pins3x3.view, C.byref(o), C.byref(h))
    assert rc == -2
    assert gpu().mocgpu_last_error(None)
