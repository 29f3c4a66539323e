"""The C-ABI library loads and exports every function include/mdiscr.h
declares; argument checks run on the host before any device work (CPU)."""
import ctypes
import os
import re

import pytest

from mdiscriminator_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mdiscr.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdiscr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libmdiscr.so not built"
    names = declared_functions()
    assert len(names) >= 13
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # The Python binding declares a signature for every one of them.
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_library_is_gfx950():
    """The embedded device code targets gfx950 (our own build's target id)."""
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_null_context_is_einval():
    lib = _lib.load()
    d = _lib.MdiscrDiscr(_lib.DISCR_MOD, 4)
    assert lib.mdiscr_partition_u32(None, None, None, 0, d, None, None) == _lib.EINVAL
    assert lib.mdiscr_ctx_destroy(None) == _lib.EINVAL
    assert lib.mdiscr_sync(None, None) == _lib.EINVAL
    assert lib.mdiscr_timing_enable(None, 1) == _lib.EINVAL
    assert lib.mdiscr_ctx_create(None, 0) == _lib.EINVAL
    assert lib.mdiscr_status_str(_lib.EBUSY) == b"busy"


def test_status_strings():
    lib = _lib.load()
    for s in range(6):
        assert lib.mdiscr_status_str(s)


def test_host_tensors_rejected_before_any_call():
    """The Python layer hands the library device addresses only: a host tensor
    raises before any call, so no host pointer reaches a kernel."""
    import torch
    import mdiscriminator_amd as md
    with pytest.raises(md.MdiscrError) as e:
        md._ptr(torch.zeros(4, dtype=torch.int32))
    assert e.value.status == _lib.EINVAL
    assert md._ptr(None) is None and md._ptr(1234) == 1234
