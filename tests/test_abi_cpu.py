"""The C-ABI library without a GPU: it loads, exports exactly what
include/*.h declares, agrees on struct sizes (the reference's abi_check,
crates/warpjq-core/src/gpu/ffi.rs:222-242) and fails cleanly with no device."""
import ctypes
import os
import re

import pytest

import warpjq_amd as w

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("warpjq_kernels.h", "warpjq_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(warpjq_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(w.LIB_PATH)
    declared = _declared()
    assert set(w.REFERENCE_SYMBOLS) <= declared
    for name in declared:
        getattr(lib, name)
    assert declared == set(w.REFERENCE_SYMBOLS) | set(w.ADDITIVE_SYMBOLS)


def test_abi_sizes_match_the_reference():
    assert w.abi_sizes() == [20, 8, 32, 8, 48, 56]
    assert w.abi_check() == 0
    assert w.abi_check([20, 8, 32, 8, 48, 57]) == 4
    assert ctypes.sizeof(w.ChunkResult) == 8 * 6 + 8 * 4 + 8 * 3 + 8 + 48 + 8 + 8


def test_compile_errors_are_reported():
    with pytest.raises(ValueError):
        w.Compiled("select(.a = 1)")
    c = w.Compiled("select(.status == 500) | count")
    assert c.view.contents.n_needed == 1


def test_probe_without_a_device_is_a_clean_error():
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(w.DeviceError):
        w.probe()
