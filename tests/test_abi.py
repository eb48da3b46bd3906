"""The C-ABI library loads and exports every symbol include/rfd.h declares
(no compute calls: this runs without a GPU)."""
import ctypes

import rfd


def test_library_exports_declared_symbols():
    lib = rfd.lib()
    missing = [s for s in rfd.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_string():
    v = rfd.lib().rfd_version()
    assert v and b"gfx950" in v


def test_null_context_is_rejected():
    lib = rfd.lib()
    assert lib.rfd_sync(None) < 0
    assert lib.rfd_zero_fm(None, None, None, 0) < 0
    f = ctypes.c_uint32()
    assert lib.rfd_read_flags(None, ctypes.byref(f)) < 0


def test_context_without_device_reports_nodev():
    import torch

    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    assert rfd.lib().rfd_ctx_create(ctypes.byref(h), 0, None) == -2
