"""ctypes binding of librfd.so (the C ABI declared in include/rfd.h).

The library is loaded from this directory only.  There is no fallback: if the
HIP library is missing or cannot be loaded, every product entry point raises.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "librfd.so"

_u32 = C.c_uint32
_f32 = C.c_float
_f64 = C.c_double
_int = C.c_int
_vp = C.c_void_p
_d2 = C.POINTER(C.c_double)
_f3 = C.POINTER(C.c_float)

# name -> argument types (return type is int unless listed in _RESTYPE)
SIGNATURES: dict[str, list] = {
    "rfd_ctx_create": [C.POINTER(_vp), _int, _vp],
    "rfd_ctx_destroy": [_vp],
    "rfd_last_error": [_vp],
    "rfd_sync": [_vp],
    "rfd_read_flags": [_vp, C.POINTER(_u32)],
    "rfd_version": [],
    "rfd_stream_copy": [_vp, _vp, _vp, C.c_size_t, _u32],
    "rfd_cell_distribution": [_vp, _vp, _u32, _f32, _u32, _u32, _u32, _vp, _vp, _vp],
    "rfd_cell_distribution_noalloc": [_vp, _vp, _u32, _f32, _u32, _u32, _u32, _vp, _vp, _vp],
    "rfd_il_construct": [_vp, _vp, _u32, _vp, _vp, _vp, _u32, _f32, _f32, _f32, _u32, _u32, _u32],
    "rfd_il_reconstruct_bpm": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32,
                               _u32, _u32, _u32, _int],
    "rfd_bonds_construct": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32,
                            _u32, _f32],
    "rfd_delete_far_links": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32,
                             _f32],
    "rfd_zero_fm": [_vp, _vp, _vp, _u32],
    "rfd_forces_bpm": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32,
                       _u32, _f32, _f32, _f32, _f32, _f32, _f32],
    "rfd_forces_contact": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _f32, _f32,
                           _f32, _f32, _f32, _f32, _f32, _f32],
    "rfd_forces_slots": [_vp, C.c_int] + [_vp] * 13 + [_u32, _u32] + [_f32] * 9,
    "rfd_sumup_forces": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32],
    "rfd_check_break": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32],
    "rfd_increments": [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32],
    "rfd_platen_uniaxial": [_vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32, _f32, _f32, _f32, _f32,
                            _f32, _f32, _d2],
    "rfd_platen_brazil": [_vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32, _f32, _f32, _f32, _f32,
                          _f32, _f32, _d2],
    "rfd_platen_uniaxial_penalty": [_vp, _vp, _vp, _vp, _u32] + [_f32] * 8 + [_d2],
    "rfd_cut_restriction": [_vp, _vp, _vp, _vp, _u32, _f3, _f3, _f3],
    "rfd_fire_increments": [_vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32],
    "rfd_fire_halfstep": [_vp, _vp, _vp, _u32, _f32],
    "rfd_fire_power": [_vp, _vp, _vp, _vp, _vp, _u32, _d2],
    "rfd_fire_ref_init": [_vp, _f32, _f32],
    "rfd_fire_relax_reference": [_vp, _vp, _vp, _vp, _u32, _f32, _vp, _u32, _vp, _vp, _vp],
}
_RESTYPE = {"rfd_last_error": C.c_char_p, "rfd_version": C.c_char_p, "rfd_fire_ref_init": None}

# int (*)(void *user): the force and restriction callbacks of the FIRE loop
CALLBACK = C.CFUNCTYPE(C.c_int, C.c_void_p)


class FireRef(C.Structure):
    """rfd_fire_ref: SetFIREData's settings and the loop state."""

    _fields_ = [("dt0", _f32), ("dtmax", _f32), ("dtmin", _f32), ("dtgrow", _f32),
                ("dtshrink", _f32), ("alpha0", _f32), ("alphashrink", _f32),
                ("n_negative_max", _u32), ("n_delay", _u32), ("dt", _f32), ("alpha", _f32),
                ("n_positive", _u32), ("n_negative", _u32), ("steps", _u32), ("stopped", _int),
                ("power", _f64)]

_lib: C.CDLL | None = None


class RfdError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, name: str, status: int, message: str):
        super().__init__(f"{name} failed with status {status}: {message}")
        self.status = status


def lib() -> C.CDLL:
    """Load librfd.so (once) and attach the signatures; raise if missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"HIP library {LIB_PATH} is not built; run __graft_entry__.build() "
                "(there is no CPU fallback)")
        handle = C.CDLL(str(LIB_PATH))
        for name, argtypes in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = handle
    return _lib


def declared_symbols() -> list[str]:
    """Function names declared in include/rfd.h (parsed from the header)."""
    import re

    hdr = pathlib.Path(__file__).resolve().parents[2] / "include" / "rfd.h"
    text = hdr.read_text()
    return sorted(set(re.findall(r"\b(rfd_[a-z0-9_]+)\s*\(", text)))
