"""ctypes binding of the HIP C ABI (include/swe_mi355x.h).

The shared library is built in-tree (samoa_amd/lib/libswe_mi355x.so) by
samoa_amd/build.py.  There is no fallback: if the library is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libswe_mi355x.so")

SWE_OK = 0
SWE_ERR_ARG, SWE_ERR_HIP, SWE_ERR_NOMEM, SWE_ERR_STATE = -1, -2, -3, -4
STATUS = {-1: "SWE_ERR_ARG", -2: "SWE_ERR_HIP", -3: "SWE_ERR_NOMEM", -4: "SWE_ERR_STATE"}


class SweError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn}: {STATUS.get(code, code)}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64),
        ("depth", C.c_void_p),
        ("plotter", C.c_void_p),
        ("nbr", C.c_void_p),
        ("scaling", C.c_double),
        ("g", C.c_double),
        ("dry_tolerance", C.c_double),
        ("courant", C.c_double),
        ("min_depth", C.c_int32),
        ("max_depth", C.c_int32),
        ("solver", C.c_int32),
        ("boundary", C.c_int32),
    ]


class DeviceMesh(C.Structure):
    """swe_device_mesh: a mesh in device memory (swe_step_device_noalloc)."""

    _fields_ = [
        ("n_cells", C.c_int64),
        ("depth", C.c_void_p),
        ("plotter", C.c_void_p),
        ("nbr", C.c_void_p),
        ("scaling", C.c_double),
        ("g", C.c_double),
        ("dry_tolerance", C.c_double),
        ("min_depth", C.c_int32),
        ("max_depth", C.c_int32),
        ("solver", C.c_int32),
        ("boundary", C.c_int32),
    ]


# entry point -> (restype, argtypes); every one declared in include/swe_mi355x.h
_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_I64 = C.c_int64
SIGNATURES = {
    "swe_create": (_I, [C.POINTER(MeshDesc), C.POINTER(_P)]),
    "swe_upload": (_I, [_P, _P, _P, _P, _P]),
    "swe_download": (_I, [_P, _P, _P, _P, _P]),
    "swe_step": (_I, [_P, _D, _I, C.POINTER(_D), C.POINTER(_I64), _P]),
    "swe_run": (_I, [_P, _I64, _D, _I, _D, C.POINTER(_D), C.POINTER(_D)]),
    "swe_step_device": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _I, _P, _P, _P, _P]),
    "swe_step_device_noalloc": (_I, [C.POINTER(DeviceMesh), _P, _P, _P, _P, _P, _P, _P, _D, _I,
                                     _P, _P, _P, _P]),
    "swe_set_stream": (_I, [_P, _P]),
    "swe_set_arithmetic": (_I, [_P, _I]),
    "swe_batch_set_arithmetic": (_I, [_P, _I]),
    "swe_adapt": (_I, [_P, _P, C.POINTER(_I64), C.POINTER(_I)]),
    "swe_keep_refinement_flags": (_I, [_P, _I]),
    "swe_keep_eta_jump_flags": (_I, [_P, _D, _D]),
    "swe_flag_eta_jumps": (_I, [_P, _D, _D]),
    "swe_set_bathymetry_raster": (_I, [_P, _I64, _I64, _D, _D, _D, _D, _D, _D, _P]),
    "swe_cells": (_I64, [_P]),
    "swe_adapt_record_device": (_I, [_P, _P, _P, C.POINTER(_I64), C.POINTER(_I), _P]),
    "swe_transfer_adjoint_device": (_I, [_I64, _P, _P, _I] + [_P] * 8 + [_P]),
    "swe_download_mesh": (_I, [_P, _P, _P, _P]),
    "swe_destroy": (_I, [_P]),
    "swe_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "swe_chunk_cells": (_I, []),
    "swe_batch_create": (_I, [C.c_int32, _P, C.POINTER(_P)]),
    "swe_batch_upload": (_I, [_P, C.c_int32, _P, _P, _P, _P]),
    "swe_batch_download": (_I, [_P, C.c_int32, _P, _P, _P, _P]),
    "swe_batch_step": (_I, [_P, _P, _I, _P, _P]),
    "swe_batch_run": (_I, [_P, _I64, _P, _I, _D, _P, _P]),
    "swe_set_tuning": (_I, [_P, _I, _I64]),
    "swe_get_tuning": (_I, [_P, C.POINTER(_I), C.POINTER(_I64)]),
    "swe_batch_set_tuning": (_I, [_P, _I, _I64]),
    "swe_batch_get_tuning": (_I, [_P, C.POINTER(_I), C.POINTER(_I64)]),
    "swe_batch_set_stream": (_I, [_P, _P]),
    "swe_batch_cells": (_I64, [_P]),
    "swe_batch_destroy": (_I, [_P]),
    "swe_batch_offset": (_I64, [_P, C.c_int32]),
    "swe_forward_device": (_I, [_P] + [_P] * 8 + [_I, _P, _P, _P]),
    "swe_adjoint_step_device": (_I, [_P] + [_P] * 13 + [_I, _P]),
    "swe_batch_forward_device": (_I, [_P] + [_P] * 8 + [_I, _P, _P, _P]),
    "swe_batch_adjoint_device": (_I, [_P] + [_P] * 13 + [_I, _P]),
}

_lib = None


def _share_runtime_with_torch() -> None:
    """One HIP runtime per process.

    PyTorch-ROCm ships its own libamdhip64 with the same SONAME
    (libamdhip64.so.7) as /opt/rocm's.  If this library is loaded first, the
    process maps /opt/rocm's runtime and a later `import torch` brings in a
    second one that sees no devices.  Importing torch first makes the dynamic
    loader satisfy this library's dependency with torch's runtime, so device
    pointers are shared between the two."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path: str = HIP_LIB):
    """Loads the HIP library once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python samoa_amd/build.py`")
    _share_runtime_with_torch()
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    lib = load()
    buf = C.create_string_buffer(1024)
    lib.swe_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(fn: str, code: int) -> None:
    if code != SWE_OK:
        raise SweError(fn, code, last_error())
