"""ctypes binding of libdgfs_hip.so (the C-ABI declared in include/dgfs_hip.h).

Status codes become exceptions the way the reference's ctypes wrappers do it
(frfs/solvers/dgfs/cufft.py:108-117, frfs/backends/cuda/cublas.py:75-80).
There is no fallback: if the shared library is missing, importing a GPU entry
point raises ``DGFSLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIB_PATH


class DGFSError(RuntimeError):
    pass


class DGFSInvalidValue(DGFSError, ValueError):
    pass


class DGFSUnsupported(DGFSError):
    pass


class DGFSAllocError(DGFSError, MemoryError):
    pass


class DGFSHipError(DGFSError):
    pass


class DGFSNotInitialized(DGFSError):
    pass


class DGFSLibraryMissing(DGFSError, ImportError):
    pass


_STATUS = {
    1: DGFSInvalidValue,
    2: DGFSUnsupported,
    3: DGFSAllocError,
    4: DGFSHipError,
    5: DGFSNotInitialized,
}


class dgfs_vhs_desc(C.Structure):
    _fields_ = [
        ("Nv", C.c_int), ("Nrho", C.c_int), ("M", C.c_int), ("precision", C.c_int),
        ("project", C.c_int), ("L", C.c_double), ("gamma", C.c_double),
        ("prefac", C.c_double), ("sw", C.c_double), ("cw", C.c_double),
        ("qz", C.POINTER(C.c_double)), ("qw", C.POINTER(C.c_double)),
        ("sz", C.POINTER(C.c_double)),
    ]


class dgfs_vss_desc(C.Structure):
    _fields_ = [
        ("Nv", C.c_int), ("Nrho", C.c_int), ("M", C.c_int), ("Mpre", C.c_int),
        ("nspcs", C.c_int), ("precision", C.c_int), ("project", C.c_int),
        ("L", C.c_double), ("sw", C.c_double), ("cw", C.c_double),
        ("qz", C.POINTER(C.c_double)), ("qw", C.POINTER(C.c_double)),
        ("sz", C.POINTER(C.c_double)), ("szpre", C.POINTER(C.c_double)),
        ("masses", C.POINTER(C.c_double)), ("gamma", C.POINTER(C.c_double)),
        ("eta", C.POINTER(C.c_double)), ("prefac", C.POINTER(C.c_double)),
    ]


class dgfs_sur_desc(C.Structure):
    _fields_ = [
        ("Nv", C.c_int), ("max_batch", C.c_int), ("L", C.c_double), ("cw", C.c_double),
        ("gamma", C.c_double), ("prefac", C.c_double),
    ]


_lib = None

_P, _I, _D = C.c_void_p, C.c_int, C.c_double

# name -> argtypes (every entry point returns int status)
_SIGNATURES = {
    "dgfs_vhs_create": [C.POINTER(dgfs_vhs_desc), C.POINTER(C.c_void_p)],
    "dgfs_vhs_eval": [_P, _P, _P, _I, _I, _I, _I, _I, _I, _P],
    "dgfs_vhs_eval_contig": [_P, _P, _P, _I, _P],
    "dgfs_vhs_eval_contig_noalloc": [_P, _P, _P, _I, _P],
    "dgfs_vhs_eval_pair": [_P, _P, _P, _I, _P, _P, _P, _I, _P],
    "dgfs_vhs_destroy": [_P],
    "dgfs_vhs_get_config": [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)],
    "dgfs_vhs_get_dims": [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)],
    "dgfs_vhs_set_groups": [_P, _I],
    "dgfs_vhs_set_slots": [_P, _I],
    "dgfs_vhs_set_timing": [_P, _I],
    "dgfs_vhs_set_algorithm": [_P, _I],
    "dgfs_vhs_get_algorithm": [_P, C.POINTER(_I)],
    "dgfs_vhs_get_timing": [_P, C.POINTER(_D), C.POINTER(_I)],
}

# extra entry points registered by other modules (FR kernels)
EXTRA_SIGNATURES: dict = {}


def register(signatures: dict):
    """Add entry points; applied at once if the library is already loaded."""
    EXTRA_SIGNATURES.update(signatures)
    if _lib is not None:
        _apply(_lib, signatures)


def _apply(lib, signatures):
    for name, args in signatures.items():
        fn = getattr(lib, name)
        fn.restype = C.c_int
        fn.argtypes = args


def lib_path() -> str:
    return os.environ.get("DGFS_HIP_LIB", LIB_PATH)


def load():
    """Load the HIP library (once).  Raises DGFSLibraryMissing if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise DGFSLibraryMissing(
            f"{path} not found: build it with dgfs_hip.build.build() "
            "(there is no CPU fallback for the HIP path)")
    lib = C.CDLL(path)
    lib.dgfs_last_error.restype = C.c_char_p
    lib.dgfs_last_error.argtypes = []
    lib.dgfs_version.restype = C.c_char_p
    lib.dgfs_version.argtypes = []
    _apply(lib, {**_SIGNATURES, **EXTRA_SIGNATURES})
    _lib = lib
    return lib


def version() -> str:
    """The library's build string (dgfs_version)."""
    return load().dgfs_version().decode(errors="replace")


def exported_symbols():
    return sorted(["dgfs_last_error", "dgfs_version", *_SIGNATURES, *EXTRA_SIGNATURES])


def check(status: int, what: str = ""):
    if status == 0:
        return
    lib = load()
    msg = lib.dgfs_last_error().decode(errors="replace")
    exc = _STATUS.get(status, DGFSError)
    raise exc(f"{what}: {msg}" if what else msg)


def call(name: str, *args):
    lib = load()
    check(getattr(lib, name)(*args), name)
