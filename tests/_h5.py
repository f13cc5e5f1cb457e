"""Minimal reader for the snapshot files (HDF5 C API through ctypes; the
image has the HDF5 1.10 library but no h5py) and the h5diff tool."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import numpy as np

CANDIDATES = [os.environ.get("HYDROFV_HDF5_LIBRARY", ""), "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103"]
_lib = None


def lib():
    global _lib
    if _lib is None:
        for name in CANDIDATES:
            if not name:
                continue
            try:
                L = C.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise OSError("HDF5 library not found")
        L.H5open.restype = C.c_int
        L.H5Fopen.restype = C.c_int64
        L.H5Fopen.argtypes = [C.c_char_p, C.c_uint, C.c_int64]
        L.H5Dopen2.restype = C.c_int64
        L.H5Dopen2.argtypes = [C.c_int64, C.c_char_p, C.c_int64]
        L.H5Dget_space.restype = C.c_int64
        L.H5Dget_space.argtypes = [C.c_int64]
        L.H5Sget_simple_extent_npoints.restype = C.c_int64
        L.H5Sget_simple_extent_npoints.argtypes = [C.c_int64]
        L.H5Dread.restype = C.c_int
        L.H5Dread.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        for n in ("H5Dclose", "H5Sclose", "H5Fclose"):
            getattr(L, n).restype = C.c_int
            getattr(L, n).argtypes = [C.c_int64]
        assert L.H5open() >= 0
        _lib = L
    return _lib


def available() -> bool:
    try:
        lib()
        return True
    except OSError:
        return False


def read(path: str, name: str, kind: str = "f8") -> np.ndarray:
    """Dataset `name` (e.g. "/fields/density") as float64 ("f8"), int64
    ("i8") or int32 ("i4")."""
    L = lib()
    types = {"f8": ("H5T_NATIVE_DOUBLE_g", np.float64), "i8": ("H5T_NATIVE_LLONG_g", np.int64),
             "i4": ("H5T_NATIVE_INT_g", np.int32)}
    gname, dtype = types[kind]
    mem_type = C.c_int64.in_dll(L, gname).value
    f = L.H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
    assert f >= 0, path
    try:
        d = L.H5Dopen2(f, name.encode(), 0)
        assert d >= 0, name
        sp = L.H5Dget_space(d)
        n = L.H5Sget_simple_extent_npoints(sp)
        out = np.empty(n, dtype=dtype)
        assert L.H5Dread(d, mem_type, 0, 0, 0, out.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Sclose(sp)
        L.H5Dclose(d)
    finally:
        L.H5Fclose(f)
    return out


def h5diff() -> str | None:
    for p in (shutil.which("h5diff"), "/opt/conda/bin/h5diff"):
        if p and os.path.exists(p):
            return p
    return None


def same_files(a: str, b: str, delta: float | None = None) -> tuple[bool, str]:
    """h5diff of two files: every object present in both, of the same type
    and shape, with equal values (within the absolute `delta` when given;
    a huge delta compares the structure only)."""
    cmd = [h5diff()] + (["-d", repr(delta)] if delta else []) + [a, b]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode == 0, (r.stdout + r.stderr)[-2000:]
