"""Device forms of the tests' custom 2D hooks (tests/native/devhooks.hip,
built into tests/native/libhfv_devhooks.so): (address, user pointer) pairs
for Simulation.set_boundary(device=...), with the same parameters as the
host forms in _cases.py."""
from __future__ import annotations

import ctypes as C
import os

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libhfv_devhooks.so")
_lib = None
_keep = []  # user data read by the hooks at every call


def available() -> bool:
    return os.path.exists(SO)


def _addr(name: str) -> int:
    global _lib
    if _lib is None:
        _lib = C.CDLL(SO)
    return C.cast(getattr(_lib, name), C.c_void_p).value


def _user(obj) -> int:
    _keep.append(obj)
    return C.cast(obj, C.c_void_p).value


def lid(lid_u=1.0, lid_v=0.0):
    return _addr("hfv_test_lid_hook"), _user((C.c_double * 2)(lid_u, lid_v))


def inflow(state=(1.2, 0.8, 0.1, 1.1)):
    return _addr("hfv_test_inflow_hook"), _user((C.c_double * 4)(*state))


def copy(side: int):
    return _addr("hfv_test_copy_hook"), _user((C.c_int * 1)(side))
