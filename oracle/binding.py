"""ctypes binding of liboracle.so (TEST INFRASTRUCTURE ONLY).

Builds the library on first use when the C compiler is available and
returns :class:`pdac_amd.solver.Solver` objects that drive it."""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

from pdac_amd import abi
from pdac_amd.solver import Solver

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ct.CDLL(LIB)
        _lib.pdac_cpu_create.argtypes = [ct.POINTER(ct.c_void_p), ct.POINTER(abi.Dims)]
        _lib.pdac_cpu_create.restype = ct.c_int
    return _lib


def cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC, **kw):
    L = lib()
    return Solver(L, "pdac_cpu_", L.pdac_cpu_create, case, iter_order=iter_order, **kw)
