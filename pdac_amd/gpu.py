"""ctypes binding of the HIP library ``libpdacgpu.so`` (include/pdacgpu.h).

The library is the product: there is no fallback.  Loading fails with an
error when it has not been built (``python -m pdac_amd.build``)."""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import abi
from .solver import Solver

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpdacgpu.so")
_lib = None

# every entry point of include/pdacgpu.h
EXPORTS = [
    "pdac_create", "pdac_destroy", "pdac_set_grid", "pdac_set_params", "pdac_put",
    "pdac_get", "pdac_field_ncomp", "pdac_cnvert", "pdac_boundary", "pdac_fieldn",
    "pdac_tilde", "pdac_iter", "pdac_ygas", "pdac_htilde", "pdac_ftem",
    "pdac_begin_step", "pdac_step", "pdac_sync", "pdac_last_error", "pdac_get_momflux",
    "pdac_profile_enable", "pdac_profile_select", "pdac_profile_reset", "pdac_profile_read",
    "pdac_profile_visits", "pdac_initialize_fieldn", "pdac_explicit_iter",
]


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with `python -m pdac_amd.build`")
        L = ct.CDLL(LIB)
        L.pdac_create.argtypes = [ct.POINTER(ct.c_void_p), ct.POINTER(abi.Dims), ct.c_int]
        L.pdac_create.restype = ct.c_int
        L.pdac_last_error.restype = ct.c_char_p
        L.pdac_get_momflux.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.POINTER(ct.c_double)]
        L.pdac_get_momflux.restype = ct.c_int
        L.pdac_profile_enable.argtypes = [ct.c_void_p, ct.c_int]
        L.pdac_profile_reset.argtypes = [ct.c_void_p]
        L.pdac_profile_select.argtypes = [ct.c_void_p, ct.c_uint32]
        L.pdac_profile_read.argtypes = [ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int64),
                                        ct.POINTER(ct.c_double)]
        L.pdac_profile_visits.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int64)]
        L.pdac_sync.argtypes = [ct.c_void_p]
        _lib = L
    return _lib


class GpuSolver(Solver):
    """:class:`Solver` on the HIP library, on one device."""

    def momflux(self, comp: int, face: int) -> np.ndarray:
        out = np.empty(self.n * (self.case.nsolid + 1))
        rc = self.lib.pdac_get_momflux(self.ctx, comp, face, abi.dptr(out))
        self._check("get_momflux", rc)
        return out.reshape(self.case.nsolid + 1, self.n)


def gpu_solver(case, device: int = 0, iter_order=abi.ORDER_REDBLACK, **kw) -> GpuSolver:
    """GPU solver of ``case``.  iter_order: the red-black sweep of iter; for
    an explicit case (time_integration = 1) the reference's stage order, or
    ``abi.ORDER_STAGE``."""
    L = lib()
    create = lambda ctx, dims: L.pdac_create(ctx, dims, device)  # noqa: E731
    return GpuSolver(L, "pdac_", create, case, iter_order=iter_order, **kw)
