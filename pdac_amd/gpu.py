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
    "pdac_ens_create", "pdac_ens_destroy", "pdac_ens_tune", "pdac_ens_set_tiles", "pdac_ens_step",
    "pdac_explicit_step_noalloc", "pdac_set_ib", "pdac_profile_checks", "pdac_sgsg",
    "pdac_snapshot_size", "pdac_snapshot_async", "pdac_snapshot_fetch", "pdac_ib_forcing",
    "pdac_ib_fill_cells",
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
        L.pdac_profile_checks.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int64)]
        L.pdac_sync.argtypes = [ct.c_void_p]
        L.pdac_ens_create.argtypes = [ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_void_p), ct.c_int]
        L.pdac_ens_destroy.argtypes = [ct.c_void_p]
        L.pdac_ens_tune.argtypes = [ct.c_void_p, ct.c_int32, ct.POINTER(ct.c_int32),
                                    ct.POINTER(ct.c_double)]
        L.pdac_ens_set_tiles.argtypes = [ct.c_void_p, ct.POINTER(ct.c_int32)]
        L.pdac_ens_step.argtypes = [ct.c_void_p, ct.POINTER(ct.c_double), ct.c_int32, ct.c_int32,
                                    ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double)]
        L.pdac_explicit_step_noalloc.argtypes = [ct.c_void_p, ct.c_double, ct.c_int32, ct.c_int32,
                                                 ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double)]
        L.pdac_explicit_step_noalloc.restype = ct.c_int
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


TILE_SHAPES = [(256, 1), (128, 2), (64, 4), (32, 8)]
TILE_KERNELS = ["boundary", "xinit", "xflux", "xrhs", "xmass", "xsolve", "ygas_flux", "ygas_update"]


class Ensemble:
    """Explicit-step domains advanced together by batched launches
    (``pdac_ens_*``).  ``members`` are :class:`GpuSolver` objects of cases
    with ``time_integration = 1`` that share dimension, particle classes,
    gas species and precision."""

    def __init__(self, members):
        self.lib = lib()
        self.members = list(members)
        arr = (ct.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.ens = ct.c_void_p()
        rc = self.lib.pdac_ens_create(ct.byref(self.ens), arr, len(self.members))
        if rc != 0:
            raise RuntimeError(f"pdac_ens_create returned {rc}: {self.lib.pdac_last_error().decode()}")

    def _check(self, where, rc):
        if rc != 0:
            raise RuntimeError(f"{where} returned {rc}: {self.lib.pdac_last_error().decode()}")

    def tune(self, reps=5):
        """Choose a tile shape per kernel class; returns ({kernel: (tx, tz)},
        {kernel: [mean us per tile shape]})."""
        tiles = (ct.c_int32 * len(TILE_KERNELS))()
        us = (ct.c_double * (len(TILE_KERNELS) * len(TILE_SHAPES)))()
        self._check("pdac_ens_tune", self.lib.pdac_ens_tune(self.ens, reps, tiles, us))
        chosen = {k: TILE_SHAPES[tiles[i]] for i, k in enumerate(TILE_KERNELS)}
        times = {k: [us[i * len(TILE_SHAPES) + t] for t in range(len(TILE_SHAPES))]
                 for i, k in enumerate(TILE_KERNELS)}
        return chosen, times

    def set_tiles(self, idx):
        arr = (ct.c_int32 * len(TILE_KERNELS))(*idx)
        self._check("pdac_ens_set_tiles", self.lib.pdac_ens_set_tiles(self.ens, arr))

    def step(self, nsteps=1, dt=None, rungekut=1):
        """Advance every member; returns (eos failures, Courant numbers) per member."""
        n = len(self.members)
        dts = (ct.c_double * n)(*[m.case.dt if dt is None else dt[i] for i, m in enumerate(self.members)])
        fail = (ct.c_int32 * n)()
        cfl = (ct.c_double * n)()
        self._check("pdac_ens_step", self.lib.pdac_ens_step(self.ens, dts, rungekut, nsteps, fail, cfl))
        for i, m in enumerate(self.members):
            for _ in range(nsteps):
                m.time = m.time + dts[i]
        return list(fail), list(cfl)

    def close(self):
        if self.ens:
            self.lib.pdac_ens_destroy(self.ens)
            self.ens = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
