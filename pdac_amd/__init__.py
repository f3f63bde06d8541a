"""MI355X (gfx950) implementation of PDAC's per-time-step hot path.

The compute path is the HIP library ``libpdacgpu.so`` built from
``pdac_amd/csrc`` (C-ABI in ``include/pdacgpu.h``); :mod:`pdac_amd.gpu` loads
it and :class:`pdac_amd.solver.Solver` drives it in the reference's
subroutine order.  The other modules read and write the reference's input
deck and restart formats and build seeded synthetic cases.
"""
__all__ = ["abi", "cases", "deck", "restart", "solver", "gpu", "build"]
