"""Physical constants and closed-form property laws used on the host.

Values follow the reference's ``setc`` (setup.f:551-612), the particle
constants of ``particles_constants_set`` (particles.f:35-78) and the molar heat
capacity polynomials of ``hcapg`` (hcapgs.f:33-85).  Single-precision literals
of the reference are reproduced as the float32 value widened to float64.
"""
from __future__ import annotations

import numpy as np

MAX_NGAS = 7
# molecular weights in kg/mol: O2, N2, CO2, H2, H2O, Air, SO2 (setup.f:565-574)
GMW = np.array([32.0, 28.0, 44.0, 2.0, 18.0, 28.96442, 64.0]) * 1.0e-3
# Lennard-Jones sigma (Angstrom) and epsilon/k (K) (setup.f:580-593)
MMUGS = np.array([3.467, 3.798, 3.941, 2.827, 2.641, 3.711, 4.112])
MMUGEK = np.array([106.7, 71.4, 195.2, 59.7, 809.1, 78.6, 335.4])
RGAS = 8.31432
GAMMAAIR = 1.4
TZERO = 0.0
HZEROG = 0.0
HZEROS = 0.0
# maximum packing, a single-precision literal in particles.f:46
PHI_PACK = float(np.float32(0.63))

# cell flags (grid.f:45-58)
FLUID, SLIP_WALL, NOSLIP_WALL, BL_CELL = 1, 2, 4, 5
FREE_IO, INLET_CELL, ZERO_GRAD, VENT_CELL = 8, 16, 32, 64
DOME_CELL, IMMB_CELL, FILLED_CELL_1, FILLED_CELL_2 = 129, 257, 512, 1024

# molar heat capacity coefficients c0 + c1 T + c2 T^2 + c3 T^3, J/(mol K)
_CP_COEF = [
    (2.811e1, -3.68e-6, 1.746e-5, -1.065e-8),
    (3.115e1, -1.357e-2, 2.680e-5, -1.168e-8),
    (1.980e1, 7.344e-2, -5.602e-5, 1.715e-8),
    (2.714e1, 9.274e-3, -1.381e-5, 7.645e-9),
    (3.224e1, 1.924e-3, 1.055e-5, -3.596e-9),
    (3.0413e1, -1.059e-2, 2.458e-5, -1.135e-8),
    (2.385e1, 6.699e-2, -4.961e-5, 1.328e-8),
]


def hcapg(tg, icpc=2, tref=298.0, species=None):
    """Specific heats (J/(kg K)) of the 7 gas species at temperature ``tg``.

    Restates hcapgs.f:33-85 with the same left-to-right evaluation order.
    ``tg`` may be a numpy array; the result has a leading axis of 7, or of
    len(species) for the 0-based species listed."""
    t1 = np.asarray(tg, dtype=np.float64) if icpc == 2 else np.full_like(
        np.asarray(tg, dtype=np.float64), tref)
    t2 = t1 * t1
    t3 = t1 * t2
    rows = list(range(7)) if species is None else [int(s) for s in species]
    cp = np.empty((len(rows),) + t1.shape)
    for r, s in enumerate(rows):
        c0, c1, c2, c3 = _CP_COEF[s]
        cp[r] = (((c0 + c1 * t1) + c2 * t2) + c3 * t3) / GMW[s]
        if icpc == 0 and s in (4, 5):
            cp[r] = 1810.0 if s == 4 else 1004.0
    return cp
