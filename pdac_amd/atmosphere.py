"""Atmospheric profile p_atm(k), t_atm(k) at cell centres.

Restates set_atmosphere, compute_profile and hydrostatic
(atmosphere.f:81-265) with the default layer table of input.f:433-446.
Only free in/outflow boundaries read the profile.
"""
from __future__ import annotations

import math

import numpy as np

from . import constants as C

# (top height [m], temperature gradient [K/m]) of the standard layers
LAYERS = [(1.1e4, -6.5e-3), (2.0e4, 0.0), (3.2e4, 1.0e-3), (4.7e4, 2.8e-3),
          (5.1e4, 0.0), (7.1e4, -2.8e-3), (8.0e4, -2.0e-3)]


def hydrostatic(zb, z, grad, tb, pb, gravz):
    """Linear temperature, hydrostatic pressure (atmosphere.f:239-263)."""
    erreair = C.RGAS / C.GMW[5]
    if grad != 0.0:
        cost = gravz / (erreair * grad)
        t = tb + grad * (z - zb)
        p = pb * (t / tb) ** cost
    else:
        cost = gravz / (erreair * tb)
        t = tb
        p = pb * math.exp(cost * (z - zb))
    return t, p


def cell_heights(dz, zzero=0.0):
    """z(k) of grid_setup (grid.f:298-303)."""
    nz = len(dz)
    z = np.empty(nz)
    zb = zzero
    z[0] = -0.5 * dz[0] + zb
    for k in range(1, nz):
        zb = zb + dz[k]
        z[k] = zb - 0.5 * dz[k]
    return z


def atmosphere_profile(case):
    """Return (p_atm, t_atm) over the nz cell rows of ``case``."""
    gravz = case.gravz
    strat = case.stratification and gravz != 0.0  # control_atmosphere
    ptop, ttop = case.p_ground, case.t_ground
    tops = []
    for l, (zt, grad) in enumerate(LAYERS):
        zbot = 0.0 if l == 0 else LAYERS[l - 1][0]
        pbot, tbot = ptop, ttop
        ttop, ptop = hydrostatic(zbot, zt, grad, tbot, pbot, gravz)
        tops.append((zt, ptop, ttop, grad))
    if getattr(case, "itp", 0):
        # the mesh lifted onto the topography (set_profile, topo.f:265-278)
        z = case._topo()[0].mesh.z
    else:
        z = cell_heights(np.asarray(case.dz, dtype=np.float64))
    nz = len(z)
    p = np.empty(nz)
    t = np.empty(nz)
    if z[0] <= tops[0][0]:
        l = 0
    else:
        l = 1
    for k in range(nz):
        za = z[k]
        if za > 0.0:
            if not strat:
                zbot, pbot, tbot, grad = 0.0, case.p_ground, case.t_ground, 0.0
            else:
                if za > tops[l][0]:
                    l += 1
                if l == 0:
                    zbot, pbot, tbot = 0.0, case.p_ground, case.t_ground
                else:
                    zbot, pbot, tbot = tops[l - 1][0], tops[l - 1][1], tops[l - 1][2]
                grad = tops[l][3]
            t[k], p[k] = hydrostatic(zbot, za, grad, tbot, pbot, gravz)
        else:
            t[k], p[k] = case.t_ground, case.p_ground
    return p, t
