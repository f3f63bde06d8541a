"""Atmospheric profile p_atm(k), t_atm(k) at cell centres.

Restates set_atmosphere / compute_profile / hydrostatic
(atmosphere.f:81-265).  Only free in/outflow boundaries read it.
"""
from __future__ import annotations

import numpy as np

from . import constants as C

# standard atmosphere layers (input.f:433-446 defaults)
_LAYERS = [(1.1e4, -6.5e-3), (2.0e4, 0.0), (3.2e4, 1.0e-3), (4.7e4, 2.8e-3),
           (5.1e4, 0.0), (7.1e4, -2.8e-3), (8.0e4, -2.0e-3)]


def _hydrostatic(zbot, ztop, gradt, tbot, pbot, gravz, rm):
    """Temperature linear in z, hydrostatic pressure (atmosphere.f:217-265)."""
    ttop = tbot + gradt * (ztop - zbot)
    if gradt == 0.0:
        ptop = pbot * np.exp(gravz / (rm * tbot) * (ztop - zbot))
    else:
        ptop = pbot * (ttop / tbot) ** (gravz / (rm * gradt))
    return ttop, ptop


def atmosphere_profile(case):
    """Return (p_atm, t_atm) over the nz cell rows of ``case``."""
    nz = case.nz
    dz = np.asarray(case.dz, dtype=np.float64)
    z = np.empty(nz)
    zb = 0.0
    for k in range(nz):
        if k == 0:
            zb = 0.0
            z[0] = -0.5 * dz[0] + zb
        else:
            zb = zb + dz[k]
            z[k] = zb - 0.5 * dz[k]
    if not case.stratification or case.gravz == 0.0:
        return np.full(nz, case.p_ground), np.full(nz, case.t_ground)
    rm = C.RGAS / C.GMW[5]  # air gas constant
    layers = []
    ptop, ttop = case.p_ground, case.t_ground
    zbot = 0.0
    for ztop, grad in _LAYERS:
        tb, pb = ttop, ptop
        ttop, ptop = _hydrostatic(zbot, ztop, grad, tb, pb, case.gravz, rm)
        layers.append((zbot, ztop, grad, tb, pb))
        zbot = ztop
    p = np.empty(nz)
    t = np.empty(nz)
    for k in range(nz):
        za = z[k]
        for zb_, zt_, g_, tb_, pb_ in layers:
            if za <= zt_ or zt_ == layers[-1][1]:
                t[k], p[k] = _hydrostatic(zb_, max(za, zb_) if za >= 0 else za, g_, tb_, pb_, case.gravz, rm)
                break
    return p, t
