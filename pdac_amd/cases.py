"""Synthetic, seeded PDAC cases.

Every case returns a :class:`~pdac_amd.deck.Case` whose initial fields are
thermodynamically self-consistent, as the reference's restart path expects
(``taperd`` followed by ``cnvert`` itd=2, setup.f:359-392, recomputes only the
molar fractions, the void fraction and the heat capacities):

    rog = p * M / (R T)   (thermal_eosg, eosg.f:114-137)
    rgp = ep * rog,  ep = 1 - sum(rlk / rho_s)
    sieg = cg(T) * T      (tzero = hzerog = 0)
    sies = cps * Ts
"""
from __future__ import annotations

import numpy as np

from . import constants as C
from .deck import Case


def _mole(ygc, gas_type):
    w = C.GMW[np.asarray(gas_type) - 1].reshape(-1, 1)
    mol = (ygc / w).sum(axis=0)
    inv = np.where(np.abs(mol) <= 1e-10, 0.0, 1.0 / np.where(mol == 0, 1, mol))
    return ygc / w * inv


def fill_state(case: Case, p, tg, ts, eps, ygc, ug=None, vg=None, wg=None,
               us=None, vs=None, ws=None):
    """Fill ``case.fields`` from primitive fields (arrays over all cells).

    ``ts``, ``eps`` are (nsolid, ntot); ``ygc`` is (ngas, ntot)."""
    n = case.ntot
    ns = case.nsolid
    z = np.zeros(n)
    rl = np.asarray(case.density, dtype=np.float64)[:ns]
    inrl = 1.0 / rl
    rlk = np.asarray(eps, dtype=np.float64).reshape(ns, n) * rl.reshape(-1, 1)
    rls = np.zeros(n)
    for s in range(ns):
        rls = rls + rlk[s] * inrl[s]
    ep = 1.0 - rls
    xgc = _mole(ygc, case.gas_type)
    w = C.GMW[np.asarray(case.gas_type) - 1]
    mg = np.zeros(n)
    for g in range(len(case.gas_type)):
        mg = mg + xgc[g] * w[g]
    rog = p / (C.RGAS * tg) * mg
    rgp = ep * rog
    cp = C.hcapg(tg, case.icpc, case.tref)
    cg = np.zeros(n)
    for g, gt in enumerate(case.gas_type):
        cg = cg + cp[gt - 1] * ygc[g]
    sieg = (tg - C.TZERO) * cg + C.HZEROG
    cps = np.asarray(case.specific_heat, dtype=np.float64)[:ns].reshape(-1, 1)
    ts = np.asarray(ts, dtype=np.float64).reshape(ns, n)
    sies = (ts - C.TZERO) * cps + C.HZEROS
    f = case.fields
    f["p"] = np.array(p, dtype=np.float64)
    f["ug"] = z.copy() if ug is None else np.array(ug, dtype=np.float64)
    f["wg"] = z.copy() if wg is None else np.array(wg, dtype=np.float64)
    if case.three_d:
        f["vg"] = z.copy() if vg is None else np.array(vg, dtype=np.float64)
    f["sieg"] = sieg
    f["rlk"] = rlk
    zs = np.zeros((ns, n))
    f["us"] = zs.copy() if us is None else np.array(us, dtype=np.float64).reshape(ns, n)
    f["ws"] = zs.copy() if ws is None else np.array(ws, dtype=np.float64).reshape(ns, n)
    if case.three_d:
        f["vs"] = zs.copy() if vs is None else np.array(vs, dtype=np.float64).reshape(ns, n)
    f["sies"] = sies
    f["ygc"] = np.array(ygc, dtype=np.float64).reshape(len(case.gas_type), n)
    f["rgp"] = rgp
    f["rog"] = rog
    f["tg"] = np.array(tg, dtype=np.float64)
    f["ts"] = ts
    return case


def shock_tube(nx=400, nz=8, dx=0.01, nsteps=1000, dt=1.0e-5, seed=1234,
               nsolid=1, muscl=1, rungekut=3, **kw) -> Case:
    """2D Cartesian multiphase shock tube (nx includes the two wall columns).

    Left half: p = 1e6 Pa, T = 1000 K, one particle class (100 um, 2500 kg/m3)
    at volume fraction 0.01 with seeded +-1 % uniform noise.  Right half:
    p = 1e5 Pa, T = 300 K, particle fraction 1e-8.  Slip walls all around,
    no gravity.  Gas: water vapour + air (gas types 5, 6)."""
    case = Case(job_type="2D", itc=0, nx=nx, ny=1, nz=nz,
                dx=np.full(nx, dx), dz=np.full(nz, dx),
                west=2, east=2, bottom=2, top=2,
                nsolid=nsolid, diameter=[100.0] * nsolid,
                density=[2500.0] * nsolid, sphericity=[1.0] * nsolid,
                viscosity=[0.5] * nsolid, specific_heat=[1200.0] * nsolid,
                thermal_conductivity=[2.0] * nsolid,
                twophase_limit=[1.0e-8] * nsolid,
                gravz=0.0, muscl=muscl, rungekut=rungekut, dt=dt,
                nsteps=nsteps, **kw)
    rng = np.random.default_rng(seed)
    n = case.ntot
    i = np.tile(np.arange(nx), nz)
    left = i < nx // 2
    p = np.where(left, 1.0e6, 1.0e5)
    tg = np.where(left, 1000.0, 300.0)
    noise = rng.uniform(-0.01, 0.01, size=(nsolid, n))
    eps = np.where(left, 0.01 * (1.0 + noise), 1.0e-8) / nsolid
    ts = np.broadcast_to(tg, (nsolid, n)).copy()
    # gas: 20 % water vapour (type 5), 80 % air (type 6) by mass
    ygc = np.stack([np.full(n, 0.2), np.full(n, 0.8)])
    return fill_state(case, p, tg, ts, eps, ygc)
