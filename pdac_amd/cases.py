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
    cp = C.hcapg(tg, case.icpc, case.tref, species=[gt - 1 for gt in case.gas_type])
    cg = np.zeros(n)
    for g in range(len(case.gas_type)):
        cg = cg + cp[g] * ygc[g]
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
               nsolid=1, muscl=1, rungekut=3, bottom=3, top=2, **kw) -> Case:
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


def _stretched(n, d0, d1, n0=1):
    """Cell sizes growing geometrically from d0 to d1 (n cells, including the
    boundary cell at index 0 which gets d0)."""
    if n <= n0 + 1 or d1 == d0:
        return np.full(n, float(d0))
    r = (d1 / d0) ** (1.0 / (n - n0 - 1))
    out = np.empty(n)
    out[:n0] = d0
    for i in range(n0, n):
        out[i] = d0 * r ** (i - n0 + 1)
    return np.minimum(out, d1)


def eruption_column(nx=512, nz=1024, nsolid=2, seed=1234, dr0=1.0, dr1=20.0,
                    vent_radius=50.0, w_vent=150.0, t_vent=1100.0, eps_vent=0.01,
                    nsteps=5000, dt=None, top=2, east=4, vent="pulse", **kw) -> Case:
    """2D cylindrical collapsing eruption column.

    Radial cells stretch from ``dr0`` at the axis to ``dr1``; vertical cells
    likewise.  Stratified standard atmosphere (gravity -9.81), air + water
    vapour, particle classes 50 and 500 um (2500 kg/m3).  The vent reaches
    out to ``vent_radius``: gas w = ``w_vent`` with seeded 2 % noise per
    cell, T = ``t_vent``, particle fraction ``eps_vent`` split evenly, water
    vapour only.  ``vent="inlet"``: a sustained eruption, the vent cells are
    a FIXED_FLOWS block of inlet cells (type 5, flag 16) in the ground row
    whose values stay fixed; ``vent="pulse"``: the same values set once in
    the lowest fluid row (an initial impulse).  The axis is a slip wall, the
    ground no-slip, the outer boundary free in/outflow; the top is a slip
    lid by default."""
    dx = _stretched(nx, dr0, dr1)
    dz = _stretched(nz, dr0, dr1)
    if dt is None:
        dt = 0.2 * dr0 / (w_vent + 350.0)
    case = Case(job_type="2D", itc=1, nx=nx, ny=1, nz=nz, dx=dx, dz=dz,
                west=2, east=east, bottom=3, top=top,
                nsolid=nsolid, diameter=[50.0, 500.0][:nsolid] + [100.0] * max(0, nsolid - 2),
                density=[2500.0] * nsolid, sphericity=[1.0] * nsolid,
                viscosity=[0.5] * nsolid, specific_heat=[1200.0] * nsolid,
                thermal_conductivity=[2.0] * nsolid, twophase_limit=[1.0e-8] * nsolid,
                gravz=-9.81, stratification=True, dt=dt, nsteps=nsteps, **kw)
    from .atmosphere import atmosphere_profile
    p_atm, t_atm = atmosphere_profile(case)
    rng = np.random.default_rng(seed)
    n = case.ntot
    kk = np.repeat(np.arange(nz), nx)
    ii = np.tile(np.arange(nx), nz)
    xb = np.cumsum(dx) - dx[0]
    r = xb - 0.5 * dx
    p = p_atm[kk].copy()
    tg = t_atm[kk].copy()
    eps = np.full((nsolid, n), 0.0)
    ygc = np.stack([np.zeros(n), np.ones(n)])
    wg = np.zeros(n)
    ws = np.zeros((nsolid, n))
    row = 0 if vent == "inlet" else 1
    vent_cells = (kk == row) & (ii >= 1) & (ii <= nx - 2) & (r[ii] <= vent_radius)
    if vent == "inlet":
        iv = int(np.max(ii[vent_cells])) + 1
        case.blocks = [{"typ": 5, "x": [2, iv], "z": [1, 1]}]
    vent = vent_cells
    noise = rng.uniform(-1.0, 1.0, size=n)
    wg[vent] = w_vent * (1.0 + 0.02 * noise[vent])
    for s_ in range(nsolid):
        ws[s_][vent] = wg[vent]
        eps[s_][vent] = eps_vent / nsolid
    tg[vent] = t_vent
    ygc[0][vent] = 1.0
    ygc[1][vent] = 0.0
    ts = np.broadcast_to(tg, (nsolid, n)).copy()
    return fill_state(case, p, tg, ts, eps, ygc, wg=wg, ws=ws)


def summit_vent_3d(nx=256, ny=256, nz=192, d=10.0, nsolid=3, seed=1234, vent_half=50.0,
                   w_vent=100.0, t_vent=1000.0, eps_vent=0.02, nsteps=1000, dt=None,
                   **kw) -> Case:
    """3D Cartesian eruption from a square vent on flat ground.

    Uniform ``d`` cells; particle classes 30, 250 and 1000 um with densities
    2700, 2000 and 1000 kg/m3; air + water vapour; stratified standard
    atmosphere.  The vent is the square of half-width ``vent_half`` at the
    centre of the lowest fluid layer: gas w = ``w_vent`` with seeded 2 %
    noise, T = ``t_vent``, particle fraction ``eps_vent`` split evenly, water
    vapour only.  The ground is a no-slip wall, the four sides free
    in/outflow, the top a slip lid.  (This is the shape of the 3D cone case
    with the topography left out: immersed boundaries are not part of this
    library.)"""
    if dt is None:
        dt = 0.2 * d / (w_vent + 350.0)
    diam = [30.0, 250.0, 1000.0][:nsolid] + [100.0] * max(0, nsolid - 3)
    dens = [2700.0, 2000.0, 1000.0][:nsolid] + [2500.0] * max(0, nsolid - 3)
    case = Case(job_type="3D", itc=0, nx=nx, ny=ny, nz=nz, dx=np.full(nx, d), dy=np.full(ny, d),
                dz=np.full(nz, d), west=4, east=4, south=4, north=4, bottom=3, top=2,
                nsolid=nsolid, diameter=diam, density=dens, sphericity=[1.0] * nsolid,
                viscosity=[0.5] * nsolid, specific_heat=[1200.0] * nsolid,
                thermal_conductivity=[2.0] * nsolid, twophase_limit=[1.0e-8] * nsolid,
                gravz=-9.81, stratification=True, dt=dt, nsteps=nsteps, **kw)
    from .atmosphere import atmosphere_profile
    p_atm, t_atm = atmosphere_profile(case)
    rng = np.random.default_rng(seed)
    n = case.ntot
    kk = np.repeat(np.arange(nz), nx * ny)
    jj = np.tile(np.repeat(np.arange(ny), nx), nz)
    ii = np.tile(np.arange(nx), ny * nz)
    xc = (np.arange(nx) - 0.5 * (nx - 1)) * d
    yc = (np.arange(ny) - 0.5 * (ny - 1)) * d
    p = p_atm[kk].copy()
    tg = t_atm[kk].copy()
    eps = np.zeros((nsolid, n))
    ygc = np.stack([np.zeros(n), np.ones(n)])
    wg = np.zeros(n)
    ws = np.zeros((nsolid, n))
    vent = (kk == 1) & (np.abs(xc[ii]) <= vent_half) & (np.abs(yc[jj]) <= vent_half)
    noise = rng.uniform(-1.0, 1.0, size=n)
    wg[vent] = w_vent * (1.0 + 0.02 * noise[vent])
    for s_ in range(nsolid):
        ws[s_][vent] = wg[vent]
        eps[s_][vent] = eps_vent / nsolid
    tg[vent] = t_vent
    ygc[0][vent] = 1.0
    ygc[1][vent] = 0.0
    ts = np.broadcast_to(tg, (nsolid, n)).copy()
    return fill_state(case, p, tg, ts, eps, ygc, wg=wg, ws=ws)


def mixed_box(three_d=False, nx=10, ny=7, nz=9, nsolid=2, seed=7, itc=0,
              gravz=-9.81, nsteps=2, dt=2.0e-4, clean_top=0, cell=1.0, vmax=20.0,
              prange=(0.9e5, 1.3e5), **kw) -> Case:
    """Small randomised domain for parity tests: non-uniform cells, several
    particle classes, seeded smooth random fields, walls of both kinds.
    ``clean_top`` rows below the lid start without particles.  Cell sizes
    are ``cell`` times U[0.8, 1.25], velocities U[-vmax, vmax] and
    pressures U[prange] before smoothing (cells of 0.1 mm make viscous
    stresses and conduction as large as the fluxes)."""
    rng = np.random.default_rng(seed)
    dx = rng.uniform(0.8, 1.25, nx) * cell
    dz = rng.uniform(0.8, 1.25, nz) * cell
    dy = rng.uniform(0.8, 1.25, ny) * cell if three_d else None
    diam = list(rng.choice([30.0, 100.0, 250.0, 1000.0], size=nsolid, replace=False))
    dens = list(rng.uniform(1000.0, 2700.0, nsolid))
    case = Case(job_type="3D" if three_d else "2D", itc=0 if three_d else itc,
                nx=nx, ny=ny if three_d else 1, nz=nz, dx=dx, dy=dy, dz=dz,
                nsolid=nsolid, diameter=diam, density=dens,
                sphericity=[1.0] * nsolid, viscosity=[0.5] * nsolid,
                specific_heat=list(rng.uniform(900.0, 1500.0, nsolid)),
                thermal_conductivity=[2.0] * nsolid, twophase_limit=[1.0e-8] * nsolid,
                gravz=gravz, dt=dt, nsteps=nsteps, **kw)
    n = case.ntot
    shp = (nz, ny if three_d else 1, nx)
    def smooth(a, b):
        f = rng.uniform(a, b, size=shp)
        for ax in range(3):
            f = 0.5 * f + 0.25 * (np.roll(f, 1, ax) + np.roll(f, -1, ax))
        return f.reshape(-1)
    p = smooth(*prange)
    tg = smooth(300.0, 900.0)
    eps = np.stack([smooth(0.0, 0.02) for _ in range(nsolid)])
    if clean_top:
        kk = np.repeat(np.arange(nz), (ny if three_d else 1) * nx)
        eps[:, kk >= nz - 1 - clean_top] = 0.0
    ts = np.stack([tg + smooth(-50.0, 50.0) for _ in range(nsolid)])
    y0 = smooth(0.0, 1.0)
    ygc = np.stack([y0, 1.0 - y0])
    vel = lambda: smooth(-vmax, vmax)
    kwv = dict(ug=vel(), wg=vel(), us=np.stack([vel() for _ in range(nsolid)]),
               ws=np.stack([vel() for _ in range(nsolid)]))
    if three_d:
        kwv["vg"] = vel()
        kwv["vs"] = np.stack([vel() for _ in range(nsolid)])
    return fill_state(case, p, tg, ts, eps, ygc, **kwv)


def summit_cone_3d(nx=256, ny=256, nz=192, d=10.0, nsolid=3, seed=1234, height=1500.0,
                   sigma=1500.0, rough=30.0, vent_half=50.0, plateau=None, w_vent=100.0,
                   t_vent=1000.0, eps_vent=0.02, nsteps=1000, dt=None, lpr=0, **kw) -> Case:
    """3D Cartesian eruption from a square summit vent on a volcanic cone,
    with immersed boundaries.

    Uniform ``d`` cells; topography from a Gaussian cone of ``height`` and
    width ``sigma`` plus seeded fractal roughness of amplitude ``rough``,
    cut flat at the summit (a platform of half-width ``plateau``, default
    twice the vent's) for the vent; the reference's topography set-up and
    immersed boundaries (itp = 1, immb = 1).  The vent is a FIXED_FLOWS
    block of inlet cells (type 5) in the platform's top solid layer, of
    half-width ``vent_half``: gas w = ``w_vent`` with seeded 2 % noise per
    cell, T = ``t_vent``, particle fraction ``eps_vent`` split evenly,
    water vapour only.  Particle classes 30, 250 and 1000 um with densities
    2700, 2000 and 1000 kg/m3; air + water vapour; stratified standard
    atmosphere above the lifted mesh.  Sides free in/outflow, ground no-slip,
    top a slip lid."""
    if dt is None:
        dt = 0.2 * d / (w_vent + 350.0)
    plateau = 2.0 * vent_half if plateau is None else plateau
    diam = [30.0, 250.0, 1000.0][:nsolid] + [100.0] * max(0, nsolid - 3)
    dens = [2700.0, 2000.0, 1000.0][:nsolid] + [2500.0] * max(0, nsolid - 3)
    case = Case(job_type="3D", itc=0, nx=nx, ny=ny, nz=nz, dx=np.full(nx, d), dy=np.full(ny, d),
                dz=np.full(nz, d), west=4, east=4, south=4, north=4, bottom=3, top=2,
                nsolid=nsolid, diameter=diam, density=dens, sphericity=[1.0] * nsolid,
                viscosity=[0.5] * nsolid, specific_heat=[1200.0] * nsolid,
                thermal_conductivity=[2.0] * nsolid, twophase_limit=[1.0e-8] * nsolid,
                gravz=-9.81, stratification=True, dt=dt, nsteps=nsteps, itp=1, immb=1,
                dem={"cell": d, "height": height, "sigma": sigma, "rough": rough,
                     "plateau": plateau, "seed": seed},
                dem_cellsize=d, lpr=lpr, **kw)
    # the platform's surface in the upper quarter of a cell layer: the vent
    # cells are that layer's cells (centre under the ground, top face at or
    # above it), so the inflow crosses an open face into fluid cells
    ts, _ = case._topo()
    m = ts.mesh
    vi = np.flatnonzero(np.abs(m.x) <= vent_half) + 1
    vj = np.flatnonzero(np.abs(m.y) <= vent_half) + 1
    if vi.size == 0 or vj.size == 0:
        raise ValueError("no cell centre within the vent")
    top = float(ts.topo2d[vi[0] - 1, vj[0] - 1])
    kv = int(np.searchsorted(m.zb, top, side="left")) + 1  # first k with zb(k) >= top
    # (the lifted mesh's heights are the DEM's: the lowest ground point is
    # the bottom face)
    case.dem = dict(case.dem, plateau_z=round(float(m.zb[kv - 1]) - 0.25 * d, 2))
    case.__dict__.pop("_topo_cache", None)
    ts, _ = case._topo()
    m = ts.mesh
    top = float(ts.topo2d[vi[0] - 1, vj[0] - 1])
    kv = int(np.searchsorted(m.zb, top, side="left")) + 1
    flat = ts.topo2d[vi - 1][:, vj - 1]
    if flat.min() != flat.max() or not (m.z[kv - 1] < top <= m.zb[kv - 1]):
        raise ValueError("the vent does not sit on a flat summit platform")
    case.blocks = [{"typ": 5, "x": [int(vi[0]), int(vi[-1])], "y": [int(vj[0]), int(vj[-1])],
                    "z": [kv, kv]}]
    case.__dict__.pop("_topo_cache", None)
    case.__dict__.pop("_ib_cache", None)
    from .atmosphere import atmosphere_profile
    p_atm, t_atm = atmosphere_profile(case)
    rng = np.random.default_rng(seed)
    n = case.ntot
    kk = np.repeat(np.arange(nz), nx * ny)
    jj = np.tile(np.repeat(np.arange(ny), nx), nz)
    ii = np.tile(np.arange(nx), ny * nz)
    p = p_atm[kk].copy()
    tg = t_atm[kk].copy()
    eps = np.zeros((nsolid, n))
    ygc = np.stack([np.zeros(n), np.ones(n)])
    wg = np.zeros(n)
    ws = np.zeros((nsolid, n))
    vent = ((kk == kv - 1) & (ii >= vi[0] - 1) & (ii <= vi[-1] - 1) &
            (jj >= vj[0] - 1) & (jj <= vj[-1] - 1))
    noise = rng.uniform(-1.0, 1.0, size=n)
    wg[vent] = w_vent * (1.0 + 0.02 * noise[vent])
    for s_ in range(nsolid):
        ws[s_][vent] = wg[vent]
        eps[s_][vent] = eps_vent / nsolid
    tg[vent] = t_vent
    ygc[0][vent] = 1.0
    ygc[1][vent] = 0.0
    ts_ = np.broadcast_to(tg, (nsolid, n)).copy()
    return fill_state(case, p, tg, ts_, eps, ygc, wg=wg, ws=ws)
