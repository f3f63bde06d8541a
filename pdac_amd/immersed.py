"""Immersed boundaries of a 3D case with topography: the reference's set-up.

With ``immb = 1`` the reference forces the velocity components that lie
under the ground next to the flow (forcing points), so that the no-slip
condition holds on the interpolated ground surface rather than on cell
faces.  Its set-up, restated here in the reference's arithmetic:

* set_forcing (immb.f:73-516): the ground on the three staggered grids
  (interpolate_dem, topo.f:970-1029), cells with all but one or two faces
  under the ground become ``filled_cell_1/2`` and get external forcing
  points (ext_forcing3d, immb.f:1071-1261), the forcing points under the
  ground (forcing3d, immb.f:913-1069), and ``immb_cell`` flags;
* local_forcing (ghost.f:2888-3050): the cell of each forcing point and the
  neighbours its interpolation reads (index_q, index_qq);
* fill_cells (ghost.f:3052-3261): per computed cell, which faces are above
  the ground (bd) and the open volume fraction vf; cells with none become
  ``filled_cell_1``;
* face_fractions (ghost.f:3263-3321): the face masks b_e..b_b and 1/vf.
  Its loop returns at the first computed cell that has both an x and a z
  forcing point, after zeroing that cell's e, w, t, b masks: every later
  cell keeps the masks 1 and 1/vf = 1 it was given by face_init.

Only 3D is restated (the 3D summit-vent case).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import constants as C
from .topography import TopoSetup, interp_2d

# neighbour offsets of the forcing interpolation directions 1..8
# (forcing3d, immb.f:940-955): W, SW, S, SE, E, NE, N, NW
DI = [0, -1, -1, 0, 1, 1, 1, 0, -1]
DJ = [0, 0, -1, -1, -1, 0, 1, 1, 1]


@dataclass
class ForcingPoints:
    """Forcing points of one velocity component, in the reference's order;
    i, j, k 1-based."""
    i: List[int] = field(default_factory=list)
    j: List[int] = field(default_factory=list)
    k: List[int] = field(default_factory=list)
    int_: List[int] = field(default_factory=list)
    di: List[int] = field(default_factory=list)
    dj: List[int] = field(default_factory=list)
    dk: List[int] = field(default_factory=list)
    nsx: List[float] = field(default_factory=list)
    nsy: List[float] = field(default_factory=list)
    nsz: List[float] = field(default_factory=list)

    def add(self, i, j, k, it, di, dj, dk, x, y, z):
        self.i.append(i); self.j.append(j); self.k.append(k); self.int_.append(it)
        self.di.append(di); self.dj.append(dj); self.dk.append(dk)
        self.nsx.append(x); self.nsy.append(y); self.nsz.append(z)

    def __len__(self):
        return len(self.i)


@dataclass
class ImmersedSetup:
    flags: np.ndarray            # final cell flags (ntot,)
    fp: List[ForcingPoints]      # x, y, z
    cx: List[np.ndarray]         # coordinates of each component's grid: (cx, cy, cz)
    num: np.ndarray              # (3, ntot) forcing point of each cell, 1-based, 0 none
    index_q: List[np.ndarray]    # 0-based cell of the first interpolation neighbour
    index_qq: List[np.ndarray]   # 0-based second neighbour, -1 where the reference sets none
    bd: np.ndarray               # (ntot,) open faces (bits e w t b n s = 1 2 4 8 16 32)
    vf: np.ndarray               # (ntot,) open volume fraction
    bmask: np.ndarray            # (6, ntot) b_e, b_w, b_n, b_s, b_t, b_b
    ivf: np.ndarray              # (ntot,)
    ib_last: int                 # 0-based cell where face_fractions returned (-1: never)


class _DemState:
    """interpolate_dem with the reference's module state: ord2d is kept
    from call to call and only set where a level lies under the ground."""

    def __init__(self, ts: TopoSetup, nx, ny, nz):
        self.ts = ts
        self.ord2d = ts.ord2d.copy()
        self.nx, self.ny, self.nz = nx, ny, nz

    def call(self, cx, cy, cz, ff):
        ts = self.ts
        nx, ny, nz = self.nx, self.ny, self.nz
        topo2d = interp_2d(ts.xtop, ts.ytop, ts.ztop2d, cx, cy)
        for k in range(nz):
            self.ord2d = np.where(cz[k] <= topo2d, k + 1, self.ord2d)
        o = self.ord2d
        if np.any(o[1:-1, 1:-1] < 1):
            raise ValueError("interpolate_dem: a column without a level under the ground")
        ffv = ff.reshape(nz, ny, nx)
        inner = o[1:-1, 1:-1]
        # lowest ground among the 8 horizontal neighbours of interior columns
        nbmin = np.minimum.reduce([o[:-2, 1:-1], o[:-2, :-2], o[1:-1, :-2], o[2:, :-2],
                                   o[2:, 1:-1], o[2:, 2:], o[1:-1, 2:], o[:-2, 2:]])
        for k in range(2, int(inner.max())):
            hit = (k <= inner - 1) & (nbmin < k)
            ffv[k - 1, 1:-1, 1:-1] |= hit.T
        ii, jj = np.meshgrid(np.arange(1, nx - 1), np.arange(1, ny - 1), indexing="ij")
        ffv[inner - 1, jj, ii] = True
        return topo2d


def _gamma(topo2d, cz, i, j, k, n):
    num = cz[k - 1] - topo2d[i - 1, j - 1]
    den = topo2d[i + DI[n] - 1, j + DJ[n] - 1] - topo2d[i - 1, j - 1]
    return num / den


def _forcing3d(o, cx, cy, cz, topo2d, fp: ForcingPoints, nx, ny):
    """forcing3d (immb.f:913-1069): loop i, j, k; the interpolation
    neighbour is the one whose gamma is nearest 0.5 (mingam, first on
    ties)."""
    cx, cy, cz = (np.asarray(a, dtype=np.float64) for a in (cx, cy, cz))
    recs = []
    kmax = int(o[1:-1, 1:-1].max())
    for k in range(2, kmax):
        ii, jj = np.nonzero(o[1:-1, 1:-1] - 1 >= k)
        ii, jj = ii + 2, jj + 2                      # 1-based
        if ii.size == 0:
            continue
        g = np.full((9, ii.size), -1.0)
        for n in range(1, 9):
            on = o[ii + DI[n] - 1, jj + DJ[n] - 1]
            sel = on < k
            num = cz[k - 1] - topo2d[ii - 1, jj - 1]
            den = topo2d[ii + DI[n] - 1, jj + DJ[n] - 1] - topo2d[ii - 1, jj - 1]
            with np.errstate(divide="ignore", invalid="ignore"):
                g[n] = np.where(sel, num / den, -1.0)
        has = np.max(g[1:], axis=0) > -1
        d = np.abs(0.5 - g[1:])
        gint = np.argmin(d, axis=0) + 1
        if np.any(has & (np.min(d, axis=0) >= 1.0)):
            raise ValueError("forcing3d: no interpolation direction (mingam)")
        for p in np.flatnonzero(has):
            recs.append((int(ii[p]), int(jj[p]), k, int(gint[p]), float(g[gint[p], p])))
    for i in range(2, nx):
        for j in range(2, ny):
            recs.append((i, j, int(o[i - 1, j - 1]), 0, 0.0))
    recs.sort(key=lambda r: (r[0], r[1], r[2]))
    for i, j, k, gi, gg in recs:
        if gi == 0:
            fp.add(i, j, k, 0, 0, 0, 1, cx[i - 1], cy[j - 1], topo2d[i - 1, j - 1])
        else:
            fp.add(i, j, k, gi, DI[gi], DJ[gi], 0,
                   gg * cx[i - 1] + (1.0 - gg) * cx[i + DI[gi] - 1],
                   gg * cy[j - 1] + (1.0 - gg) * cy[j + DJ[gi] - 1],
                   cz[k - 1])


def _ext_forcing3d(o, i, j, k, cx, cy, cz, topo2d, fp: ForcingPoints, axis):
    """ext_forcing3d (immb.f:1071-1261) for a cell whose component lies
    above the ground."""
    g = [None] + [-1.0] * 8
    g_ext = [None, 0.0, 0.0, 0.0, 0.0]
    maxg = 1.0
    g_min = None
    gint = fpint = None
    if k > o[i - 1, j - 1]:
        for n in range(1, 9):
            if o[i + DI[n] - 1, j + DJ[n] - 1] >= k:
                g[n] = _gamma(topo2d, cz, i, j, k, n)
        if axis != 1:
            g_ext[1] = g[1] * g[5]
        g_ext[2] = g[2] * g[6]
        if axis != 2:
            g_ext[3] = g[3] * g[7]
        g_ext[4] = g[4] * g[8]
        for l in range(1, 5):
            if g_ext[l] < 0 and -g_ext[l] < maxg:
                g_min = l
                maxg = -g_ext[l]
        if min(g_ext[1:]) < 0 and g_min is not None:
            a, b = g_min, g_min + 4
            if g[a] < 0:
                gint, fpint = b, 20 + a
            else:
                gint, fpint = a, 24 + a
    if k == o[i - 1, j - 1] + 1:
        maxg_z = (cz[k - 1] - topo2d[i - 1, j - 1]) / (cz[k] - topo2d[i - 1, j - 1])
    else:
        maxg_z = 1.0
    if maxg < 1 and maxg < maxg_z:
        gg = g[gint]
        fp.add(i, j, k, fpint, DI[fpint - 20], DJ[fpint - 20], 0,
               (1.0 - gg) * cx[i - 1] + gg * cx[i + DI[gint] - 1],
               (1.0 - gg) * cy[j - 1] + gg * cy[j + DJ[gint] - 1],
               cz[k - 1])
    else:
        fp.add(i, j, k, 20, 0, 0, 1, cx[i - 1], cy[j - 1], topo2d[i - 1, j - 1])


def immersed_setup(case, ts: TopoSetup, flags_in: np.ndarray) -> ImmersedSetup:
    """set_forcing, local_forcing, fill_cells and face_fractions for a 3D
    case; ``flags_in`` are the flags after export_topography."""
    if not case.three_d:
        raise ValueError("immersed boundaries are restated for 3D cases only")
    nx, ny, nz = case.nx, case.ny, case.nz
    n = nx * ny * nz
    m = ts.mesh
    fl = flags_in.copy()
    flv = fl.reshape(nz, ny, nx)
    st = _DemState(ts, nx, ny, nz)
    force = [np.zeros(n, dtype=bool) for _ in range(3)]
    ext = [np.zeros(n, dtype=bool) for _ in range(3)]
    grids = [(m.xb, m.y, m.z), (m.x, m.yb, m.z), (m.x, m.y, m.zb)]
    topo = [st.call(*grids[a], force[a]) for a in range(3)]
    tx, ty, tc = topo
    # filled cells (immb.f:242-358): all but one or two faces of a cell
    # under the ground
    keep = (C.INLET_CELL, C.VENT_CELL)
    zc = m.z[1:-1, None, None]
    zt = m.zb[1:-1, None, None]
    zbm = m.zb[:-2, None, None]
    tci = tc[1:-1, 1:-1].T[None]
    bds = ((zc > tx[1:-1, 1:-1].T[None]) * 1 + (zc > tx[:-2, 1:-1].T[None]) * 2 +
           (zt > tci) * 4 + (zbm >= tci) * 8 +
           (zc > ty[1:-1, 1:-1].T[None]) * 16 + (zc > ty[1:-1, :-2].T[None]) * 32)
    if np.any(bds == 8):
        raise ValueError("set_forcing: inconsistent immersed boundary (bds = 8)")
    full = np.zeros((nz, ny, nx), dtype=np.int64)
    full[1:-1, 1:-1, 1:-1] = bds
    full[0, :, :] = -1
    full[-1, :, :] = -1
    full[:, 0, :] = -1
    full[:, -1, :] = -1
    full[:, :, 0] = -1
    full[:, :, -1] = -1
    b = full.reshape(-1)
    own_x = np.isin(b, (1, 17, 33, 3))
    west_x = np.isin(b, (2, 18, 34, 3))
    own_y = np.isin(b, (16, 18, 17, 48))
    south_y = np.isin(b, (32, 34, 33, 48))
    ext[0] |= own_x
    ext[0][np.flatnonzero(west_x) - 1] = True
    ext[1] |= own_y
    ext[1][np.flatnonzero(south_y) - nx] = True
    ext[2] |= b == 4
    f1 = np.isin(b, (1, 2, 4, 16, 32)) & ~np.isin(fl, keep)
    f2 = np.isin(b, (17, 18, 34, 33, 48, 3)) & ~np.isin(fl, keep)
    fl[f1] = C.FILLED_CELL_1
    fl[f2] = C.FILLED_CELL_2
    counts = [int(force[a].sum() + ext[a].sum()) for a in range(3)]
    fps = []
    for a in range(3):
        topo2d = st.call(*grids[a], force[a])
        cx, cy, cz = grids[a]
        fp = ForcingPoints()
        for c in np.flatnonzero(ext[a]):
            k = c // (nx * ny) + 1
            j = (c // nx) % ny + 1
            i = c % nx + 1
            _ext_forcing3d(st.ord2d, i, j, k, cx, cy, cz, topo2d, fp, a + 1)
        _forcing3d(st.ord2d, cx, cy, cz, topo2d, fp, nx, ny)
        if len(fp) != counts[a]:
            raise ValueError(f"set_forcing: {len(fp)} forcing points along axis {a + 1}, "
                             f"{counts[a]} counted (the reference would fail)")
        for p in range(len(fp)):
            i, j, k = fp.i[p], fp.j[p], fp.k[p]
            c = (i - 1) + (j - 1) * nx + (k - 1) * nx * ny
            if i > 1 and j > 1 and k > 1 and fl[c] not in (C.INLET_CELL, C.VENT_CELL,
                                                             C.FILLED_CELL_1, C.FILLED_CELL_2):
                fl[c] = C.IMMB_CELL
        fps.append(fp)
    # write_profile (topo.f:1225-1255): the cells under a specified-flow
    # block of type 5 are no-slip walls again, whatever set_forcing made
    # of them
    for b in getattr(case, "blocks", []):
        if b["typ"] != 5:
            continue
        (xl, xh), (yl, yh), (zl, _) = b["x"], b["y"], b["z"]
        flv[:zl - 1, yl - 1:yh, xl - 1:xh] = C.NOSLIP_WALL
    # local_forcing (ghost.f:2888-3050)
    num = np.zeros((3, n), dtype=np.int64)
    iq, iqq = [], []
    for a in range(3):
        fp = fps[a]
        q = np.empty(len(fp), dtype=np.int64)
        qq = np.full(len(fp), -1, dtype=np.int64)
        for p in range(len(fp)):
            i, j, k = fp.i[p], fp.j[p], fp.k[p]
            c = (i - 1) + (j - 1) * nx + (k - 1) * nx * ny
            num[a, c] = p + 1
            q[p] = c + fp.di[p] + fp.dj[p] * nx + fp.dk[p] * nx * ny
            it = fp.int_[p]
            if it % 2 != 0 or it == 0:
                qq[p] = c + 2 * fp.di[p] + 2 * fp.dj[p] * nx + 2 * fp.dk[p] * nx * ny
        iq.append(q)
        iqq.append(qq)
    # fill_cells (ghost.f:3052-3261), in the reference's cell order: a cell
    # left with no open face becomes filled_cell_1 and later cells see it
    # so; within one level that cannot change a neighbour's count (the
    # shared face is closed for both), across levels it can, so the levels
    # go in order
    bd = np.zeros(n, dtype=np.int64)
    vf = np.zeros(n)
    flv = fl.reshape(nz, ny, nx)
    bdv = bd.reshape(nz, ny, nx)
    vfv = vf.reshape(nz, ny, nx)
    for k in range(2, nz):
        lay = flv[k - 1]
        comp = (lay & 1) == 1
        comp[0, :] = comp[-1, :] = comp[:, 0] = comp[:, -1] = False
        zk, zbk, zbm1 = m.z[k - 1], m.zb[k - 1], m.zb[k - 2]
        bb = np.zeros((ny, nx), dtype=np.int64)
        vv = np.zeros((ny, nx))

        def side(open_, nbf, bit):
            nonlocal bb, vv
            f1_ = open_ & (nbf == C.FILLED_CELL_1)
            f2_ = open_ & (nbf == C.FILLED_CELL_2)
            other = open_ & ~f1_ & ~f2_
            vv = vv + np.where(f1_, 2.0, 0.0) + np.where(f2_, 1.0, 0.0) + np.where(other, 1.0, 0.0)
            bb = bb + np.where(other, bit, 0)

        sl = np.s_[1:-1, 1:-1]
        op = np.zeros((ny, nx), dtype=bool)
        nb = np.zeros((ny, nx), dtype=np.int64)
        # east, west
        op[sl] = zk > tx[1:-1, 1:-1].T
        nb[sl] = lay[1:-1, 2:]
        side(op.copy(), nb.copy(), 1)
        op[sl] = zk > tx[:-2, 1:-1].T
        nb[sl] = lay[1:-1, :-2]
        side(op.copy(), nb.copy(), 2)
        # top
        op[sl] = zbk > tc[1:-1, 1:-1].T
        vv = vv + np.where(op, 1.0, 0.0)
        bb = bb + np.where(op, 4, 0)
        # bottom
        op[sl] = zbm1 >= tc[1:-1, 1:-1].T
        nb[sl] = flv[k - 2][1:-1, 1:-1]
        side(op.copy(), nb.copy(), 8)
        # north, south
        op[sl] = zk > ty[1:-1, 1:-1].T
        nb[sl] = lay[2:, 1:-1]
        side(op.copy(), nb.copy(), 16)
        op[sl] = zk > ty[1:-1, :-2].T
        nb[sl] = lay[:-2, 1:-1]
        side(op.copy(), nb.copy(), 32)
        vv = vv / 6
        bdv[k - 1] = np.where(comp, bb, 0)
        vfv[k - 1] = np.where(comp, vv, 0.0)
        lay[comp & (vv == 0.0)] = C.FILLED_CELL_1
    # face_fractions (ghost.f:3263-3321)
    bmask = np.ones((6, n), dtype=np.int64)   # e, w, n, s, t, b
    ivf = np.ones(n)
    comp = (fl & 1) == 1
    stop = np.flatnonzero(comp & (num[0] != 0) & (num[2] != 0))
    ib_last = int(stop[0]) if stop.size else n - 1
    upto = np.zeros(n, dtype=bool)
    upto[:ib_last if stop.size else n] = True
    sel = comp & upto
    for row, bit in ((0, 1), (1, 2), (4, 4), (5, 8), (2, 16), (3, 32)):
        bmask[row, sel] = (bd[sel] & bit != 0).astype(np.int64)
    with np.errstate(divide="ignore"):
        ivf[sel] = np.where(vf[sel] > 0.0, 1.0 / np.where(vf[sel] > 0.0, vf[sel], 1.0), 0.0)
    if stop.size:
        bmask[[0, 1, 4, 5], ib_last] = 0
    return ImmersedSetup(fl, fps, [list(g) for g in grids], num, iq, iqq, bd, vf, bmask, ivf, ib_last)
