"""Oracle (test infrastructure only): the forcing points of one velocity
component and fill_cells restated in numpy and Python loops, the checkers of
the library's pdac_ib_forcing and pdac_ib_fill_cells.  Only tests import
this module.

forcing3d (immb.f:913-1069): for each interior column and each level below
its ground level, the interpolation direction whose gamma is nearest 0.5
among the neighbours whose ground is below the level (first on ties), then
a point at the ground level interpolating along z (type 0); points sorted by
(i, j, k).  ext_forcing3d (immb.f:1071-1261): for a cell above the ground,
the line through the cell (W-E, SW-NE, S-N, SE-NW) along which the ground is
nearest, if nearer than along z (types 21..28), else along z (type 20).

Pinned by the reference's own forcing-point dumps of two cone fixtures
(tests/test_immersed_setup.py compares both this restatement and the
library with them) and by a randomised comparison with the library on a
rough cone.
"""
from __future__ import annotations

from typing import List

import numpy as np

# neighbour offsets of the interpolation directions 1..8 (immb.f:940-955):
# W, SW, S, SE, E, NE, N, NW
DI = [0, -1, -1, 0, 1, 1, 1, 0, -1]
DJ = [0, 0, -1, -1, -1, 0, 1, 1, 1]


class Points:
    """Forcing points as lists, in the order they are added."""

    def __init__(self):
        self.i: List[int] = []; self.j: List[int] = []; self.k: List[int] = []
        self.int_: List[int] = []; self.di: List[int] = []; self.dj: List[int] = []
        self.dk: List[int] = []; self.nsx: List[float] = []; self.nsy: List[float] = []
        self.nsz: List[float] = []

    def add(self, i, j, k, it, di, dj, dk, x, y, z):
        self.i.append(i); self.j.append(j); self.k.append(k); self.int_.append(it)
        self.di.append(di); self.dj.append(dj); self.dk.append(dk)
        self.nsx.append(x); self.nsy.append(y); self.nsz.append(z)

    def __len__(self):
        return len(self.i)


def forcing_points(o, grid, topo2d, ext_cells, axis, nx, ny):
    """The points of one component: ext_forcing3d for each cell of
    ext_cells (0-based, in that order), then forcing3d."""
    cx, cy, cz = grid
    fp = Points()
    for c in ext_cells:
        k = c // (nx * ny) + 1
        j = (c // nx) % ny + 1
        i = c % nx + 1
        _ext_forcing3d(o, int(i), int(j), int(k), cx, cy, cz, topo2d, fp, axis)
    _forcing3d(o, cx, cy, cz, topo2d, fp, nx, ny)
    return fp


def _gamma(topo2d, cz, i, j, k, n):
    num = cz[k - 1] - topo2d[i - 1, j - 1]
    den = topo2d[i + DI[n] - 1, j + DJ[n] - 1] - topo2d[i - 1, j - 1]
    return num / den


def _forcing3d(o, cx, cy, cz, topo2d, fp: "Points", nx, ny):
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


def _ext_forcing3d(o, i, j, k, cx, cy, cz, topo2d, fp: "Points", axis):
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


FILLED_CELL_1, FILLED_CELL_2 = 512, 1024


def fill_cells(fl, tx, ty, tc, z, zb, nx, ny, nz):
    """fill_cells (ghost.f:3052-3261) level by level in numpy: the flags with
    closed cells made filled_cell_1, the open-face bits bd and the open
    volume fraction vf (the checker of the library's pdac_ib_fill_cells)."""
    fl = np.array(fl, copy=True)
    n = nx * ny * nz
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
        zk, zbk, zbm1 = z[k - 1], zb[k - 1], zb[k - 2]
        bb = np.zeros((ny, nx), dtype=np.int64)
        vv = np.zeros((ny, nx))

        def side(open_, nbf, bit):
            nonlocal bb, vv
            f1_ = open_ & (nbf == FILLED_CELL_1)
            f2_ = open_ & (nbf == FILLED_CELL_2)
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
        lay[comp & (vv == 0.0)] = FILLED_CELL_1
    return fl, bd, vf
