"""Immersed boundaries of a 3D case with topography: the reference's set-up.

With ``immb = 1`` the reference forces the velocity components that lie
under the ground next to the flow (forcing points), so that the no-slip
condition holds on the interpolated ground surface rather than on cell
faces.  Its set-up, restated here in the reference's arithmetic:

* set_forcing (immb.f:73-516): the ground on the three staggered grids
  (interpolate_dem, topo.f:970-1029), cells with all but one or two faces
  under the ground become ``filled_cell_1/2`` and get external forcing
  points (ext_forcing3d, immb.f:1071-1261), the forcing points under the
  ground (forcing3d, immb.f:913-1069), and ``immb_cell`` flags; the
  forcing points come from the library's pdac_ib_forcing (C++, host
  threads over the mesh columns);
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

from dataclasses import dataclass
from typing import List

import numpy as np

from . import constants as C
from .topography import TopoSetup, interp_2d

@dataclass
class ForcingPoints:
    """Forcing points of one velocity component, in the reference's order
    (TYPE forcing_point, immb.f:33-45): int32 arrays i, j, k (1-based),
    int_ (interpolation type), di, dj, dk (direction), float64 arrays nsx,
    nsy, nsz (the no-slip point)."""
    i: np.ndarray
    j: np.ndarray
    k: np.ndarray
    int_: np.ndarray
    di: np.ndarray
    dj: np.ndarray
    dk: np.ndarray
    nsx: np.ndarray
    nsy: np.ndarray
    nsz: np.ndarray

    def __len__(self):
        return len(self.i)


def _cells(fp: ForcingPoints, nx, ny):
    """0-based cell of each point."""
    i, j, k = (np.asarray(v, dtype=np.int64) for v in (fp.i, fp.j, fp.k))
    return (i - 1) + (j - 1) * nx + (k - 1) * nx * ny


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


def _forcing_points(o, grid, topo2d, ext_cells, axis, nz, cap) -> ForcingPoints:
    """The forcing points of one velocity component from the library's
    pdac_ib_forcing: ext_forcing3d for each cell of ext_cells, then forcing3d
    over the interior columns (immb.f:913-1261), on host threads."""
    import ctypes as ct

    from .gpu import lib
    L = lib()
    f = L.pdac_ib_forcing
    i32, i64, f64 = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int64), ct.POINTER(ct.c_double)
    f.argtypes = [ct.c_int32] * 4 + [i32, f64, f64, f64, f64, i64, ct.c_int64, ct.c_int64, i32, f64, i64]
    f.restype = ct.c_int
    nx, ny = o.shape
    o32 = np.ascontiguousarray(o, dtype=np.int32)
    cx, cy, cz = (np.ascontiguousarray(g, dtype=np.float64) for g in grid)
    t = np.ascontiguousarray(topo2d, dtype=np.float64)
    ext = np.ascontiguousarray(ext_cells, dtype=np.int64)
    pts = np.empty((max(cap, 1), 7), dtype=np.int32)
    nsl = np.empty((max(cap, 1), 3), dtype=np.float64)
    count = ct.c_int64(0)
    p = lambda a, ty: a.ctypes.data_as(ty)  # noqa: E731
    rc = f(nx, ny, nz, axis, p(o32, i32), p(cx, f64), p(cy, f64), p(cz, f64), p(t, f64), p(ext, i64),
           ext.size, cap, p(pts, i32), p(nsl, f64), ct.byref(count))
    if rc != 0:
        if count.value > cap:
            raise ValueError(f"set_forcing: {count.value} forcing points along axis {axis}, "
                             f"{cap} counted (the reference would fail)")
        raise ValueError(L.pdac_last_error().decode())
    n = count.value
    return ForcingPoints(*(pts[:n, c].copy() for c in range(7)), *(nsl[:n, c].copy() for c in range(3)))


def _fill_cells(fl, tx, ty, tc, z, zb, nx, ny, nz):
    """fill_cells from the library's pdac_ib_fill_cells (the reference's
    serial cell loop): flags with the closed cells made filled_cell_1, the
    open-face bits bd and the open volume fraction vf."""
    import ctypes as ct

    from .gpu import lib
    L = lib()
    f = L.pdac_ib_fill_cells
    i32, f64 = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double)
    f.argtypes = [ct.c_int32] * 3 + [f64] * 5 + [i32, i32, f64]
    f.restype = ct.c_int
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (tx, ty, tc, z, zb)]
    out_fl = np.ascontiguousarray(fl, dtype=np.int32).copy()
    bd = np.empty(out_fl.size, dtype=np.int32)
    vf = np.empty(out_fl.size, dtype=np.float64)
    rc = f(nx, ny, nz, *(a.ctypes.data_as(f64) for a in arrs), out_fl.ctypes.data_as(i32),
           bd.ctypes.data_as(i32), vf.ctypes.data_as(f64))
    if rc != 0:
        raise ValueError(L.pdac_last_error().decode())
    return out_fl.astype(fl.dtype), bd.astype(np.int64), vf


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
    full = np.full((nz, ny, nx), 255, dtype=np.uint8)   # 255: the mesh's outer cells
    bds = full[1:-1, 1:-1, 1:-1]
    bds[...] = zc > tx[1:-1, 1:-1].T[None]
    bds |= (zc > tx[:-2, 1:-1].T[None]).view(np.uint8) << 1
    bds |= (zt > tci).view(np.uint8) << 2
    bds |= (zbm >= tci).view(np.uint8) << 3
    bds |= (zc > ty[1:-1, 1:-1].T[None]).view(np.uint8) << 4
    bds |= (zc > ty[1:-1, :-2].T[None]).view(np.uint8) << 5
    if np.any(bds == 8):
        raise ValueError("set_forcing: inconsistent immersed boundary (bds = 8)")
    b = full.reshape(-1)

    def of(*vals):   # cells whose open-face bits are one of vals
        lut = np.zeros(256, dtype=bool)
        lut[list(vals)] = True
        return lut[b]

    ext[0] |= of(1, 17, 33, 3)
    ext[0][np.flatnonzero(of(2, 18, 34, 3)) - 1] = True
    ext[1] |= of(16, 18, 17, 48)
    ext[1][np.flatnonzero(of(32, 34, 33, 48)) - nx] = True
    ext[2] |= b == 4
    free = ~np.isin(fl, keep)
    f1 = of(1, 2, 4, 16, 32) & free
    f2 = of(17, 18, 34, 33, 48, 3) & free
    fl[f1] = C.FILLED_CELL_1
    fl[f2] = C.FILLED_CELL_2
    counts = [int(force[a].sum() + ext[a].sum()) for a in range(3)]
    fps = []
    for a in range(3):
        topo2d = st.call(*grids[a], force[a])
        fp = _forcing_points(st.ord2d, grids[a], topo2d, np.flatnonzero(ext[a]), a + 1, nz, counts[a])
        if len(fp) != counts[a]:
            raise ValueError(f"set_forcing: {len(fp)} forcing points along axis {a + 1}, "
                             f"{counts[a]} counted (the reference would fail)")
        # the cells of the points become immb_cell unless inlet, vent or filled
        c = _cells(fp, nx, ny)
        c = c[(fp.i > 1) & (fp.j > 1) & (fp.k > 1)]
        c = c[~np.isin(fl[c], (C.INLET_CELL, C.VENT_CELL, C.FILLED_CELL_1, C.FILLED_CELL_2))]
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
        c = _cells(fp, nx, ny)
        # num: the last point of a cell (points are numbered in order)
        np.maximum.at(num[a], c, np.arange(1, len(fp) + 1, dtype=np.int64))
        di, dj, dk = (np.asarray(v, dtype=np.int64) for v in (fp.di, fp.dj, fp.dk))
        step = di + dj * nx + dk * nx * ny
        it = np.asarray(fp.int_, dtype=np.int64)
        iq.append(c + step)
        iqq.append(np.where((it % 2 != 0) | (it == 0), c + 2 * step, -1))
    # fill_cells (ghost.f:3052-3261), in the reference's cell order
    fl, bd, vf = _fill_cells(fl, tx, ty, tc, m.z, m.zb, nx, ny, nz)
    # face_fractions (ghost.f:3263-3321)
    bmask = np.empty((6, n), dtype=np.int32)   # e, w, n, s, t, b
    comp = (fl & 1) == 1
    stop = np.flatnonzero(comp & (num[0] != 0) & (num[2] != 0))
    ib_last = int(stop[0]) if stop.size else n - 1
    upto = np.zeros(n, dtype=bool)
    upto[:ib_last if stop.size else n] = True
    sel = comp & upto
    for row, bit in ((0, 1), (1, 2), (4, 4), (5, 8), (2, 16), (3, 32)):
        np.logical_or(~sel, (bd & bit) != 0, out=bmask[row], casting="unsafe")
    pos = vf > 0.0
    ivf = np.where(sel, np.where(pos, 1.0 / np.where(pos, vf, 1.0), 0.0), 1.0)
    if stop.size:
        bmask[[0, 1, 4, 5], ib_last] = 0
    return ImmersedSetup(fl, fps, [list(g) for g in grids], num, iq, iqq, bd, vf, bmask, ivf, ib_last)
