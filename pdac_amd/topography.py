"""Topography of 3D cases: the DEM file and the reference's set-up from it.

A 3D case with topography (``itp = 1``) reads a digital elevation model in
the reference's ASCII format (read_dem_ascii, topo.f:350-469), resamples it
to the resolution ``cellsize`` (resize_dem, topo.f:474-580), moves the mesh
to the DEM's coordinates (compute_UTM_coords, topo.f:731-751), interpolates
it to the cell centres, rounds it to centimetres and lifts the mesh so that
the lowest ground point is the bottom face (set_profile, topo.f:145-304),
and marks the cells under the ground as no-slip walls with a three-cell
boundary layer above (set_flags, topo.f:1031-1112).  This module restates
that set-up in the reference's arithmetic (same operations, same order, in
IEEE double), so the flags and the immersed-boundary geometry built from it
(:mod:`pdac_amd.immersed`) are those of the reference.

The DEM of the 3D summit-vent case is synthetic: a Gaussian cone with a
flat summit plateau for the vent, plus seeded fractal roughness.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import constants as C


@dataclass
class Dem:
    """An ASCII DEM: node coordinates xll + (i-1) cell (west to east) and
    yll + (j-1) cell (south to north); z[j, i] in metres."""
    xll: float
    yll: float
    cell: float
    z: np.ndarray          # (ny, nx), south row first
    noval: int = -9999

    def write(self, path: str) -> None:
        """The reference's ASCII DEM: nx, ny, xll, yll, cellsize, noval on
        one line each, then the elevations row by row from north to south
        (topo.f:372-445)."""
        ny, nx = self.z.shape
        with open(path, "w") as fh:
            fh.write(f"{nx}\n{ny}\n{_num(self.xll)}\n{_num(self.yll)}\n{_num(self.cell)}\n{self.noval}\n")
            for j in range(ny - 1, -1, -1):
                fh.write(" ".join(f"{v:.2f}" for v in self.z[j]) + "\n")

    @staticmethod
    def read(path: str) -> "Dem":
        with open(path) as fh:
            toks = fh.read().split()
        nx, ny = int(toks[0]), int(toks[1])
        xll, yll, cell = float(toks[2]), float(toks[3]), float(toks[4])
        noval = int(float(toks[5]))
        vals = np.array([float(t) for t in toks[6:6 + nx * ny]])
        z = vals.reshape(ny, nx)[::-1].copy()
        return Dem(xll, yll, cell, z, noval)

    def nodes(self):
        ny, nx = self.z.shape
        x = np.array([self.xll + i * self.cell for i in range(nx)])
        y = np.array([self.yll + j * self.cell for j in range(ny)])
        return x, y


def _num(v: float) -> str:
    return repr(float(v))


def fractal_roughness(shape, amplitude, seed, slope=2.0):
    """Seeded fractal (power-law spectrum) surface of the given shape with
    maximum absolute value ``amplitude``: Fourier amplitudes proportional to
    |k|^-slope with uniform random phases."""
    rng = np.random.default_rng(seed)
    ny, nx = shape
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    kk = np.sqrt(kx * kx + ky * ky)
    kk[0, 0] = np.inf
    amp = kk ** (-slope)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    f = np.fft.irfft2(amp * np.exp(1j * phase), s=shape)
    f = f - f.mean()
    return f * (amplitude / np.max(np.abs(f)))


def cone_dem(half_x, half_y, cell=10.0, height=1500.0, sigma=1500.0, rough=30.0,
             plateau=None, seed=1234, margin=None, plateau_z=None) -> Dem:
    """Gaussian cone of ``height`` and width ``sigma`` centred on (0, 0),
    cut flat at the height it reaches at radius ``plateau`` (or at
    ``plateau_z`` when given; a summit platform for the vent; None: no cut),
    plus fractal roughness of amplitude ``rough`` outside the platform.
    Nodes every ``cell`` metres covering [-half - margin, half + margin] in
    x and y; elevations rounded to centimetres."""
    margin = 4 * cell if margin is None else margin
    nx = int(round(2 * (half_x + margin) / cell)) + 1
    ny = int(round(2 * (half_y + margin) / cell)) + 1
    xll = -(half_x + margin)
    yll = -(half_y + margin)
    x = np.array([xll + i * cell for i in range(nx)])
    y = np.array([yll + j * cell for j in range(ny)])
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    z = height * np.exp(-r2 / (2.0 * sigma * sigma))
    if rough:
        z = z + fractal_roughness(z.shape, rough, seed)
    if plateau is not None:
        top = height * np.exp(-(plateau ** 2) / (2.0 * sigma * sigma)) if plateau_z is None else plateau_z
        flat = (np.abs(x)[None, :] <= plateau) & (np.abs(y)[:, None] <= plateau)
        z = np.where(flat, top, np.minimum(z, top))
    z = np.round(z * 100.0) / 100.0
    return Dem(xll, yll, cell, z)


# ----------------------------------------------------------------------
# the reference's arithmetic
def fortran_nint(a):
    """NINT: nearest integer, halves away from zero."""
    a = np.asarray(a, dtype=np.float64)
    return (np.sign(a) * np.floor(np.abs(a) + 0.5)).astype(np.int64)


def interp_2d(x1, y1, f1, x2, y2):
    """Bilinear interpolation of f1 (x1 x y1 nodes, f1[i, j]) to the nodes
    x2 x y2 (interp_2d, arrfil.f:410-511): tx(ii) is the first node i of x1
    with x1(i) >= x2(ii) (1-based), and the weights are the distances to
    the nodes tx-1 and tx.  Returns f2[i, j]."""
    x1, y1, x2, y2 = (np.asarray(a, dtype=np.float64) for a in (x1, y1, x2, y2))
    tx = np.searchsorted(x1, x2, side="left")  # 0-based index of the first x1 >= x2
    ty = np.searchsorted(y1, y2, side="left")
    if np.any(tx < 1) or np.any(ty < 1) or np.any(tx >= x1.size) or np.any(ty >= y1.size):
        raise ValueError("interp_2d: a target node lies outside the interior of the source grid")
    d1x = x2 - x1[tx - 1]
    d2x = x1[tx] - x2
    alpha = d1x / (d1x + d2x)
    d1y = y2 - y1[ty - 1]
    d2y = y1[ty] - y2
    beta = d1y / (d1y + d2y)
    a = alpha[:, None]
    b = beta[None, :]
    tp1 = a * f1[np.ix_(tx, ty)] + (1.0 - a) * f1[np.ix_(tx - 1, ty)]
    tp2 = a * f1[np.ix_(tx, ty - 1)] + (1.0 - a) * f1[np.ix_(tx - 1, ty - 1)]
    return b * tp1 + (1.0 - b) * tp2


@dataclass
class MeshCoords:
    """Cell-centre and face coordinates of the mesh (grid_setup,
    grid.f:284-303, 1-based arrays stored 0-based)."""
    x: np.ndarray
    xb: np.ndarray
    y: np.ndarray
    yb: np.ndarray
    z: np.ndarray
    zb: np.ndarray


def mesh_coords(dx, dy, dz, zzero=0.0) -> MeshCoords:
    def axis(d, origin):
        d = np.asarray(d, dtype=np.float64)
        n = d.size
        b = np.empty(n)
        c = np.empty(n)
        b[0] = origin
        c[0] = -0.5 * d[0] + b[0]
        for i in range(1, n):
            b[i] = b[i - 1] + d[i]
            c[i] = b[i] - 0.5 * d[i]
        return c, b
    x, xb = axis(dx, 0.0)
    y, yb = axis(dy, 0.0)
    z, zb = axis(dz, zzero)
    return MeshCoords(x, xb, y, yb, z, zb)


@dataclass
class TopoSetup:
    """Everything the reference derives from the DEM before the immersed
    boundaries: the moved mesh, the resampled DEM (xtop, ytop, ztop2d), the
    ground at the cell centres (rounded), ord2d and the cell flags."""
    mesh: MeshCoords
    xtop: np.ndarray
    ytop: np.ndarray
    ztop2d: np.ndarray     # [i, j]
    topo2d: np.ndarray     # [i, j], cell centres, centimetres
    ord2d: np.ndarray      # [i, j], 1-based last cell whose bottom face is under the ground
    flags: np.ndarray      # (ntot,) i fastest


def topo_setup(case, dem: Dem) -> TopoSetup:
    """import_topography for a 3D case (topo.f:93-143) followed by
    set_flags; ``case.flags_flic()`` gives the flags before topography."""
    nx, ny, nz = case.nx, case.ny, case.nz
    dx, dy, dz = case.cell_sizes()
    dx, dy, dz = (np.asarray(a, dtype=np.float64) for a in (dx, dy, dz))
    m = mesh_coords(dx, dy, dz)
    # grid_setup with grigen = 0 (grid.f:228-239)
    dxmax, dymax = float(np.max(dx)), float(np.max(dy))
    domain_x = float(np.sum(dx))
    domain_y = float(np.sum(dy))
    iv = int(fortran_nint(nx / 2.0))
    jv = int(fortran_nint(ny / 2.0))
    cellsize = case.dem_cellsize
    center_x, center_y, alpha_x, alpha_y = 0.0, 0.0, 0.5, 0.5
    # read_dem_ascii: node coordinates
    xdem, ydem = dem.nodes()
    zdem = dem.z.T  # [i, j]
    # resize_dem (topo.f:474-580)
    newsizex = domain_x + 2.0 * max(dxmax, cellsize)
    newsizey = domain_y + 2.0 * max(dymax, cellsize)
    ntx = int(newsizex / cellsize) + 1
    nty = int(newsizey / cellsize) + 1
    xll = center_x - alpha_x * domain_x - dxmax
    yll = center_y - alpha_y * domain_y - dymax
    xtop = np.empty(ntx)
    ytop = np.empty(nty)
    xtop[0], ytop[0] = xll, yll
    for i in range(1, ntx):
        xtop[i] = xtop[i - 1] + cellsize
    for j in range(1, nty):
        ytop[j] = ytop[j - 1] + cellsize
    if xll < xdem[0] or xtop[-1] > xdem[-1] or yll < ydem[0] or ytop[-1] > ydem[-1]:
        raise ValueError("the mesh extends beyond the DEM (resize_dem)")
    ztop2d = interp_2d(xdem, ydem, zdem, xtop, ytop)
    if case.dem_filtersize >= cellsize:
        raise ValueError("DEM filtering (filtersize >= cellsize) is not restated here")
    # compute_UTM_coords (topo.f:731-751)
    tx = center_x - m.x[iv - 1]
    ty = center_y - m.y[jv - 1]
    m.x = m.x + tx
    m.xb = m.xb + tx
    m.y = m.y + ty
    m.yb = m.yb + ty
    # set_profile (topo.f:228-295)
    if np.max(xtop) < np.max(m.x) or np.max(ytop) < np.max(m.y):
        raise ValueError("computational domain exceeds topography")
    topo2d = interp_2d(xtop, ytop, ztop2d, m.x, m.y)
    topo2d = fortran_nint(topo2d * 1.0e2).astype(np.float64) * 1.0e-2
    transl_z = float(np.min(topo2d)) - m.zb[0]
    m.z = m.z + transl_z
    m.zb = m.zb + transl_z
    ord2d = np.zeros((nx, ny), dtype=np.int64)
    for k in range(nz):
        ord2d = np.where(m.zb[k] <= topo2d, k + 1, ord2d)
    # set_flags (topo.f:1031-1112) on the flags of flic
    fl = case.flags_flic().copy()
    flv = fl.reshape(nz, ny, nx)
    kk = np.arange(1, nz + 1)[:, None, None]
    below = kk <= ord2d.T[None, :, :]
    flv[below & (flv != C.INLET_CELL)] = C.NOSLIP_WALL
    flv[(~below) & (flv == C.NOSLIP_WALL)] = C.FLUID
    # the three rows above the ground of each column and the 7 x 7 columns
    # around it, by linear index as the reference computes it (an offset
    # past a mesh edge lands in the neighbouring row)
    ntot = nx * ny * nz
    ii, jj = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="ij")
    q = ord2d
    idx = []
    for dk in (1, 2, 3):
        for dj in range(-3, 4):
            for di in range(-3, 4):
                idx.append((ii + di) + (jj + dj - 1) * nx + (q + dk - 1) * nx * ny)
    idx = np.concatenate([a.reshape(-1) for a in idx])
    if idx.min() < 1 or idx.max() > ntot:
        raise ValueError("set_flags: boundary layer outside the mesh")
    hit = np.zeros(ntot, dtype=bool)
    hit[idx - 1] = True
    fl[hit & (fl == C.FLUID)] = C.BL_CELL
    return TopoSetup(m, xtop, ytop, ztop2d, topo2d, ord2d, fl)


def export_layer(case, fl):
    """export_topography (topo.f:1114-1162): fluid cells within three rows
    above the last no-slip cell met (in the loop over j, i, k, the row
    counter carrying over from column to column) become bl_cell.

    The serial loop, vectorised: along a column the counter is the highest
    no-slip row met so far (rows only grow within a column); before the
    column's first no-slip cell it is the last one of the nearest earlier
    column that has one (the initial bl_row0 if none)."""
    nx, ny, nz = case.nx, case.ny, case.nz
    flv = fl.reshape(nz, ny * nx)                    # columns in the loop's j, i order
    rows = np.arange(1, nz + 1, dtype=np.int64)[:, None]
    ns = flv == C.NOSLIP_WALL
    q_in = np.maximum.accumulate(np.where(ns, rows, 0), axis=0)   # 0 until the first no-slip
    last = q_in[-1]                                  # last no-slip row of each column, 0 if none
    idx = np.where(last > 0, np.arange(last.size), -1)
    np.maximum.accumulate(idx, out=idx)              # nearest column at or before with one
    carry = np.where(idx >= 0, last[np.maximum(idx, 0)], case.bl_row0)
    carry = np.concatenate([[case.bl_row0], carry[:-1]])   # what the previous columns leave
    q = np.where(q_in > 0, q_in, carry[None, :])
    flv[(flv == C.FLUID) & (rows <= q + 3)] = C.BL_CELL
    return fl
