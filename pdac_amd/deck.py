"""Case description and ``pdac.dat`` input-deck writer.

A :class:`Case` holds what the reference reads from its input deck
(namelists ``&control &model &mesh &boundaries &topography &inlet &dome
&atmosphere &particles &numeric`` and the ``MESH`` card, input.f:159-209 and
input.f:790-823) plus the initial primary fields, which enter through a
``pdac.res`` restart file (``restart_mode='restart'``).  The same object drives
the GPU solver, so a deck run by the reference and a run on the GPU start from
identical bits.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import constants as C
from .restart import RestartState, write_restart


@dataclass
class Case:
    job_type: str = "2D"          # '2D' or '3D'
    itc: int = 0                  # 1: 2D cylindrical
    nx: int = 10
    ny: int = 1
    nz: int = 10
    dx: Optional[np.ndarray] = None   # cell sizes including boundary cells
    dy: Optional[np.ndarray] = None
    dz: Optional[np.ndarray] = None
    # boundary types as written in the deck (flag = 2**(type-1)), grid.f:385-470
    west: int = 2
    east: int = 2
    south: int = 2
    north: int = 2
    bottom: int = 2
    top: int = 2
    ord: int = 0
    bl_row0: int = 1
    # particles (input.f:399-407 names)
    nsolid: int = 1
    diameter: List[float] = field(default_factory=lambda: [100.0])
    density: List[float] = field(default_factory=lambda: [2500.0])
    sphericity: List[float] = field(default_factory=lambda: [1.0])
    viscosity: List[float] = field(default_factory=lambda: [0.5])
    specific_heat: List[float] = field(default_factory=lambda: [1200.0])
    thermal_conductivity: List[float] = field(default_factory=lambda: [2.0])
    twophase_limit: List[float] = field(default_factory=lambda: [1.0e-8])
    # gas: species present, as reference gas-type indices (1-based)
    gas_type: List[int] = field(default_factory=lambda: [5, 6])
    # model
    icpc: int = 2
    tref: float = 298.0
    # viscous stresses and heat conduction (&model, input.f:247-251);
    # the turbulence models stay off (iturb = iss = 0)
    gas_viscosity: bool = False
    part_viscosity: bool = False
    repulsive_model: int = 0
    # gas sub-grid turbulence: 0 off, 1 classical Smagorinsky (modturbo = 1)
    iturb: int = 0
    cmut: float = 0.1
    drag_model: int = 1
    pmodel: int = 2
    gravx: float = 0.0
    gravy: float = 0.0
    gravz: float = 0.0
    # atmosphere used by free in/outflow boundaries
    p_ground: float = 1.01325e5
    t_ground: float = 288.15
    stratification: bool = False
    void_fraction: float = 1.0 - 1.0e-10
    atm_ygc: List[float] = field(default_factory=lambda: [0, 0, 0, 0, 0, 1.0, 0])
    # numerics (input.f:452-466)
    time_integration: int = 0
    rungekut: int = 1
    beta: float = 0.25
    muscl: int = 1
    mass_limiter: int = 2
    vel_limiter: int = 2
    inmax: int = 8
    maxout: int = 500
    delg: float = 1.0e-8
    omega: float = 1.1
    rlim: float = 1.0e-8
    flim: float = 1.0e-6
    alpha: float = 1.0
    alphagrav: float = 0.0
    # control
    dt: float = 1.0e-5
    nsteps: int = 1
    time: float = 0.0
    cstop: bool = False
    # output files every tpr_steps steps (tpr = tpr_steps * dt); 0: only the
    # one of the first step
    tpr_steps: int = 0
    # topography (&topography itp, &boundaries immb): the DEM is a Gaussian
    # cone made by pdac_amd.topography.cone_dem from the parameters in
    # 'dem' (written as the reference's ASCII DEM file beside the deck)
    itp: int = 0
    immb: int = 0
    dem: Optional[dict] = None
    dem_cellsize: float = 10.0
    dem_filtersize: float = 0.0
    lpr: int = 0
    # FIXED_FLOWS blocks (input.f:832-906): dicts with 'typ' (cell flag
    # 2**(typ-1), grid.f:415-434) and 1-based inclusive bounds 'x', 'y', 'z'
    # as (lo, hi); typ 1, 5 and 8 also carry the block's fixed values, which
    # a restart run reads but does not apply (setup.f:106-140: the restart
    # file holds every cell's values)
    blocks: List[dict] = field(default_factory=list)
    # initial primary fields, keyed like RestartState.fields
    fields: Dict[str, np.ndarray] = field(default_factory=dict)

    # ------------------------------------------------------------------
    @property
    def three_d(self) -> bool:
        return self.job_type.upper() == "3D"

    @property
    def ntot(self) -> int:
        return self.nx * self.ny * self.nz

    def cell_sizes(self):
        dx = np.asarray(self.dx, dtype=np.float64)
        dz = np.asarray(self.dz, dtype=np.float64)
        dy = np.asarray(self.dy, dtype=np.float64) if self.three_d else np.zeros(1)
        return dx, dy, dz

    def flags(self) -> np.ndarray:
        """Cell flags, i fastest, as the reference's set-up leaves them:
        ``flic`` (grid.f:375-470), with topography ``set_flags``
        (topo.f:1031-1112), the near-ground layer of ``export_topography``
        (topo.f:1116-1162) and, with immersed boundaries, the flags of
        ``set_forcing`` and ``fill_cells`` (see :meth:`immersed`)."""
        if self.itp:
            if self.immb:
                return self.immersed().flags.copy()
            return self._topo()[1].copy()
        from .topography import export_layer
        return export_layer(self, self.flags_flic())

    def dem_object(self):
        from .topography import cone_dem
        d = dict(self.dem or {})
        dx, dy, _ = self.cell_sizes()
        return cone_dem(0.5 * float(np.sum(dx)), 0.5 * float(np.sum(dy)), **d)

    def _topo(self):
        """(TopoSetup, flags after export_topography), computed once."""
        cached = self.__dict__.get("_topo_cache")
        if cached is None:
            from .topography import export_layer, topo_setup
            ts = topo_setup(self, self.dem_object())
            cached = (ts, export_layer(self, ts.flags.copy()))
            self.__dict__["_topo_cache"] = cached
        return cached

    def immersed(self):
        """The immersed-boundary set-up (pdac_amd.immersed), computed once."""
        cached = self.__dict__.get("_ib_cache")
        if cached is None:
            from .immersed import immersed_setup
            ts, fl = self._topo()
            cached = immersed_setup(self, ts, fl)
            self.__dict__["_ib_cache"] = cached
        return cached

    def flags_flic(self) -> np.ndarray:
        """Cell flags of ``flic`` (grid.f:375-470): boundary faces and
        FIXED_FLOWS blocks.  The near-ground layer of export_topography
        turns the fluid cells of each column that lie within three rows above
        the last no-slip cell met so far into ``bl_cell`` (5), which disables
        the MUSCL corrections there.  Its row counter is not initialised in
        the reference and carries over from column to column; it starts at
        ``bl_row0`` here (1, the bottom boundary row, which is what the
        reference build does for a deck whose bottom is not a no-slip
        wall).

        The latter turns the fluid cells of each column that lie within three
        rows above the last no-slip cell met so far into ``bl_cell`` (5),
        which disables the MUSCL corrections there.  Its row counter is not
        initialised in the reference and carries over from column to column;
        it starts at ``bl_row0`` here (1, the bottom boundary row, which is
        what the reference build does for a deck whose bottom is not a no-slip
        wall)."""
        nx, ny, nz = self.nx, self.ny, self.nz
        if not self.three_d:
            fl = np.full((nz, 1, nx), C.FLUID, dtype=np.int32)
            fl[0, 0, 1:nx - 1] = 2 ** (self.bottom - 1)
            fl[nz - 1, 0, 1:nx - 1] = 2 ** (self.top - 1)
            fl[:, 0, 0] = 2 ** (self.west - 1)
            fl[:, 0, nx - 1] = 2 ** (self.east - 1)
        else:
            fl = np.full((nz, ny, nx), C.FLUID, dtype=np.int32)
            fl[0, 1:ny - 1, 1:nx - 1] = 2 ** (self.bottom - 1)
            fl[nz - 1, 1:ny - 1, 1:nx - 1] = 2 ** (self.top - 1)
            fl[:, :, 0] = 2 ** (self.west - 1)
            fl[:, :, nx - 1] = 2 ** (self.east - 1)
            fl[:, 0, :] = 2 ** (self.south - 1)
            fl[:, ny - 1, :] = 2 ** (self.north - 1)
        for b in self.blocks:
            (xl, xh), (zl, zh) = b["x"], b["z"]
            yl, yh = b["y"] if self.three_d else (1, 1)
            v = 2 ** (b["typ"] - 1)
            if self.three_d and b["typ"] == 8:
                v = C.DOME_CELL
            fl[zl - 1:zh, yl - 1:yh, xl - 1:xh] = v
        return fl.reshape(-1)

    def _block_values(self, b) -> List[str]:
        """The three value records of a fixed-flow block (input.f:845-868),
        taken from the initial fields at the block's first cell."""
        def d(x):
            return repr(float(x)).replace("e", "D") if "e" in repr(float(x)) else repr(float(x)) + "D0"
        f = self.fields
        yl = b["y"][0] if self.three_d else 1
        c = (b["x"][0] - 1) + (yl - 1) * self.nx + (b["z"][0] - 1) * self.nx * self.ny
        g = lambda k, s=None: (0.0 if k not in f else  # noqa: E731
                               float(np.asarray(f[k]).reshape(-1, self.ntot)[s or 0][c]))
        rl = np.asarray(self.density, dtype=np.float64)
        eps = [g("rlk", s) / rl[s] for s in range(self.nsolid)]
        gas = 1.0 - sum(eps)
        rows = []
        if self.three_d:
            rows.append(" " + " ".join(d(x) for x in (g("ug"), g("vg"), g("wg"), g("p"), gas, g("tg"))))
            rows.append(" " + " ".join(" ".join(d(x) for x in (g("us", s), g("vs", s), g("ws", s), eps[s],
                                                                   g("ts", s))) for s in range(self.nsolid)))
        else:
            rows.append(" " + " ".join(d(x) for x in (g("ug"), g("wg"), g("p"), gas, g("tg"))))
            rows.append(" " + " ".join(" ".join(d(x) for x in (g("us", s), g("ws", s), eps[s], g("ts", s)))
                                       for s in range(self.nsolid)))
        conc = [0.0] * 7
        ygc = np.asarray(f.get("ygc", np.zeros((len(self.gas_type), self.ntot)))).reshape(-1, self.ntot)
        for k, t in enumerate(self.gas_type):
            conc[t - 1] = float(ygc[k][c])
        rows.append(" " + " ".join(d(x) for x in conc))
        return rows

    # ------------------------------------------------------------------
    def scalars(self) -> dict:
        """All non-field settings, JSON-serialisable."""
        out = {}
        for k, v in self.__dict__.items():
            if k == "fields" or k.startswith("_"):
                continue
            if isinstance(v, np.ndarray):
                v = v.tolist()
            elif isinstance(v, (list, tuple)):
                v = [float(x) if isinstance(x, (float, np.floating)) else
                     (int(x) if isinstance(x, (int, np.integer)) else x) for x in v]
            elif isinstance(v, np.generic):
                v = v.item()
            out[k] = v
        return out

    @classmethod
    def from_scalars(cls, d: dict, fields=None) -> "Case":
        kw = dict(d)
        for k in ("dx", "dy", "dz"):
            if kw.get(k) is not None:
                kw[k] = np.asarray(kw[k], dtype=np.float64)
        c = cls(**kw)
        if fields:
            c.fields = {k: np.array(v, dtype=np.float64) for k, v in fields.items()}
        return c

    def restart_state(self) -> RestartState:
        st = RestartState(self.time, self.nx, self.ny, self.nz, self.nsolid,
                          len(self.gas_type), 0, list(self.gas_type),
                          self.three_d)
        st.fields = {k: np.array(v, dtype=np.float64) for k, v in self.fields.items()}
        return st

    def write_restart(self, path: str) -> None:
        write_restart(path, self.restart_state())

    def write_deck(self, path: str) -> None:
        def b(x):
            return "T" if x else "F"

        def d(x):
            return repr(float(x)).replace("e", "D") if "e" in repr(float(x)) else repr(float(x)) + "D0"

        def arr(xs):
            return ", ".join(d(x) for x in xs)

        tstop = self.time + self.nsteps * self.dt
        lines = []
        lines.append("&control")
        lines.append(f" run_name = 'pdac_amd', job_type = '{self.job_type}',")
        lines.append(" restart_mode = 'restart',")
        lines.append(f" time = {d(self.time)}, tstop = {d(tstop)}, dt = {d(self.dt)},")
        tpr = self.tpr_steps * self.dt if self.tpr_steps > 0 else 1.0e6
        lines.append(f" lpr = {self.lpr}, imr = 0, tpr = {d(tpr)}, tdump = {d(self.nsteps * self.dt)},")
        lines.append(f" nfil = 0, max_seconds = {d(1.0e7)}, cstop = {b(self.cstop)}")
        lines.append("/")
        lines.append("&model")
        lines.append(f" icpc = {self.icpc}, tref = {d(self.tref)}, irex = 1,")
        lines.append(f" gas_viscosity = {b(self.gas_viscosity)}, part_viscosity = {b(self.part_viscosity)},")
        lines.append(f" iss = 0, iturb = {self.iturb}, modturbo = 1, cmut = {d(self.cmut)},")
        lines.append(f" repulsive_model = {self.repulsive_model},")
        lines.append(f" drag_model = {self.drag_model}, pmodel = {self.pmodel},")
        lines.append(f" gravx = {d(self.gravx)}, gravy = {d(self.gravy)}, gravz = {d(self.gravz)},")
        lines.append(f" ngas = {len(self.gas_type)}, isink = 0")
        lines.append("/")
        lines.append("&mesh")
        dx, dy, dz = self.cell_sizes()
        lines.append(f" nx = {self.nx}, ny = {self.ny}, nz = {self.nz}, itc = {self.itc},")
        lines.append(" iuni = 0, grigen = 0, zzero = 0.0D0")
        lines.append("/")
        lines.append("&boundaries")
        lines.append(f" west = {self.west}, east = {self.east}, south = {self.south},")
        lines.append(f" north = {self.north}, bottom = {self.bottom}, top = {self.top},")
        lines.append(f" immb = {self.immb}, ibl = 0, ord = {self.ord}")
        lines.append("/")
        lines.append("&topography")
        if self.itp:
            lines.append(f" itp = {self.itp}, dem_file = 'topo.dat', iavv = 0, ismt = 1,")
            lines.append(f" cellsize = {d(self.dem_cellsize)}, filtersize = {d(self.dem_filtersize)},")
            lines.append(" itrans = T, nocrater = T")
            self.dem_object().write(os.path.join(os.path.dirname(os.path.abspath(path)), "topo.dat"))
        else:
            lines.append(" itp = 0")
        lines.append("/")
        lines.append("&inlet")
        lines.append(" ivent = 0")
        lines.append("/")
        lines.append("&dome")
        lines.append(" idome = 0")
        lines.append("/")
        lines.append("&atmosphere")
        lines.append(f" p_ground = {d(self.p_ground)}, t_ground = {d(self.t_ground)},")
        lines.append(f" stratification = {b(self.stratification)}, void_fraction = {d(self.void_fraction)},")
        names = ["atm_O2", "atm_N2", "atm_CO2", "atm_H2", "atm_H2O", "atm_Air", "atm_SO2"]
        lines.append(" " + ", ".join(f"{n} = {d(v)}" for n, v in zip(names, self.atm_ygc)))
        lines.append("/")
        lines.append("&particles")
        lines.append(f" nsolid = {self.nsolid},")
        lines.append(f" diameter = {arr(self.diameter)},")
        lines.append(f" density = {arr(self.density)},")
        lines.append(f" sphericity = {arr(self.sphericity)},")
        lines.append(f" viscosity = {arr(self.viscosity)},")
        lines.append(f" specific_heat = {arr(self.specific_heat)},")
        lines.append(f" thermal_conductivity = {arr(self.thermal_conductivity)},")
        lines.append(f" twophase_limit = {arr(self.twophase_limit)}")
        lines.append("/")
        lines.append("&numeric")
        lines.append(f" time_integration = {self.time_integration}, rungekut = {self.rungekut},")
        lines.append(f" beta = {d(self.beta)}, muscl = {self.muscl}, mass_limiter = {self.mass_limiter},")
        lines.append(f" vel_limiter = {self.vel_limiter}, inmax = {self.inmax}, maxout = {self.maxout},")
        lines.append(f" omega = {d(self.omega)}, delg = {d(self.delg)}, optimization = 1,")
        lines.append(" implicit_fluxes = F, implicit_enthalpy = F, update_eosg = F,")
        lines.append(f" rlim = {d(self.rlim)}, flim = {d(self.flim)}, alpha = {d(self.alpha)},")
        lines.append(f" alphagrav = {d(self.alphagrav)}, ctu = 0, adapt_dt = F")
        lines.append("/")
        lines.append("'MESH'")

        def rows(vals):
            # a few values per record; list-directed input reads on
            vals = [d(x) for x in vals]
            return [" " + " ".join(vals[i:i + 4]) for i in range(0, len(vals), 4)]

        lines += rows(dx)
        if self.three_d:
            lines += rows(dy)
        lines += rows(dz)
        lines.append("'FIXED_FLOWS'")
        lines.append(f" {len(self.blocks)}")
        for b in self.blocks:
            (xl, xh), (zl, zh) = b["x"], b["z"]
            if self.three_d:
                (yl, yh) = b["y"]
                lines.append(f" {b['typ']} {xl} {xh} {yl} {yh} {zl} {zh}")
            else:
                lines.append(f" {b['typ']} {xl} {xh} {zl} {zh}")
            if b["typ"] in (1, 5, 8):
                lines.extend(self._block_values(b))
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
