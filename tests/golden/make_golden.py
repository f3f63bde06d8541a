"""Write the golden fixtures of tests/golden/*.npz.

Each fixture holds a small seeded case (its settings and initial fields) and
the state the reference solver reached after ``nsteps`` steps, read from the
restart file it wrote, plus the number of pressure-velocity sweeps it
reported per step.  The reference executable is built in the build container
from the reference's own sources by oracle/build_ref.sh; it never leaves it.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pdac_amd.cases import eruption_column, mixed_box, shock_tube, summit_cone_3d  # noqa: E402
from tests.golden.refrun import available, run_reference  # noqa: E402

def viscous_box(nsteps=4, **kw):
    """mixed_box on cells of about 0.1 mm (pressures within 10 Pa,
    velocities within 1 m/s, dt 0.2 us) with the viscous stresses, the
    repulsive term and conduction switched on."""
    opts = dict(nsteps=nsteps, bottom=3, cell=1.0e-4, vmax=1.0, prange=(1.0e5, 1.0001e5),
                dt=2.0e-7, gas_viscosity=True, part_viscosity=True, repulsive_model=1)
    opts.update(kw)
    return mixed_box(**opts)


def explicit(case):
    """The same case advanced by the explicit Runge-Kutta solver
    (time_integration = 1, explicit_solver.f)."""
    case.time_integration = 1
    return case


CASES = {
    # 2D Cartesian shock tube (slip walls), one class, MUSCL + minmod
    "shock_tube_2d": lambda: shock_tube(nx=40, nz=8, nsteps=4, dt=1.0e-5, rungekut=1),
    # the same with the 3-stage Runge-Kutta of prog.f:235-318
    "shock_tube_2d_rk3": lambda: shock_tube(nx=40, nz=8, nsteps=2, dt=1.0e-5, rungekut=3),
    # 2D Cartesian, 2 classes, gravity, non-uniform cells, no-slip ground
    "mixed_2d": lambda: mixed_box(nsteps=3, bottom=3),
    # 2D cylindrical, free in/outflow east boundary, stratified atmosphere
    "mixed_2d_cyl_freeio": lambda: mixed_box(nsteps=3, itc=1, west=2, east=4, bottom=3,
                                             top=2, stratification=True),
    # Syamlal-O'Brien drag, pressure model A, Crank-Nicolson interphase terms
    "mixed_2d_variants": lambda: mixed_box(nsteps=2, bottom=3, drag_model=2, pmodel=1,
                                           alpha=0.5, alphagrav=0.5, rungekut=2),
    # other limiters (van Leer mass, superbee velocity; not converging in 500)
    "mixed_2d_limiters": lambda: mixed_box(nsteps=1, bottom=3, mass_limiter=1,
                                           vel_limiter=3, beta=0.3),
    # 3D, 3 classes, walls of both kinds and free in/outflow sides
    "mixed_3d": lambda: mixed_box(three_d=True, nsteps=2, nsolid=3, bottom=3, west=3,
                                  east=4, south=4, north=3, top=2, stratification=True),
    # small eruption column: vent inflow, stratified atmosphere, 2 classes
    "eruption_2d_small": lambda: eruption_column(nx=24, nz=40, nsteps=3, dt=2.0e-3,
                                                 vent_radius=6.0),
    # sustained vent: inlet cells (FIXED_FLOWS type 5) in the ground row
    "eruption_2d_vent": lambda: eruption_column(nx=24, nz=40, nsteps=3, dt=2.0e-3,
                                                vent_radius=6.0, vent="inlet"),
    # free in/outflow lid (bdry.f:540-570): w extrapolated with ord 0 and 1.
    # The reference sets the particle w of the ghost row above the lid from
    # ws(:, nsolid+1), past the end of the array (stale loop index,
    # bdry.f:561); without particles near the lid that read cannot reach the
    # computed cells
    "mixed_2d_freetop": lambda: mixed_box(nsteps=2, bottom=3, top=4, east=4, clean_top=9),
    "mixed_2d_freetop_ord1": lambda: mixed_box(nsteps=2, bottom=3, top=4, ord=1, seed=11, clean_top=9),
    # zero-gradient boundaries (flag 32, inoutflow.f zerograd)
    "mixed_2d_zerograd": lambda: mixed_box(nsteps=2, bottom=3, east=6, top=6, seed=13, clean_top=3),
    # constant heat capacities (icpc 0) and heat capacities at tref (icpc 1)
    "mixed_2d_icpc0": lambda: mixed_box(nsteps=2, bottom=3, icpc=0, seed=17),
    "mixed_2d_icpc1": lambda: mixed_box(nsteps=2, bottom=3, icpc=1, tref=350.0, seed=19),
    # ultrabeta mass limiter and beta-unlimited velocity limiter
    "mixed_2d_beta": lambda: mixed_box(nsteps=2, bottom=3, mass_limiter=4, vel_limiter=0,
                                       beta=0.3, seed=23),
    # 3D cone with immersed boundaries (itp = 1, immb = 1): DEM, forcing
    # points, filled cells and face masks from the reference's own set-up,
    # square inlet vent on the summit platform
    "cone_3d_ib": lambda: summit_cone_3d(nx=24, ny=24, nz=16, d=10.0, height=100.0, sigma=80.0,
                                         rough=3.0, vent_half=15.0, nsteps=5, lpr=1),
    "cone_3d_ib_steep": lambda: summit_cone_3d(nx=20, ny=22, nz=18, d=10.0, height=120.0,
                                               sigma=60.0, rough=5.0, vent_half=15.0, nsteps=4,
                                               seed=2, lpr=1),
    # explicit Runge-Kutta steps (prog.f:192-268 with time_integration = 1)
    "explicit_mixed_2d": lambda: explicit(mixed_box(nsteps=3, bottom=3)),
    "explicit_mixed_2d_cyl_freeio": lambda: explicit(mixed_box(
        nsteps=3, itc=1, west=2, east=4, bottom=3, top=2, stratification=True)),
    "explicit_mixed_2d_variants": lambda: explicit(mixed_box(
        nsteps=2, bottom=3, drag_model=2, pmodel=1, alpha=0.5, alphagrav=0.5, rungekut=2)),
    "explicit_mixed_3d": lambda: explicit(mixed_box(
        three_d=True, nsteps=2, nsolid=3, bottom=3, west=3, east=4, south=4, north=3, top=2,
        stratification=True)),
    "explicit_shock_tube_2d_rk3": lambda: explicit(shock_tube(nx=40, nz=8, nsteps=2, dt=1.0e-5,
                                                              rungekut=3)),
    "explicit_eruption_2d_small": lambda: explicit(eruption_column(nx=24, nz=40, nsteps=3,
                                                                   dt=2.0e-4, vent_radius=6.0)),
    "explicit_eruption_2d_vent": lambda: explicit(eruption_column(nx=24, nz=40, nsteps=3, dt=2.0e-4,
                                                                  vent_radius=6.0, vent="inlet")),
    "explicit_mixed_2d_icpc0": lambda: explicit(mixed_box(nsteps=2, bottom=3, icpc=0, seed=17)),
    "explicit_mixed_2d_zerograd": lambda: explicit(mixed_box(nsteps=2, bottom=3, east=6, west=6,
                                                            seed=13)),
    # viscous stresses (gas and particles, Coulombic repulsion) and heat
    # conduction on cells of 0.1 mm, where they are as large as the fluxes
    "visc_2d": lambda: viscous_box(),
    "visc_2d_gas": lambda: viscous_box(part_viscosity=False, seed=29),
    "visc_2d_cyl_freeio": lambda: viscous_box(itc=1, west=2, east=4, top=2, stratification=True),
    "visc_3d": lambda: viscous_box(three_d=True, nsolid=3, west=3, east=4, south=4, north=3, top=2),
    "explicit_visc_2d_cyl": lambda: explicit(viscous_box(itc=1, west=2, east=4, top=2,
                                                         stratification=True, repulsive_model=0)),
    "explicit_visc_3d": lambda: explicit(viscous_box(three_d=True, nsolid=3, west=3, east=4,
                                                     south=4, north=3, top=2, nsteps=3)),
    # the classical Smagorinsky sub-grid model (iturb = 1) with the viscous
    # terms; cmut large enough that the turbulent viscosity outweighs mug
    "visc_2d_sgs": lambda: viscous_box(iturb=1, cmut=0.5, seed=31),
    "visc_3d_sgs": lambda: viscous_box(three_d=True, nsolid=2, west=3, east=4, south=4, north=3,
                                       top=2, iturb=1, cmut=0.5, seed=37, ny=6),
    "visc_2d_sgs_rk2": lambda: viscous_box(iturb=1, cmut=0.5, seed=41, rungekut=2, itc=1, west=2,
                                           east=4, top=2),
    "explicit_visc_2d_sgs": lambda: explicit(viscous_box(iturb=1, cmut=0.5, seed=43, itc=1, west=2,
                                                         east=4, top=2)),
    "explicit_visc_3d_sgs": lambda: explicit(viscous_box(three_d=True, nsolid=2, west=3, east=4,
                                                         south=4, north=3, top=2, iturb=1, cmut=0.5,
                                                         seed=47, ny=6, nsteps=3)),
    # output files at the first and the last step (outp) and the restart
    # file (taperd) kept with the state: 2D cylindrical explicit, 3D
    # semi-implicit with three classes
    "explicit_output_2d": lambda: explicit(mixed_box(nsteps=3, itc=1, west=2, east=4, bottom=3,
                                                     top=2, stratification=True, seed=53,
                                                     tpr_steps=3)),
    "output_3d": lambda: mixed_box(three_d=True, nsteps=2, nsolid=3, bottom=3, west=3, east=4,
                                   south=4, north=3, top=2, stratification=True, seed=59,
                                   tpr_steps=2),
}

# Fixtures with a free in/outflow or zero-gradient lid: the reference sets
# the particle w of the ghost row above the lid from ws(:, nsolid+1), past
# the end of the array (the DO index after its loop, bdry.f:561), so those
# ghost values are unspecified and are left out of every comparison; they
# would reach the computed cells through the top row's particle momentum
# flux, so these fixtures keep particles away from the lid
STALE_TOP_GHOST = {"mixed_2d_freetop", "mixed_2d_freetop_ord1", "mixed_2d_zerograd"}


def top_ghost_cells(case):
    """Cell indices of the mesh's top row (k = nz)."""
    nxy = case.nx * (case.ny if case.three_d else 1)
    return np.arange((case.nz - 1) * nxy, case.nz * nxy)


def save(name, case, ref_state, nits, extras=None):
    data = {"settings": np.frombuffer(json.dumps(case.scalars()).encode(), dtype=np.uint8),
            "nits": np.asarray(nits, dtype=np.int32)}
    for k, v in (extras or {}).items():
        data["ref_" + k] = np.asarray(v)
    for k, v in case.fields.items():
        data["in_" + k] = np.asarray(v)
    for k, v in ref_state.fields.items():
        data["out_" + k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)


def load_extras(name):
    """Other reference outputs kept with a fixture (forcing points)."""
    z = np.load(os.path.join(HERE, name + ".npz"))
    return {k[4:]: z[k] for k in z.files if k.startswith("ref_")}


def load(name):
    from pdac_amd.deck import Case
    z = np.load(os.path.join(HERE, name + ".npz"))
    settings = json.loads(bytes(z["settings"]).decode())
    fields = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return Case.from_scalars(settings, fields), out, [int(x) for x in z["nits"]]


def main(names=None):
    if not available():
        print("reference executable not built (oracle/build_ref.sh)")
        return 1
    for name, make in CASES.items():
        if names and name not in names:
            continue
        case = make()
        extras = {}
        st, nits = run_reference(case, extras=extras)
        save(name, case, st, nits, extras)
        print(f"{name}: {case.ntot} cells, {case.nsteps} steps, nit {nits}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
