"""Ensembles of explicit-step domains (pdac_ens_*): batched, graph-replayed
launches over members of uneven size.

An ensemble step runs, for every member, the same per-cell bodies in the
same order as the member's own pdac_step, so the results are compared bit
for bit: against the members stepped one by one, across tile shapes, and
after the autotuner (which must leave the state untouched)."""
import numpy as np
import pytest

from pdac_amd import abi

pytestmark = pytest.mark.gpu


def members_2d(precision=abi.PREC_FP64, order=abi.ORDER_REDBLACK):
    from pdac_amd.cases import eruption_column, mixed_box
    from pdac_amd.gpu import gpu_solver
    cases = [eruption_column(nx=24, nz=40, nsteps=3, dt=2.0e-4, vent_radius=6.0),
             eruption_column(nx=40, nz=80, nsteps=3, dt=1.5e-4, vent_radius=10.0, seed=5),
             mixed_box(nsteps=3, bottom=3, nx=13, nz=11),
             eruption_column(nx=70, nz=33, nsteps=3, dt=1.0e-4, vent_radius=20.0, seed=9)]
    out = []
    for c in cases:
        c.time_integration = 1
        out.append(gpu_solver(c, device=0, precision=precision, iter_order=order))
    return out


def states(ms):
    return [m.state() for m in ms]


def assert_same(a, b, where):
    for i, (sa, sb) in enumerate(zip(a, b)):
        for k in sb:
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"{where}: member {i} field {k}")


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_FP32FLUX])
def test_ensemble_matches_members(precision):
    from pdac_amd.gpu import Ensemble
    print(f"[ensemble vs members] precision {precision}", flush=True)
    solo = members_2d(precision)
    for m in solo:
        m.step(3)
    want = states(solo)
    batch = members_2d(precision)
    e = Ensemble(batch)
    fails, cfl = e.step(3)
    got = states(batch)
    assert_same(got, want, "ensemble vs one by one")
    assert all(f == 0 for f in fails)
    assert all(c > 0.0 for c in cfl)
    e.close()


def test_ensemble_stage_order_and_rk2():
    from pdac_amd.gpu import Ensemble
    print("[ensemble stage order rk2]", flush=True)
    solo = members_2d(order=abi.ORDER_STAGE)
    for m in solo:
        m.step(2, rungekut=2)
    batch = members_2d(order=abi.ORDER_STAGE)
    e = Ensemble(batch)
    e.step(2, rungekut=2)
    assert_same(states(batch), states(solo), "stage order, rungekut 2")
    e.close()


def test_tiles_and_tuner_do_not_change_results():
    from pdac_amd.gpu import TILE_KERNELS, Ensemble
    print("[ensemble tiles]", flush=True)
    ref = members_2d()
    e0 = Ensemble(ref)
    e0.step(2)
    want = states(ref)
    e0.close()
    for t in range(4):
        ms = members_2d()
        e = Ensemble(ms)
        e.set_tiles([t] * len(TILE_KERNELS))
        e.step(2)
        assert_same(states(ms), want, f"tile {t}")
        e.close()
    ms = members_2d()
    before = states(ms)
    e = Ensemble(ms)
    chosen, times = e.tune(reps=3)
    print("tuned:", chosen, flush=True)
    after = states(ms)
    for i, (sa, sb) in enumerate(zip(after, before)):
        for k in ("p", "ug", "wg", "rgp", "rlk", "tg", "sieg", "sies", "ts", "us", "ws", "ygc"):
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"tuning changed member {i} {k}")
    assert all(len(v) == 4 and all(x > 0.0 for x in v) for k, v in times.items()
               if k in ("xinit", "xflux", "xrhs", "xmass", "ygas_flux"))
    e.step(2)
    assert_same(states(ms), want, "after tuning")
    e.close()


def test_ensemble_3d_members():
    from pdac_amd.cases import mixed_box
    from pdac_amd.gpu import Ensemble, gpu_solver
    print("[ensemble 3D]", flush=True)

    def make():
        out = []
        for seed, (nx, ny, nz) in ((7, (10, 7, 9)), (3, (8, 9, 7))):
            c = mixed_box(three_d=True, nsteps=2, nsolid=3, bottom=3, west=3, east=4, south=4,
                          north=3, top=2, stratification=True, nx=nx, ny=ny, nz=nz, seed=seed)
            c.time_integration = 1
            out.append(gpu_solver(c, device=0))
        return out
    solo = make()
    for m in solo:
        m.step(2)
    batch = make()
    e = Ensemble(batch)
    e.step(2)
    assert_same(states(batch), states(solo), "3D")
    e.close()


def test_ensemble_refuses_viscous_members():
    """The ensemble kernels carry no viscous stresses or conduction: a
    member with gas_viscosity set is refused with an error, not stepped
    without them."""
    from pdac_amd.cases import mixed_box
    from pdac_amd.gpu import Ensemble, gpu_solver
    ms = members_2d()
    c = mixed_box(nsteps=1, bottom=3, gas_viscosity=True)
    c.time_integration = 1
    ms.append(gpu_solver(c, device=0))
    with pytest.raises(Exception, match="viscous"):
        Ensemble(ms)
