"""pdac_explicit_step_noalloc: explicit steps whose Runge-Kutta workspace is
the library's static device array, against pdac_step on a context that holds
its own workspace.  The explicit step's bar: 1e-12 relative per field."""
import copy
import ctypes as ct

import numpy as np
import pytest

from pdac_amd import abi
from tests.golden.make_golden import load

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    d = np.max(np.abs(a - b))
    return d if scale == 0.0 else d / scale


def noalloc(g, nsteps):
    fail, cfl = ct.c_int32(-1), ct.c_double(-1.0)
    rc = g.lib.pdac_explicit_step_noalloc(g.ctx, g.case.dt, g.case.rungekut, nsteps, ct.byref(fail),
                                          ct.byref(cfl))
    return rc, fail.value, cfl.value


def solver(case, time_integration):
    from pdac_amd.gpu import gpu_solver
    c = copy.deepcopy(case)
    c.time_integration = time_integration
    return gpu_solver(c, device=0)


def compare(g, h, where):
    sg, sh = g.state(), h.state()
    worst = max(rel(sg[k], sh[k]) for k in sh)
    bits = all(np.array_equal(sg[k], sh[k]) for k in sh)
    print(f"[noalloc] {where}: max rel diff {worst:.2e}, bit-identical {bits}", flush=True)
    bad = {k: rel(sg[k], sh[k]) for k in sh if rel(sg[k], sh[k]) > TOL}
    assert not bad, f"{where}: " + ", ".join(f"{k}={v:.2e}" for k, v in sorted(bad.items()))


def test_noalloc_matches_step_two_calls_then_smaller():
    """Two calls in a row on one device, the second on a smaller mesh, so the
    array holds the first call's values when the second starts."""
    for name in ("explicit_eruption_2d_small", "explicit_mixed_3d", "explicit_mixed_2d"):
        case, _, _ = load(name)
        print(f"[noalloc] {name}: {case.ntot} cells, {case.nsteps} steps", flush=True)
        want = solver(case, 1)
        want.step(case.nsteps)
        got = solver(case, 0)  # no workspace of its own
        rc, fail, cfl = noalloc(got, case.nsteps)
        assert rc == 0, got.last_error()
        assert fail == 0 and cfl > 0.0
        compare(got, want, name)
        want.close()
        got.close()


def test_noalloc_leaves_the_context_workspace_alone():
    """On a context with its own workspace: noalloc steps, then pdac_step,
    equal pdac_step throughout."""
    case, _, _ = load("explicit_mixed_2d_variants")
    a, b = solver(case, 1), solver(case, 1)
    rc, _, _ = noalloc(a, 1)
    assert rc == 0, a.last_error()
    a.step(1)
    b.step(2)
    compare(a, b, "noalloc then step")
    a.close()
    b.close()


def test_noalloc_refuses_a_mesh_beyond_the_array():
    from pdac_amd.cases import eruption_column
    from pdac_amd.gpu import gpu_solver
    case = eruption_column(nx=600, nz=600, dr0=2.0, dr1=2.0, nsteps=1, dt=1.0e-4)
    case.time_integration = 0
    g = gpu_solver(case, device=0)
    rc, _, _ = noalloc(g, 1)
    assert rc == abi.ST_BAD_ARGUMENT
    assert "PDAC_NOALLOC_BYTES" in g.last_error()
    g.close()
