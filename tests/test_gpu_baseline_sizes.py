"""The HIP path at the full sizes of the measured workloads, against the CPU
checker for the first steps from the initial state (the checker's cell loops
run on OpenMP threads, so a step at these sizes takes seconds) and through
launch-shape invariance where the checker would take minutes:

- the 2D cylindrical eruption column, 512 x 1024 stretched cells (1 m at the
  axis to 20 m), sustained inlet-cell vent, free in/outflow lid,
  semi-implicit: ten steps in the red-black order, the same sweep counts,
  fields to the semi-implicit tolerance;
- the 3D summit cone with immersed boundaries, 256 x 256 x 192 cells of 10 m,
  three particle classes: two steps in the red-black order against the
  checker, and five steps in several sweep launch shapes (flag words per
  thread, list visits between checks, graph replay, the persistent kernel),
  bit-identical to the default shape;
- the explicit eruption column, 256 x 512 uniform 2 m cells: ten steps in the
  reference's cell order, fields to 1e-11 (the explicit step's 1e-12 per
  step, over ten steps)."""
import os

import numpy as np
import pytest

from oracle.binding import cpu_solver
from pdac_amd import abi

pytestmark = pytest.mark.gpu

SHAPES = ["launches", "launches:12", "launches:24:flush", "launches:7:late", "graph", "persistent"]


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    d = np.max(np.abs(a - b)) if b.size else 0.0
    return d / scale if scale > 0.0 else d


def gpu(case, mode="launches", **kw):
    """GPU solver with the sweeps launched as 'mode' says: "launches" (one
    launch per colour phase), "graph" (batches of those replayed as HIP
    graphs), "persistent" (one cooperative kernel); "launches:N" fixes the
    flag words per thread of the 3D sweep kernel, ":late" reads the sweep
    counters before the first checks, ":flush" visits the block's list as soon
    as it holds more than one cell."""
    from pdac_amd.gpu import gpu_solver
    mode, _, rest = mode.partition(":")
    words, _, order = rest.partition(":")
    env = {"PDAC_ITER_MODE": mode}
    if words:
        env["PDAC_SWEEP_WORDS"] = words
    if order == "flush":
        env["PDAC_SWEEP_FLUSH"] = "1"
    elif order:
        env["PDAC_SWEEP_EARLY"] = "0" if order == "late" else "1"
    os.environ.update(env)
    try:
        return gpu_solver(case, device=0, **kw)
    finally:
        for k in env:
            os.environ.pop(k, None)


def full_cone():
    from pdac_amd.cases import summit_cone_3d
    return summit_cone_3d(nx=256, ny=256, nz=192, nsteps=1)


def test_eruption_column_512x1024_against_checker():
    from pdac_amd.cases import eruption_column
    case = eruption_column(nx=512, nz=1024, nsteps=10, vent="inlet", top=4)
    print(f"[column 512x1024] dt {case.dt:.4g}", flush=True)
    g = gpu(case)
    c = cpu_solver(case, iter_order=abi.ORDER_REDBLACK)
    ng, nc = [], []
    for step in range(case.nsteps):
        ng.append(g.step(1))
        nc.append(c.step(1))
        print(f"[column 512x1024] step {step}: sweeps gpu {ng[-1]} checker {nc[-1]}", flush=True)
    sg, sc = g.state(), c.state()
    g.close()
    c.close()
    assert ng == nc
    assert max(rel(sc[k], case.fields[k]) for k in ("p", "wg")) > 1e-8  # the vent has acted
    bad = {k: rel(sg[k], sc[k]) for k in sc if rel(sg[k], sc[k]) > 1e-10}
    assert not bad, bad


@pytest.mark.timeout(900)
def test_summit_cone_full_size_against_checker():
    case = full_cone()
    print(f"[cone 256x256x192] checker step, dt {case.dt:.4g}", flush=True)
    g = gpu(case)
    ng = [g.step(1) for _ in range(2)]
    sg = g.state()
    g.close()
    c = cpu_solver(case, iter_order=abi.ORDER_REDBLACK)
    nc = [c.step(1) for _ in range(2)]
    sc = c.state()
    c.close()
    print(f"[cone 256x256x192] sweeps gpu {ng} checker {nc}", flush=True)
    assert ng == nc
    bad = {k: rel(sg[k], sc[k]) for k in sc if rel(sg[k], sc[k]) > 1e-10}
    assert not bad, bad


@pytest.mark.timeout(600)
def test_summit_cone_full_size_launch_shapes():
    case = full_cone()
    ref = ref_nits = None
    for mode in SHAPES:
        print(f"[cone 256x256x192 shapes] {mode}", flush=True)
        g = gpu(case, mode)
        nits = [g.step(1) for _ in range(5)]
        st = g.state()
        g.close()
        print(f"[cone 256x256x192 shapes] {mode}: sweeps {nits}", flush=True)
        if ref is None:
            ref, ref_nits = st, nits
            continue
        assert nits == ref_nits, (mode, nits, ref_nits)
        for k, v in ref.items():
            assert np.array_equal(st[k], v), (mode, k)


def test_explicit_eruption_column_256x512_against_checker():
    from pdac_amd.cases import eruption_column
    case = eruption_column(nx=256, nz=512, dr0=2.0, dr1=2.0, nsteps=10, dt=1.0e-4, vent="inlet")
    case.time_integration = 1
    case.rungekut = 1
    print("[explicit 256x512]", flush=True)
    g = gpu(case)
    c = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
    g.step(case.nsteps)
    c.step(case.nsteps)
    sg, sc = g.state(), c.state()
    g.close()
    c.close()
    assert max(rel(sc[k], case.fields[k]) for k in ("p", "wg")) > 1e-8  # the vent has acted
    bad = {k: rel(sg[k], sc[k]) for k in sc if rel(sg[k], sc[k]) > 1e-11}
    assert not bad, bad
