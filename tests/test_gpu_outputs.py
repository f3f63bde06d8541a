"""Output and restart files from device snapshots (pdac_snapshot_*):
against the files the reference wrote for two fixtures (output.0000 after
the first step's boundary call, output.0001 and pdac.res after the last
step), and the snapshot's stream order: a snapshot fetched after further
steps holds the state at the time it was queued."""
import numpy as np
import pytest

from pdac_amd import abi
from pdac_amd.outputs import read_output
from pdac_amd.restart import read_restart
from tests.golden.make_golden import load, load_extras

pytestmark = pytest.mark.gpu

# explicit steps agree with the reference to 1e-12 (float32 output: a few
# values may round to the neighbouring float); semi-implicit ones within the
# converged residual tolerance of the two sweep orders
TOL = {"explicit_output_2d": (1e-6, 1e-11), "output_3d": (1e-5, 1e-5)}


def rel_rows(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.max(np.abs(b), axis=1), 1e-300)
    return np.max(np.abs(a - b), axis=1) / scale


def ref_file(tmp_path, name, key):
    p = tmp_path / f"ref_{key}"
    p.write_bytes(bytes(load_extras(name)[key]))
    return str(p)


@pytest.mark.parametrize("name", sorted(TOL))
def test_files_match_the_reference(name, tmp_path):
    from pdac_amd.gpu import gpu_solver
    print(f"[outputs] {name}", flush=True)
    case, _, _ = load(name)
    tol_out, tol_res = TOL[name]
    # the first file: after the first step's boundary call (prog.f:160-165)
    g = gpu_solver(case, device=0)
    g.boundary()
    g.write_output(str(tmp_path / "output.0000"))
    g.close()
    t0, got0 = read_output(str(tmp_path / "output.0000"))
    r0, want0 = read_output(ref_file(tmp_path, name, "output_0000"))
    assert t0 == r0 and got0.shape == want0.shape
    assert np.all(rel_rows(got0, want0) <= 1e-6), rel_rows(got0, want0)
    # the last output and the restart file
    g = gpu_solver(case, device=0)
    g.step(case.nsteps)
    g.snapshot_async(abi.SNAP_OUTPUT)
    g.snapshot_async(abi.SNAP_RESTART)
    g.write_output(str(tmp_path / "output.0001"))
    g.write_restart(str(tmp_path / "pdac.res"), nfil=2)
    g.close()
    t1, got1 = read_output(str(tmp_path / "output.0001"))
    r1, want1 = read_output(ref_file(tmp_path, name, "output_0001"))
    assert t1 == r1 and got1.shape == want1.shape
    assert np.all(rel_rows(got1, want1) <= tol_out), rel_rows(got1, want1)
    got = read_restart(str(tmp_path / "pdac.res"), case.three_d)
    want = read_restart(ref_file(tmp_path, name, "restart"), case.three_d)
    assert (got.time, got.nx, got.ny, got.nz, got.nsolid, got.ngas, got.nfil, list(got.gas_type)) == \
        (want.time, want.nx, want.ny, want.nz, want.nsolid, want.ngas, want.nfil, list(want.gas_type))
    for k, v in want.fields.items():
        a, b = np.asarray(got.fields[k]).reshape(-1, case.ntot), np.asarray(v).reshape(-1, case.ntot)
        assert np.all(rel_rows(a, b) <= tol_res), (k, rel_rows(a, b))


def test_snapshot_holds_the_state_when_queued():
    """Steps queued after a snapshot do not reach it; the fetched arrays
    are the state at the snapshot, bit for bit, in the restart file's
    order."""
    from pdac_amd.cases import mixed_box
    from pdac_amd.gpu import gpu_solver
    case = mixed_box(nsteps=5, bottom=3, itc=1, west=2, east=4, top=2)
    case.time_integration = 1
    g = gpu_solver(case, device=0)
    g.step(2)
    st = g.state()
    g.snapshot_async(abi.SNAP_RESTART)
    g.snapshot_async(abi.SNAP_OUTPUT)
    g.step(3)
    rows = g.snapshot_fetch(abi.SNAP_RESTART)
    out = g.snapshot_fetch(abi.SNAP_OUTPUT)
    later = g.state()
    g.close()
    n, s = case.ntot, case.nsolid
    want = [st["p"], st["ug"], st["wg"], st["sieg"]] + list(st["rlk"].reshape(s, n))
    for k in range(s):
        want += [st["us"].reshape(s, n)[k], st["ws"].reshape(s, n)[k]]
    want += list(st["sies"].reshape(s, n)) + list(st["ygc"].reshape(-1, n))
    want += [st["rgp"], st["rog"], st["tg"]] + list(st["ts"].reshape(s, n))
    assert np.array_equal(rows, np.stack([np.asarray(w).reshape(-1) for w in want]))
    assert np.array_equal(out[0], st["p"].astype(np.float32))
    assert not np.array_equal(later["p"], st["p"])
    with pytest.raises(Exception, match="no snapshot"):
        g2 = gpu_solver(case, device=0)
        try:
            g2.snapshot_fetch(abi.SNAP_OUTPUT)
        finally:
            g2.close()
