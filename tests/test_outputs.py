"""The writers of the reference's output and restart files
(pdac_amd/outputs.py) against the files the reference itself wrote for two
fixtures: output.0000 (outp at the first step, outp.f:194-384 through
write_array, io_parallel.f:18-65) rebuilt from the initial state as the
checker converts it (cnvert) and the first step's boundary call leaves it,
and pdac.res (taperd, io_restart.f:31-70)
rebuilt from the reference's final state, byte for byte."""
import os

import numpy as np
import pytest

from oracle.binding import cpu_solver
from pdac_amd import abi
from pdac_amd.outputs import output_names, read_output, write_output, write_restart_arrays
from pdac_amd.restart import RestartState, read_restart
from tests.golden.make_golden import load, load_extras

OUTPUT_FIXTURES = ["explicit_output_2d", "output_3d"]


def output_rows(case, st):
    """The arrays of an output file from a state (host restatement of outp's
    list: eps = rlk * inrl, REAL(x, 4))."""
    n = case.ntot
    rows = [st["p"], st["ug"]] + ([st["vg"]] if case.three_d else []) + [st["wg"], st["tg"]]
    rows += list(np.asarray(st["xgc"]).reshape(-1, n))
    for s in range(case.nsolid):
        inrl = 1.0 / case.density[s]
        rows.append(np.asarray(st["rlk"]).reshape(-1, n)[s] * inrl)
        for c in (["us", "vs", "ws"] if case.three_d else ["us", "ws"]):
            rows.append(np.asarray(st[c]).reshape(-1, n)[s])
        rows.append(np.asarray(st["ts"]).reshape(-1, n)[s])
    return np.stack([np.asarray(r, dtype=np.float64).reshape(-1) for r in rows]).astype(np.float32)


@pytest.mark.parametrize("name", OUTPUT_FIXTURES)
def test_first_output_file_matches_the_reference(name, tmp_path):
    case, _, _ = load(name)
    ref = bytes(load_extras(name)["output_0000"])
    c = cpu_solver(case, iter_order=abi.ORDER_LEXICOGRAPHIC)
    c.boundary()  # main_prog writes the first file after the step's boundary call (prog.f:160-165)
    st = c.state(["p", "ug", "vg", "wg", "tg", "xgc", "rlk", "us", "vs", "ws", "ts"])
    c.close()
    rows = output_rows(case, st)
    assert rows.shape[0] == len(output_names(case))
    path = tmp_path / "output.0000"
    write_output(str(path), case.time, rows)
    assert path.read_bytes() == ref
    t, back = read_output(str(path))
    assert t == 0.0 and np.array_equal(back, rows)


@pytest.mark.parametrize("name", OUTPUT_FIXTURES)
def test_restart_file_matches_the_reference(name, tmp_path):
    case, out, _ = load(name)
    ref_path = tmp_path / "ref.res"
    ref_path.write_bytes(bytes(load_extras(name)["restart"]))
    ref = read_restart(str(ref_path), case.three_d)
    st = RestartState(ref.time, ref.nx, ref.ny, ref.nz, ref.nsolid, ref.ngas, ref.nfil,
                      ref.gas_type, case.three_d)
    n = st.ntot
    rows = []
    comps = ["us", "vs", "ws"] if case.three_d else ["us", "ws"]
    for nm, nc in st.order():
        if nm == "__svel__":
            for s in range(case.nsolid):
                rows += [np.asarray(out[c]).reshape(-1, n)[s] for c in comps]
            continue
        rows += list(np.asarray(out[nm]).reshape(-1, n)[:nc])
    path = tmp_path / "pdac.res"
    write_restart_arrays(str(path), case, ref.time, ref.nfil, np.stack(rows))
    assert path.read_bytes() == ref_path.read_bytes()
    assert ref.nfil == 2  # output.0000 and output.0001 before the dump
