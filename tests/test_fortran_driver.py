"""The Fortran host binding (pdac_amd/fortran/pdac_gpu.f90, iso_c_binding)
and a driver shaped like the reference's time loop (main_prog,
prog.f:143-329; pdac_amd/fortran/pdac_drive.f90), both compiled with the
image's flang by pdac_amd/build.py.

CPU: the BIND(C) types have the C structs' sizes (C_SIZEOF against sizeof
from gcc and ctypes), and the module binds every function the header
declares.  GPU: the driver, reading a domain written here, advances it
through the C-ABI from Fortran, subroutine by subroutine in main_prog's order
and through pdac_step, and its fields equal, bit for bit, those of the same
calls made from Python."""
from __future__ import annotations

import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from pdac_amd import abi
from pdac_amd import build as hipbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pdacgpu.h")
MODULE = os.path.join(ROOT, "pdac_amd", "fortran", "pdac_gpu.f90")


def driver():
    if not os.path.exists(hipbuild.DRIVER):
        pytest.fail("pdac_amd/fortran/pdac_drive.x is not built (python -m pdac_amd.build)")
    return hipbuild.DRIVER


def c_sizes(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "pdacgpu.h"\n'
                   'int main(void) { printf("%zu %zu %zu\\n", sizeof(pdac_dims), sizeof(pdac_params),'
                   ' offsetof(pdac_params, time_integration)); return 0; }\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]


def test_fortran_types_match_the_c_structs(tmp_path):
    out = subprocess.run([driver(), "--sizes"], capture_output=True, text=True, check=True).stdout
    m = re.match(r"pdac_dims (\d+) pdac_params (\d+)", out.strip())
    assert m, out
    fd, fp = int(m.group(1)), int(m.group(2))
    cd, cp, off = c_sizes(tmp_path)
    assert (fd, fp) == (cd, cp) == (ct.sizeof(abi.Dims), ct.sizeof(abi.Params))
    assert off == abi.Params.time_integration.offset


def test_module_binds_every_declared_function():
    text = open(HEADER).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(pdac_\w+)\(", text, re.M))
    bound = set(re.findall(r"name='(pdac_\w+)'", open(MODULE).read()))
    assert declared and declared == bound, (declared - bound, bound - declared)


def test_driver_reports_a_refused_call(tmp_path):
    """A refused call ends the driver as the reference's CALL error does:
    the message of pdac_last_error, status in the output file, exit 1 (the
    dims are checked before any device call)."""
    from tests.golden.make_golden import load
    case, _, _ = load("mixed_2d")
    gets = write_domain(tmp_path / "in.bin", case, 1, 0)
    raw = bytearray(open(tmp_path / "in.bin", "rb").read())
    raw[20:24] = np.asarray([abi.PDAC_MAX_SOLID + 1], dtype="<i4").tobytes()  # dims.nsolid
    open(tmp_path / "in.bin", "wb").write(bytes(raw))
    res = subprocess.run([driver(), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 1
    assert res.stdout.startswith(f"pdac_create: status {abi.ST_BAD_ARGUMENT}: "), res.stdout
    raw = open(tmp_path / "out.bin", "rb").read()
    assert len(raw) == 16 and int(np.frombuffer(raw[:4], dtype="<i4")[0]) == abi.ST_BAD_ARGUMENT
    assert gets


# ----------------------------------------------------------------------
def write_domain(path, case, nsteps, mode, dt=None, rungekut=None, precision=abi.PREC_FP64):
    """The driver's input file: dims, params, grid, immersed boundaries,
    state fields, and the ids of the fields to write back."""
    from pdac_amd.atmosphere import atmosphere_profile
    from pdac_amd.solver import ib_arrays
    d = abi.dims_from_case(case, precision)
    p = abi.params_from_case(case, abi.ORDER_REDBLACK)
    dx, dy, dz = case.cell_sizes()
    p_atm, t_atm = atmosphere_profile(case)
    i32 = lambda v: np.asarray(v, dtype="<i4").tobytes()  # noqa: E731
    f64 = lambda v: np.asarray(v, dtype="<f8").tobytes()  # noqa: E731
    parts = [i32([d.job_type, d.itc, d.nx, d.ny, d.nz, d.nsolid, d.ngas, d.precision]),
             f64([p.alpha, p.alphagrav, p.gravx, p.gravy, p.gravz, p.beta, p.omega, p.delg,
                  p.rlim, p.flim, p.tref, p.void_fraction])]
    for name in ("dk", "rl", "phis", "cps", "kap", "plim", "atm_ygc"):
        parts.append(f64(list(getattr(p, name))))
    parts.append(i32(list(p.gas_type)))
    parts.append(i32([p.muscl, p.lm, p.lv, p.inmax, p.maxout, p.pmodel, p.drag_model, p.icpc,
                      p.ord, p.cstop, p.iter_order, p.time_integration, p.gas_viscosity,
                      p.part_viscosity, p.repulsive_model, p.iturb]))
    parts.append(f64(list(p.mus) + [p.cmut]))
    parts.append(i32([nsteps, case.rungekut if rungekut is None else rungekut, mode]))
    parts.append(f64([case.dt if dt is None else dt, case.time]))
    immb = int(getattr(case, "immb", 0) or 0)
    parts.append(i32([immb]))
    parts += [f64(dx), f64(dy if case.three_d else np.zeros(1)), f64(dz), f64(p_atm), f64(t_atm),
              i32(case.flags())]
    if immb:
        nfp, fpi, fpd, coords, bmask, ivf = ib_arrays(case)
        parts += [i32(nfp), i32(fpi), f64(fpd), f64(coords), i32(bmask), f64(ivf)]
    puts = [nm for nm in abi.STATE_FIELDS if nm in case.fields]
    parts.append(i32([len(puts)]))
    for nm in puts:
        a = np.asarray(case.fields[nm], dtype="<f8").reshape(-1)
        parts += [i32([abi.FIELD_ID[nm]]), np.asarray([a.size], dtype="<i8").tobytes(), a.tobytes()]
    gets = [nm for nm in abi.STATE_FIELDS if case.three_d or nm not in ("vg", "vs")]
    parts.append(i32([len(gets)]))
    parts.append(i32([abi.FIELD_ID[nm] for nm in gets]))
    with open(path, "wb") as f:
        f.write(b"".join(parts))
    return gets


def read_result(path, gets):
    raw = open(path, "rb").read()
    status, nit_last = np.frombuffer(raw[:8], dtype="<i4")
    nit_total = int(np.frombuffer(raw[8:16], dtype="<i8")[0])
    out, o = {}, 16
    for nm in gets:
        fid = int(np.frombuffer(raw[o:o + 4], dtype="<i4")[0])
        cnt = int(np.frombuffer(raw[o + 4:o + 12], dtype="<i8")[0])
        assert fid == abi.FIELD_ID[nm]
        out[nm] = np.frombuffer(raw[o + 12:o + 12 + 8 * cnt], dtype="<f8").copy()
        o += 12 + 8 * cnt
    assert o == len(raw)
    return int(status), int(nit_last), nit_total, out


DRIVER_CASES = [("mixed_2d", 0, 3), ("mixed_2d", 1, 3), ("explicit_mixed_2d", 0, 3),
                ("eruption_2d_vent", 0, 2), ("cone_3d_ib", 0, 2), ("cone_3d_ib", 1, 2)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,mode,nsteps", DRIVER_CASES)
def test_fortran_driver_matches_python_host(tmp_path, name, mode, nsteps):
    from pdac_amd.gpu import gpu_solver
    from tests.golden.make_golden import load
    case, _, _ = load(name)
    gets = write_domain(tmp_path / "in.bin", case, nsteps, mode)
    print(f"[fortran driver] {name} mode {mode} {nsteps} steps", flush=True)
    res = subprocess.run([driver(), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    status, nit_last, nit_total, fort = read_result(tmp_path / "out.bin", gets)
    assert status == 0
    g = gpu_solver(case, device=0)
    if mode == 1:
        g.step(nsteps)
        assert (nit_last, nit_total) == (g.nit_last, g.nit_total)
    else:
        for _ in range(nsteps):
            g.step_by_parts()
    for nm in gets:
        py = np.asarray(g.get(nm)).reshape(-1)
        assert np.array_equal(fort[nm], py), nm
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["explicit_output_2d", "output_3d"])
def test_fortran_driver_writes_the_reference_files(tmp_path, name):
    """The driver writes output.0000 (after the first step's boundary call),
    output.0001 and pdac.res with Fortran unformatted records from device
    snapshots, byte for byte the files the Python host writes for the same
    steps; their layout is the reference's (the Python writers match the
    reference's own files, tests/test_outputs.py)."""
    from pdac_amd.gpu import gpu_solver
    from tests.golden.make_golden import load
    case, _, _ = load(name)
    write_domain(tmp_path / "in.bin", case, case.nsteps, 0)
    fdir = tmp_path / "fortran"
    fdir.mkdir()
    print(f"[fortran driver files] {name}", flush=True)
    res = subprocess.run([driver(), str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(fdir)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    g = gpu_solver(case, device=0)
    g.boundary()
    g.write_output(str(tmp_path / "output.0000"))
    g.close()
    g = gpu_solver(case, device=0)
    for _ in range(case.nsteps):
        g.step_by_parts()
    g.write_output(str(tmp_path / "output.0001"))
    g.write_restart(str(tmp_path / "pdac.res"), nfil=2)
    g.close()
    for f in ("output.0000", "output.0001", "pdac.res"):
        assert (fdir / f).read_bytes() == (tmp_path / f).read_bytes(), f
