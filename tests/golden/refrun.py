"""Run the reference solver (oracle/_ref/pdac.x, built here from the
reference's sources by oracle/build_ref.sh) on a Case and read back its
restart file.  Used only in the build container to write golden fixtures."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import numpy as np

from pdac_amd.restart import read_restart

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
PDACX = os.path.join(REPO, "oracle", "_ref", "pdac.x")


def available():
    return os.path.exists(PDACX)


def read_forcing_points(wd):
    """The forcing points the reference reports (set_forcing with lpr > 0,
    immb.f:478-503): per component an array of (i, j, k, int) and one of
    the no-slip points as printed (F12.3)."""
    out = {}
    for a in "xyz":
        path = os.path.join(wd, f"fpt{a}.dat")
        if not os.path.exists(path):
            continue
        rows = [l.split() for l in open(path) if l.strip()]
        out[f"fpt{a}_int"] = np.array([[int(v) for v in r[1:5]] for r in rows], dtype=np.int32)
        out[f"fpt{a}_nsl"] = np.array([[float(v) for v in r[5:8]] for r in rows])
    return out


def run_reference(case, workdir=None, keep=False, timeout=3600, extras=None):
    wd = workdir or tempfile.mkdtemp(prefix="pdacref_")
    os.makedirs(wd, exist_ok=True)
    case.write_deck(os.path.join(wd, "pdac.dat"))
    case.write_restart(os.path.join(wd, "pdac.res"))
    r = subprocess.run([PDACX], cwd=wd, capture_output=True, text=True, timeout=timeout)
    log = open(os.path.join(wd, "pdac.log")).read() if os.path.exists(os.path.join(wd, "pdac.log")) else ""
    if r.returncode != 0 or "writing to restart file" not in log:
        raise RuntimeError(f"reference run failed in {wd}: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    st = read_restart(os.path.join(wd, "pdac.res"), case.three_d)
    nits = [int(l.split("nit:")[1]) for l in log.splitlines() if "nit:" in l]
    if extras is not None:
        extras.update(read_forcing_points(wd))
        # the output files (outp) and the restart file (taperd) as written
        for name in sorted(os.listdir(wd)):
            if name.startswith("output.") or name == "pdac.res":
                key = "restart" if name == "pdac.res" else name.replace(".", "_")
                extras[key] = np.frombuffer(open(os.path.join(wd, name), "rb").read(), dtype=np.uint8)
    if not keep and workdir is None:
        shutil.rmtree(wd, ignore_errors=True)
    return st, nits
