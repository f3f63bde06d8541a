"""Run the reference solver (oracle/_ref/pdac.x, built here from the
reference's sources by oracle/build_ref.sh) on a Case and read back its
restart file.  Used only in the build container to write golden fixtures."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

from pdac_amd.restart import read_restart

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
PDACX = os.path.join(REPO, "oracle", "_ref", "pdac.x")


def available():
    return os.path.exists(PDACX)


def run_reference(case, workdir=None, keep=False, timeout=3600):
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
    if not keep and workdir is None:
        shutil.rmtree(wd, ignore_errors=True)
    return st, nits
