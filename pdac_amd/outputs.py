"""Writers of the reference's output and restart files from the arrays of a
device snapshot (pdac_snapshot_async / pdac_snapshot_fetch).

- ``output.nnnn``, unformatted (outp, outp.f:194-384 with
  formatted_output = .FALSE., the default of input.f:232): one Fortran
  sequential record holding the time as a 4-byte real, then one record per
  array of ntot 4-byte reals (write_array, io_parallel.f:18-65): p, ug, [vg],
  wg, tg, xgc of each species, then per class rlk/rl, us, [vs], ws, ts.
- ``pdac.res`` (taperd, io_restart.f:31-70): a record with time (8 bytes)
  and nx, ny, nz, nsolid, ngas, nfil (4 bytes each), a record with the gas
  types, then one record per array of ntot 8-byte reals in the order of
  pdac_amd.restart.RestartState.order.

Sequential records carry their byte count as a 4-byte integer before and
after the payload (the Fortran compilers' default record markers)."""
from __future__ import annotations

import struct

import numpy as np

from .restart import RestartState, write_restart


def _record(fh, payload: bytes) -> None:
    n = struct.pack("<i", len(payload))
    fh.write(n)
    fh.write(payload)
    fh.write(n)


def write_output(path: str, time: float, arrays: np.ndarray) -> None:
    """arrays: float32, one row of ntot values per array of the file."""
    a = np.ascontiguousarray(arrays, dtype="<f4")
    with open(path, "wb") as fh:
        _record(fh, np.float32(time).astype("<f4").tobytes())
        for row in a:
            _record(fh, row.tobytes())


def read_output(path: str):
    """(time, arrays) of an unformatted output file: the time record and
    one row per following record."""
    rows = []
    with open(path, "rb") as fh:
        data = fh.read()
    off, recs = 0, []
    while off < len(data):
        (n,) = struct.unpack_from("<i", data, off)
        recs.append(data[off + 4:off + 4 + n])
        (n2,) = struct.unpack_from("<i", data, off + 4 + n)
        if n2 != n:
            raise ValueError("corrupt record markers in output file")
        off += 8 + n
    time = float(np.frombuffer(recs[0], dtype="<f4")[0])
    rows = [np.frombuffer(r, dtype="<f4") for r in recs[1:]]
    return time, np.stack(rows) if rows else np.zeros((0, 0), dtype=np.float32)


def output_names(case) -> list:
    """Names of the arrays of an output file, in file order."""
    names = ["p", "ug"] + (["vg"] if case.three_d else []) + ["wg", "tg"]
    names += [f"xgc{g + 1}" for g in range(len(case.gas_type))]
    for s in range(case.nsolid):
        names += [f"eps{s + 1}", f"us{s + 1}"] + ([f"vs{s + 1}"] if case.three_d else [])
        names += [f"ws{s + 1}", f"ts{s + 1}"]
    return names


def write_restart_arrays(path: str, case, time: float, nfil: int, arrays: np.ndarray) -> None:
    """pdac.res from the rows of a restart snapshot (float64, file order)."""
    st = RestartState(time, case.nx, case.ny if case.three_d else 1, case.nz, case.nsolid,
                      len(case.gas_type), nfil, list(case.gas_type), case.three_d)
    n = st.ntot
    rows = iter(np.asarray(arrays, dtype=np.float64))
    comps = ["us", "vs", "ws"] if case.three_d else ["us", "ws"]
    fields = {}
    for name, nc in st.order():
        if name == "__svel__":
            sv = {c: np.empty((case.nsolid, n)) for c in comps}
            for s in range(case.nsolid):
                for c in comps:
                    sv[c][s] = next(rows)
            fields.update({c: v.reshape(-1) for c, v in sv.items()})
            continue
        fields[name] = np.stack([next(rows) for _ in range(nc)]).reshape(-1)
    st.fields = fields
    write_restart(path, st)
