"""Reader and writer for the PDAC binary restart file ``pdac.res``.

The layout follows the reference writer ``tapewr`` (io_restart.f:31-124) and
reader ``taperd`` (io_restart.f:126-339): Fortran unformatted sequential
records, each framed by 4-byte little-endian length markers.

    record 1: time (f8), nx, ny, nz, nsolid, ngas, nfil (i4)
    record 2: gas_type(ngas) (i4)
    then one record of ntot float64 values per field, in this order:
    p, ug, [vg], wg, sieg, rlk(is), (us, [vs], ws)(is), sies(is), ygc(ig),
    rgp, rog, tg, ts(is)

Multi-component fields are kept as Fortran column-major ``(ntot, ncomp)``
arrays, i.e. ``numpy`` arrays of shape ``(ncomp, ntot)`` here, one contiguous
plane per component, which is also the device layout.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np


def _write_record(fh, payload: bytes) -> None:
    n = len(payload)
    fh.write(struct.pack("<i", n))
    fh.write(payload)
    fh.write(struct.pack("<i", n))


def _read_record(fh) -> bytes:
    head = fh.read(4)
    if len(head) != 4:
        raise EOFError("truncated restart file")
    (n,) = struct.unpack("<i", head)
    payload = fh.read(n)
    (n2,) = struct.unpack("<i", fh.read(4))
    if n2 != n or len(payload) != n:
        raise ValueError("corrupt Fortran record markers in restart file")
    return payload


@dataclass
class RestartState:
    """Primary fields of one PDAC state, as stored in ``pdac.res``."""

    time: float
    nx: int
    ny: int
    nz: int
    nsolid: int
    ngas: int
    nfil: int
    gas_type: List[int]
    three_d: bool
    fields: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def ntot(self) -> int:
        return self.nx * self.ny * self.nz

    def order(self):
        """Field names and component counts in file order."""
        s, g = self.nsolid, self.ngas
        seq = [("p", 1), ("ug", 1)]
        if self.three_d:
            seq.append(("vg", 1))
        seq += [("wg", 1), ("sieg", 1), ("rlk", s)]
        # particle velocities are interleaved per class in the file
        seq.append(("__svel__", s))
        seq += [("sies", s), ("ygc", g), ("rgp", 1), ("rog", 1), ("tg", 1),
                ("ts", s)]
        return seq


def write_restart(path: str, st: RestartState) -> None:
    n = st.ntot
    with open(path, "wb") as fh:
        _write_record(fh, struct.pack("<d6i", st.time, st.nx, st.ny, st.nz,
                                      st.nsolid, st.ngas, st.nfil))
        _write_record(fh, np.asarray(st.gas_type, dtype="<i4").tobytes())
        for name, nc in st.order():
            if name == "__svel__":
                comps = ["us", "vs", "ws"] if st.three_d else ["us", "ws"]
                for is_ in range(st.nsolid):
                    for c in comps:
                        a = np.asarray(st.fields[c], dtype="<f8").reshape(-1, n)
                        _write_record(fh, a[is_].tobytes())
                continue
            a = np.asarray(st.fields[name], dtype="<f8").reshape(-1, n)
            if a.shape[0] != nc:
                raise ValueError(f"field {name}: expected {nc} components")
            for c in range(nc):
                _write_record(fh, a[c].tobytes())


def read_restart(path: str, three_d: bool) -> RestartState:
    with open(path, "rb") as fh:
        rec = _read_record(fh)
        time, nx, ny, nz, ns, ng, nfil = struct.unpack("<d6i", rec)
        gas_type = list(np.frombuffer(_read_record(fh), dtype="<i4"))
        st = RestartState(time, nx, ny, nz, ns, ng, nfil, gas_type, three_d)
        n = st.ntot
        for name, nc in st.order():
            if name == "__svel__":
                comps = ["us", "vs", "ws"] if three_d else ["us", "ws"]
                bufs = {c: np.empty((ns, n)) for c in comps}
                for is_ in range(ns):
                    for c in comps:
                        bufs[c][is_] = np.frombuffer(_read_record(fh), dtype="<f8")
                st.fields.update(bufs)
                continue
            a = np.empty((nc, n))
            for c in range(nc):
                a[c] = np.frombuffer(_read_record(fh), dtype="<f8")
            st.fields[name] = a if nc > 1 or name in ("rlk", "sies", "ygc", "ts") else a[0]
        return st
