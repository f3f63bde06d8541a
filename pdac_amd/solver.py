"""Host-side driver of the PDAC hot path through its C-ABI.

:class:`Solver` mirrors the reference time loop ``main_prog``
(prog.f:143-329): it owns one context of a library that implements the
``include/pdacgpu.h`` entry points, uploads a :class:`~pdac_amd.deck.Case`,
and calls ``boundary``, ``fieldn`` and, per Runge-Kutta stage, ``tilde``,
``iter``, ``ygas``, ``htilde`` and ``ftem`` in the reference's order.

The class is library-agnostic (it is given a ``ctypes.CDLL`` and the symbol
prefix) so that the same driver runs the HIP library and any other library
that exports the same ABI.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import abi
from .atmosphere import atmosphere_profile


class PdacError(RuntimeError):
    def __init__(self, where, code, msg=""):
        super().__init__(f"{where} returned {code}{': ' + msg if msg else ''}")
        self.code = code


def bind(lib: ct.CDLL, prefix: str):
    """Declare argument types of the C-ABI entry points on ``lib``."""
    P = ct.c_void_p
    dp, ip = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32)
    sig = {
        "destroy": [P], "set_grid": [P, dp, dp, dp, ip, dp, dp],
        "set_params": [P, ct.POINTER(abi.Params)],
        "put": [P, ct.c_int, dp, ct.c_int64], "get": [P, ct.c_int, dp, ct.c_int64],
        "field_ncomp": [P, ct.c_int], "cnvert": [P], "boundary": [P], "fieldn": [P], "sgsg": [P],
        "begin_step": [P], "tilde": [P, ct.c_double],
        "iter": [P, ct.c_double, ip, ip], "ygas": [P, ct.c_double],
        "htilde": [P, ct.c_double], "ftem": [P, ct.c_double, ip],
        "step": [P, ct.c_double, ct.c_int32, ct.c_int32, ip, ct.POINTER(ct.c_int64)],
        "initialize_fieldn": [P, ct.c_double],
        "explicit_iter": [P, ct.c_double, ip, ct.POINTER(ct.c_double)],
    }
    sig["set_ib"] = [P, ip, ip, dp, dp, ip, dp]
    # device staging of output and restart files (the GPU library only)
    sig["snapshot_size"] = [P, ct.c_int, ct.POINTER(ct.c_int64)]
    sig["snapshot_async"] = [P, ct.c_int]
    sig["snapshot_fetch"] = [P, ct.c_int, ct.c_void_p, ct.c_int64]
    out = {}
    for name, args in sig.items():
        fn = getattr(lib, prefix + name, None)
        if fn is None and name in ("set_ib", "snapshot_size", "snapshot_async", "snapshot_fetch"):
            continue  # a library without immersed boundaries or snapshots
        fn.argtypes = args
        fn.restype = ct.c_int
        out[name] = fn
    return out


def ib_arrays(case):
    """The immersed-boundary arguments of set_ib for a case with
    ``immb = 1`` (pdac_amd.immersed): point counts, the integer and
    real planes of the forcing points, the grid coordinates, the face masks
    and 1/vf."""
    ib = case.immersed()
    ts, _ = case._topo()
    m = ts.mesh
    nfp = np.array([len(f) for f in ib.fp], dtype=np.int32)
    tot = int(nfp.sum())
    fpi = np.empty((7, tot), dtype=np.int32)
    fpd = np.empty((3, tot), dtype=np.float64)
    o = 0
    for f in ib.fp:
        k = len(f)
        for r, col in enumerate((f.i, f.j, f.k, f.int_, f.di, f.dj, f.dk)):
            fpi[r, o:o + k] = col
        for r, col in enumerate((f.nsx, f.nsy, f.nsz)):
            fpd[r, o:o + k] = col
        o += k
    coords = np.concatenate([m.x, m.xb, m.y, m.yb, m.z, m.zb]).astype(np.float64)
    bmask = np.ascontiguousarray(ib.bmask, dtype=np.int32)
    ivf = np.ascontiguousarray(ib.ivf, dtype=np.float64)
    return nfp, np.ascontiguousarray(fpi), np.ascontiguousarray(fpd), coords, bmask, ivf


class Solver:
    """One PDAC domain driven through a C-ABI library.

    ``create`` is a callable ``create(byref(ctx_ptr), byref(dims))`` that
    allocates the context (the device entry point takes an extra device
    index, the caller binds it)."""

    def __init__(self, lib, prefix, create, case, precision=abi.PREC_FP64,
                 iter_order=abi.ORDER_REDBLACK, p_atm=None, t_atm=None):
        self.lib = lib
        self.prefix = prefix
        self.fn = bind(lib, prefix)
        self.case = case
        self.dims = abi.dims_from_case(case, precision)
        self.params = abi.params_from_case(case, iter_order)
        self.ctx = ct.c_void_p()
        rc = create(ct.byref(self.ctx), ct.byref(self.dims))
        if rc != 0:
            raise PdacError(prefix + "create", rc, self.last_error())
        self.n = case.ntot
        dx, dy, dz = case.cell_sizes()
        self._dx = np.ascontiguousarray(dx, dtype=np.float64)
        self._dy = np.ascontiguousarray(dy if case.three_d else np.zeros(1), dtype=np.float64)
        self._dz = np.ascontiguousarray(dz, dtype=np.float64)
        self._flag = np.ascontiguousarray(case.flags(), dtype=np.int32)
        if p_atm is None or t_atm is None:
            p_atm, t_atm = atmosphere_profile(case)
        self._patm = np.ascontiguousarray(p_atm, dtype=np.float64)
        self._tatm = np.ascontiguousarray(t_atm, dtype=np.float64)
        self._check("set_grid", self.fn["set_grid"](
            self.ctx, abi.dptr(self._dx), abi.dptr(self._dy), abi.dptr(self._dz),
            abi.iptr(self._flag), abi.dptr(self._patm), abi.dptr(self._tatm)))
        self._check("set_params", self.fn["set_params"](self.ctx, ct.byref(self.params)))
        if getattr(case, "immb", 0):
            if "set_ib" not in self.fn:
                raise PdacError(prefix + "set_ib", abi.ST_BAD_ARGUMENT, "library without immersed boundaries")
            self._ib = ib_arrays(case)
            nfp, fpi, fpd, coords, bmask, ivf = self._ib
            self._check("set_ib", self.fn["set_ib"](
                self.ctx, abi.iptr(nfp), abi.iptr(fpi), abi.dptr(fpd), abi.dptr(coords),
                abi.iptr(bmask), abi.dptr(ivf)))
        self.time = case.time
        self.nit_last = 0
        self.nit_total = 0
        self._queued = set()  # snapshot kinds queued and not fetched
        if case.fields:
            self.load_state(case.fields)

    # ------------------------------------------------------------------
    def last_error(self):
        f = getattr(self.lib, "pdac_last_error", None)
        if f is None:
            return ""
        f.restype = ct.c_char_p
        msg = f()
        return msg.decode() if msg else ""

    def _check(self, where, rc):
        if rc != 0:
            raise PdacError(self.prefix + where, rc, self.last_error())
        return rc

    def close(self):
        if self.ctx:
            self.fn["destroy"](self.ctx)
            self.ctx = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def ncomp(self, name):
        return self.fn["field_ncomp"](self.ctx, abi.FIELD_ID[name])

    def put(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1))
        self._check("put(" + name + ")", self.fn["put"](self.ctx, abi.FIELD_ID[name], abi.dptr(a), a.size))

    def get(self, name):
        nc = self.ncomp(name)
        a = np.empty(nc * self.n, dtype=np.float64)
        self._check("get(" + name + ")", self.fn["get"](self.ctx, abi.FIELD_ID[name], abi.dptr(a), a.size))
        return a.reshape(nc, self.n) if nc > 1 or name in (
            "rlk", "ts", "sies", "us", "vs", "ws", "ygc", "xgc", "ck", "rlkn",
            "siesn", "rusn", "rvsn", "rwsn", "rgpgcn", "rus", "rvs", "rws", "rhs",
            "rsfe", "rsfn", "rsft", "appu", "appv", "appw") else a

    def load_state(self, fields):
        for name in abi.STATE_FIELDS:
            if name in fields:
                self.put(name, fields[name])
        self._check("cnvert", self.fn["cnvert"](self.ctx))

    def state(self, names=None):
        names = names or [f for f in abi.STATE_FIELDS
                          if self.case.three_d or f not in ("vg", "vs")]
        return {nm: self.get(nm) for nm in names}

    # ------------------------------------------------------------------
    def boundary(self):
        self._check("boundary", self.fn["boundary"](self.ctx))

    def fieldn(self):
        self._check("fieldn", self.fn["fieldn"](self.ctx))

    def sgsg(self):
        self._check("sgsg", self.fn["sgsg"](self.ctx))

    # -- output and restart files (outp, taperd) ------------------------------
    def snapshot_async(self, kind):
        """Pack the arrays of an output (abi.SNAP_OUTPUT) or restart
        (abi.SNAP_RESTART) file on the device and queue their download;
        returns at once."""
        self._check("snapshot_async", self.fn["snapshot_async"](self.ctx, kind))
        self._queued.add(kind)

    def snapshot_fetch(self, kind):
        """The arrays queued by snapshot_async (waits for their download):
        float32 for an output file, float64 for a restart file, one row per
        array of the file."""
        nb = ct.c_int64()
        self._check("snapshot_size", self.fn["snapshot_size"](self.ctx, kind, ct.byref(nb)))
        dtype = np.float32 if kind == abi.SNAP_OUTPUT else np.float64
        out = np.empty(nb.value // np.dtype(dtype).itemsize, dtype=dtype)
        self._check("snapshot_fetch", self.fn["snapshot_fetch"](self.ctx, kind, out.ctypes.data, nb.value))
        self._queued.discard(kind)
        return out.reshape(-1, self.case.ntot)

    def write_output(self, path):
        """The unformatted output file of outp (outp.f:194-384) for the
        current state and time, from a snapshot queued by
        ``snapshot_async(abi.SNAP_OUTPUT)`` (queued here if none is)."""
        from .outputs import write_output
        if abi.SNAP_OUTPUT not in self._queued:
            self.snapshot_async(abi.SNAP_OUTPUT)
        write_output(path, self.time, self.snapshot_fetch(abi.SNAP_OUTPUT))

    def write_restart(self, path, nfil=0):
        """pdac.res as taperd writes it (io_restart.f:31-70) for the current
        state and time, from a restart snapshot (queued here if none is)."""
        from .outputs import write_restart_arrays
        if abi.SNAP_RESTART not in self._queued:
            self.snapshot_async(abi.SNAP_RESTART)
        write_restart_arrays(path, self.case, self.time, nfil, self.snapshot_fetch(abi.SNAP_RESTART))

    def begin_step(self):
        self._check("begin_step", self.fn["begin_step"](self.ctx))

    def tilde(self, dt):
        self._check("tilde", self.fn["tilde"](self.ctx, dt))

    def iter(self, dt):
        nit, conv = ct.c_int32(), ct.c_int32()
        rc = self.fn["iter"](self.ctx, dt, ct.byref(nit), ct.byref(conv))
        self._check("iter", rc)
        return nit.value, bool(conv.value)

    def ygas(self, dt):
        self._check("ygas", self.fn["ygas"](self.ctx, dt))

    def htilde(self, dt):
        self._check("htilde", self.fn["htilde"](self.ctx, dt))

    def ftem(self, dt):
        fail = ct.c_int32()
        self._check("ftem", self.fn["ftem"](self.ctx, dt, ct.byref(fail)))
        return fail.value

    def initialize_fieldn(self, dt):
        self._check("initialize_fieldn", self.fn["initialize_fieldn"](self.ctx, dt))

    def explicit_iter(self, dt):
        """One explicit Runge-Kutta update (explicit_solver.f:996); returns
        the equation-of-state failure count and the largest Courant number."""
        fail, cfl = ct.c_int32(), ct.c_double()
        self._check("explicit_iter", self.fn["explicit_iter"](self.ctx, dt, ct.byref(fail), ct.byref(cfl)))
        return fail.value, cfl.value

    def step(self, nsteps=1, dt=None, rungekut=None):
        """Advance ``nsteps`` steps in one library call (explicit ones when the
        case sets ``time_integration = 1``)."""
        dt = self.case.dt if dt is None else dt
        rk = self.case.rungekut if rungekut is None else rungekut
        nl, nt = ct.c_int32(), ct.c_int64()
        self._check("step", self.fn["step"](self.ctx, dt, rk, nsteps, ct.byref(nl), ct.byref(nt)))
        self.nit_last = nl.value
        self.nit_total += nt.value
        for _ in range(nsteps):
            self.time = self.time + dt
        return nl.value

    def step_by_parts(self, dt=None, rungekut=None):
        """One step as separate subroutine calls, in main_prog's order."""
        dt = self.case.dt if dt is None else dt
        rk_n = self.case.rungekut if rungekut is None else rungekut
        if self.case.time_integration == 1:
            self.boundary()
            self.initialize_fieldn(dt)
            self.sgsg()
            for rk in range(1, rk_n + 1):
                dts = dt / (rk_n + 1 - rk)
                self.explicit_iter(dts)
                self.ygas(dts)
            self.time = self.time + dt
            return 0
        self.boundary()
        self.fieldn()
        self.sgsg()
        self.begin_step()
        nit = 0
        for rk in range(1, rk_n + 1):
            dts = dt / (rk_n + 1 - rk)
            self.tilde(dts)
            nit, _ = self.iter(dts)
            self.ygas(dts)
            self.htilde(dts)
            self.ftem(dts)
        self.time = self.time + dt
        return nit
