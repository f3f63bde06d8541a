"""ctypes mirror of the C-ABI structures declared in include/pdacgpu.h."""
from __future__ import annotations

import ctypes as ct

import numpy as np

PDAC_MAX_SOLID = 4
PDAC_MAX_GAS = 7

PREC_FP64 = 0
PREC_FP32FLUX = 1
ORDER_LEXICOGRAPHIC = 0
ORDER_REDBLACK = 1
ORDER_STAGE = 2

ST_ITER_NOT_CONVERGED = 1
ST_EOS_FAILURE = 2
ST_BAD_ARGUMENT = 3
SNAP_OUTPUT = 0   # PDAC_SNAP_OUTPUT
SNAP_RESTART = 1  # PDAC_SNAP_RESTART

FIELDS = [
    "P", "EP", "RGP", "ROG", "TG", "SIEG", "UG", "VG", "WG",
    "RLK", "TS", "SIES", "US", "VS", "WS", "YGC", "XGC",
    "CG", "CK", "MUG", "KAPG",
    "PN", "EPN", "RGPN", "SIEGN", "RLKN", "SIESN",
    "RUGN", "RVGN", "RWGN", "RUSN", "RVSN", "RWSN", "RGPGCN",
    "RUG", "RVG", "RWG", "RUS", "RVS", "RWS", "APPU", "APPV", "APPW",
    "RHG", "RHS",
    "RGFE", "RGFN", "RGFT", "RSFE", "RSFN", "RSFT",
]
FIELD_ID = {name.lower(): i for i, name in enumerate(FIELDS)}
NFIELDS = len(FIELDS)

# fields of the primary state (restart file) and their enum names
STATE_FIELDS = ["p", "ug", "vg", "wg", "sieg", "rlk", "us", "vs", "ws", "sies",
                "ygc", "rgp", "rog", "tg", "ts"]


class Dims(ct.Structure):
    _fields_ = [("job_type", ct.c_int32), ("itc", ct.c_int32),
                ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
                ("nsolid", ct.c_int32), ("ngas", ct.c_int32),
                ("precision", ct.c_int32)]


class Params(ct.Structure):
    _fields_ = [
        ("alpha", ct.c_double), ("alphagrav", ct.c_double),
        ("gravx", ct.c_double), ("gravy", ct.c_double), ("gravz", ct.c_double),
        ("beta", ct.c_double), ("omega", ct.c_double), ("delg", ct.c_double),
        ("rlim", ct.c_double), ("flim", ct.c_double), ("tref", ct.c_double),
        ("void_fraction", ct.c_double),
        ("dk", ct.c_double * PDAC_MAX_SOLID), ("rl", ct.c_double * PDAC_MAX_SOLID),
        ("phis", ct.c_double * PDAC_MAX_SOLID), ("cps", ct.c_double * PDAC_MAX_SOLID),
        ("kap", ct.c_double * PDAC_MAX_SOLID), ("plim", ct.c_double * PDAC_MAX_SOLID),
        ("atm_ygc", ct.c_double * PDAC_MAX_GAS),
        ("gas_type", ct.c_int32 * PDAC_MAX_GAS),
        ("muscl", ct.c_int32), ("lm", ct.c_int32), ("lv", ct.c_int32),
        ("inmax", ct.c_int32), ("maxout", ct.c_int32),
        ("pmodel", ct.c_int32), ("drag_model", ct.c_int32), ("icpc", ct.c_int32),
        ("ord", ct.c_int32), ("cstop", ct.c_int32), ("iter_order", ct.c_int32),
        ("time_integration", ct.c_int32),
        ("gas_viscosity", ct.c_int32), ("part_viscosity", ct.c_int32),
        ("repulsive_model", ct.c_int32), ("mus", ct.c_double * PDAC_MAX_SOLID),
        ("cmut", ct.c_double), ("iturb", ct.c_int32),
    ]


def dims_from_case(case, precision=PREC_FP64) -> Dims:
    return Dims(3 if case.three_d else 2, case.itc, case.nx,
                case.ny if case.three_d else 1, case.nz, case.nsolid,
                len(case.gas_type), precision)


def params_from_case(case, iter_order=ORDER_REDBLACK) -> Params:
    p = Params()
    p.alpha, p.alphagrav = case.alpha, case.alphagrav
    p.gravx, p.gravy, p.gravz = case.gravx, case.gravy, case.gravz
    p.beta, p.omega, p.delg = case.beta, case.omega, case.delg
    p.rlim, p.flim, p.tref = case.rlim, case.flim, case.tref
    p.void_fraction = case.void_fraction
    for s in range(case.nsolid):
        p.dk[s] = case.diameter[s] * 1.0e-6  # initc, input.f:1095
        p.rl[s] = case.density[s]
        p.phis[s] = case.sphericity[s]
        p.cps[s] = case.specific_heat[s]
        p.kap[s] = case.thermal_conductivity[s]
        p.plim[s] = case.twophase_limit[s]
    for g in range(7):
        p.atm_ygc[g] = case.atm_ygc[g]
    for g, t in enumerate(case.gas_type):
        p.gas_type[g] = t
    p.muscl, p.lm, p.lv = case.muscl, case.mass_limiter, case.vel_limiter
    p.inmax, p.maxout = case.inmax, case.maxout
    p.pmodel, p.drag_model, p.icpc, p.ord = case.pmodel, case.drag_model, case.icpc, case.ord
    p.cstop = 1 if case.cstop else 0
    p.iter_order = iter_order
    p.time_integration = case.time_integration
    p.gas_viscosity = 1 if case.gas_viscosity else 0
    p.part_viscosity = 1 if case.part_viscosity else 0
    p.repulsive_model = case.repulsive_model
    for s in range(case.nsolid):
        p.mus[s] = case.viscosity[s]  # cmus, input.f:1100
    p.cmut, p.iturb = case.cmut, case.iturb
    return p


def dptr(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))
