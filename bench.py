#!/usr/bin/env python3
"""Throughput of the PDAC semi-implicit step on one MI355X.

Workload: the 3D three-class summit-vent case with topography, 256 x 256 x
192 cells of 10 m: a Gaussian cone (height 1500 m, sigma 1500 m) with seeded
fractal roughness of 30 m amplitude and a flat summit platform, treated by
immersed boundaries (the reference's own topography and immersed-boundary
set-up, restated); a square inlet vent of 11 x 11 cells on the summit at
w = 100 m/s (seeded 2 % noise per cell), T = 1000 K, particle fraction 0.02
(30, 250, 1000 um; 2700, 2000, 1000 kg/m3), air + water vapour, stratified
atmosphere; rungekut = 1, MUSCL + minmod, red-black iteration.  Synthetic,
seeded initial state, resident in HBM before the timed region.

One step is main_prog's step (prog.f:143-329): boundary, fieldn, tilde,
iter (red-black sweeps until every cell's mass residual is below
delg*rgp), ygas, htilde, ftem.  The metric is Mcell-steps/s over the whole
mesh including boundary cells, in the FP32-flux form (value) and the plain
FP64 form (fp64_value), with the relative L2 difference of p and of each
particle volume fraction between the forms after 1000 steps.

Each form runs a fixed spin-up (--spinup steps) before --warmup and the
--steps timed ones, so the timed window starts from a developed flow
whatever the warm-up, then a second window of the same length with HIP
events on the dominant kernel, then steps on to --error-steps in all.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line with a "roofline" object for the dominant kernel
(algorithmic bytes per launch over its mean HIP-event duration, with the
HBM traffic rocprofv3 measured for it) and a "cpu_baseline" object (the CPU
restatement of the same step, one core, on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# HBM bytes per working launch of the sweep kernel (k_sweep, one colour
# phase) on the default workload (FP32-flux form): rocprofv3 FETCH_SIZE and
# WRITE_SIZE over its dispatches, each scaled by what the counter reads for a
# copy of known size (--copy-calibration): 45.47 MB read + 6.55 MB written
# over the 50,772 working dispatches of the default event region (steps
# 260-360); None until measured
SWEEP_TRAFFIC_PER_LAUNCH = 52.02e6
# order of enum pdac_kernel (include/pdacgpu.h)
KERNELS = ["boundary", "fieldn", "momflux", "tilde", "predict", "gasflux", "snapshot",
           "sweep", "ygas_flux", "ygas_update", "htilde_flux", "htilde_update", "ftem", "fill",
           "sweep_check", "iterate", "xinit", "xflux", "xrhs", "xchain", "xmass", "xsolve",
           "stress"]


def words(D, S, G):
    """Algorithmic words per cell of each reference subroutine: the distinct
    per-cell fields read and written once (fieldn, tilde, iter setup, one
    iter sweep, ygas, htilde+ftem)."""
    P = S + 1
    A = P * (P + 1) // 2
    return {
        "F": (5 + 2 * S + D * P + 2 * G) + (6 + 2 * S + G + D * P),
        "T": (2 * P + 2 * D * P + 3) + (D * P + D * A),
        "I0": (D * P + D * A + 4 + S) + (D * P + D + 2),
        "Is": (5 + 2 * S + G + 3 * D * P + D * A + 2) + (5 + S + 2 * D * P),
        "Y": (1 + 2 * G + D) + (2 * G),
        "H": (13 + 5 * S + D * P + G) + (3 + 3 * S + G),
    }


def check_words(D, S):
    """Words a sweep's residual check reads and writes for a cell whose
    inputs changed (calc_res and calc_eps, iter.f:743-771, 927-994): rgp,
    ep, rog, rgpn and the tolerance; the gas mass fluxes of the 2D faces; rlk,
    rlkn and the particle mass fluxes of the 2D faces per class; rlk, ep and
    rgp written."""
    return 5 + 2 * D + S * (2 + 2 * D) + S + 2


# kernels of each reference subroutine
GROUPS = {"F": ["fieldn"], "T": ["momflux", "tilde"], "I0": ["predict", "gasflux"],
          "Is": ["snapshot", "sweep_check", "sweep", "iterate"], "Y": ["ygas_flux", "ygas_update"],
          "H": ["htilde_flux", "htilde_update", "ftem"]}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_case(args):
    """The default workload: the 3D summit-vent cone with immersed boundaries."""
    from pdac_amd.cases import summit_cone_3d
    return summit_cone_3d(nx=args.nx, ny=args.ny, nz=args.nz, nsteps=1)


def profile(g, ids):
    import ctypes as ct
    out = {}
    for name in KERNELS:
        n, ms = ct.c_int64(), ct.c_double()
        rc = g.lib.pdac_profile_read(g.ctx, ids[name], ct.byref(n), ct.byref(ms))
        if rc != 0:
            raise RuntimeError(f"pdac_profile_read: {g.last_error()}")
        out[name] = (n.value, ms.value)
    return out


def counters(g):
    import ctypes as ct
    v, c = ct.c_int64(), ct.c_int64()
    if g.lib.pdac_profile_visits(g.ctx, ct.byref(v)) != 0 or g.lib.pdac_profile_checks(g.ctx, ct.byref(c)) != 0:
        raise RuntimeError(g.last_error())
    return v.value, c.value


def cpu_sample(args):
    """The CPU restatement (oracle, red-black order, its order-independent
    cell loops on OpenMP threads) on a bounded sample of the workload: the
    same case on a 64 x 64 x 48 mesh of the same 10 m cells with the cone
    scaled by 1/4 (same slopes), from its initial state."""
    from oracle.binding import cpu_solver
    from pdac_amd import abi
    from pdac_amd.cases import summit_cone_3d
    case = summit_cone_3d(nx=64, ny=64, nz=48, height=375.0, sigma=375.0, rough=7.5, nsteps=1)
    c = cpu_solver(case, iter_order=abi.ORDER_REDBLACK)
    threads = int(c.lib.pdac_cpu_threads())
    nits, secs = [], []
    for _ in range(args.cpu_steps):
        t0 = time.perf_counter()
        c.step(1)
        secs.append(time.perf_counter() - t0)
        nits.append(c.nit_last)
    c.close()
    return case, secs, nits, threads


def run_form(args, case, precision, ids):
    """Spin-up, warm-up, the timed region and the event-timed region of one
    precision form on a fresh copy of the workload, then steps on to
    --error-steps; returns timings and the final p and particle volume
    fractions."""
    from pdac_amd.gpu import gpu_solver
    n = case.ntot
    tag = "fp32-flux" if precision else "fp64"
    t0 = time.perf_counter()
    g = gpu_solver(case, device=0, precision=precision)
    g.lib.pdac_sync(g.ctx)
    log(f"[{tag}] uploaded in {time.perf_counter() - t0:.1f} s")
    done = 0
    t0 = time.perf_counter()
    if args.spinup:
        g.step(args.spinup)
    if args.warmup:
        g.step(args.warmup)
    g.lib.pdac_sync(g.ctx)
    done += args.spinup + args.warmup
    log(f"[{tag}] spin-up {args.spinup} + warm-up {args.warmup} steps: {time.perf_counter() - t0:.2f} s")
    # timed region: K steps, nothing recorded but the stream synchronisations
    nit_before = g.nit_total
    g.lib.pdac_sync(g.ctx)
    t0 = time.perf_counter()
    g.step(args.steps)
    g.lib.pdac_sync(g.ctx)
    elapsed = time.perf_counter() - t0
    sweeps = g.nit_total - nit_before
    done += args.steps
    log(f"[{tag}] timed {args.steps} steps: {elapsed:.3f} s, {sweeps} sweeps")
    # second region of the same length with HIP events around the kernels
    # (events cost host time per launch, so the timed region records none)
    if args.kernel_timing != "none":
        mask = 0xFFFFFFFF if args.kernel_timing == "all" else (1 << ids["sweep"]) | (1 << ids["iterate"])
        g.lib.pdac_profile_select(g.ctx, mask)
        g.lib.pdac_profile_enable(g.ctx, 1)
    g.lib.pdac_profile_reset(g.ctx)
    nit_before = g.nit_total
    t1 = time.perf_counter()
    # events are folded into the totals every few steps (the event pool
    # holds a bounded number of launches)
    left = args.steps
    prof = profile(g, ids)
    while left:
        k = min(left, 5)
        g.step(k)
        prof = profile(g, ids)
        left -= k
    g.lib.pdac_sync(g.ctx)
    elapsed_prof = time.perf_counter() - t1
    sweeps_prof = g.nit_total - nit_before
    done += args.steps
    visits, checks = counters(g)
    g.lib.pdac_profile_enable(g.ctx, 0)
    log(f"[{tag}] event-timed {args.steps} steps: {elapsed_prof:.3f} s, {sweeps_prof} sweeps, "
        f"{checks} cells checked, {visits} visited")
    rest = max(0, args.error_steps - done)
    if rest:
        t2 = time.perf_counter()
        g.step(rest)
        g.lib.pdac_sync(g.ctx)
        done += rest
        log(f"[{tag}] {rest} more steps to {done}: {time.perf_counter() - t2:.2f} s")
    rho = np.asarray(case.density[:case.nsolid], dtype=np.float64).reshape(-1, 1)
    final = {"p": g.get("p"), "eps_s": g.get("rlk").reshape(case.nsolid, -1) / rho}
    g.close()
    return {"n": n, "elapsed": elapsed, "sweeps": sweeps, "elapsed_prof": elapsed_prof,
            "sweeps_prof": sweeps_prof, "prof": prof, "visits": visits, "checks": checks,
            "final": final, "steps_total": done}


def make_explicit_case(nx, nz):
    """2D cylindrical eruption column on uniform 2 m cells advanced by the
    explicit solver (time_integration = 1, rungekut = 1: one three-stage
    Runge-Kutta update per step)."""
    from pdac_amd.cases import eruption_column
    case = eruption_column(nx=nx, nz=nz, dr0=2.0, dr1=2.0, nsteps=1, dt=1.0e-4, vent="inlet")
    case.time_integration = 1
    case.rungekut = 1
    return case


def run_explicit(args, precision, order, ids):
    """Warm-up and timed region of the explicit step on a fresh copy of its
    workload, then a second region with HIP events around every kernel."""
    from pdac_amd.gpu import gpu_solver
    case = make_explicit_case(args.explicit_nx, args.explicit_nz)
    g = gpu_solver(case, device=0, precision=precision, iter_order=order)
    if args.warmup:
        g.step(args.warmup)
    g.lib.pdac_sync(g.ctx)
    t0 = time.perf_counter()
    g.step(args.explicit_steps)
    g.lib.pdac_sync(g.ctx)
    elapsed = time.perf_counter() - t0
    g.lib.pdac_profile_select(g.ctx, 0xFFFFFFFF)
    g.lib.pdac_profile_enable(g.ctx, 1)
    g.lib.pdac_profile_reset(g.ctx)
    g.step(args.explicit_steps)
    g.lib.pdac_sync(g.ctx)
    prof = profile(g, ids)
    g.lib.pdac_profile_enable(g.ctx, 0)
    g.close()
    return case, elapsed, prof


def explicit_summary(args, ids):
    from pdac_amd import abi
    out = {}
    runs = {"reference order": (abi.PREC_FP64, abi.ORDER_REDBLACK),
            "stage order": (abi.PREC_FP64, abi.ORDER_STAGE),
            "reference order FP32-flux": (abi.PREC_FP32FLUX, abi.ORDER_REDBLACK)}
    for label, (prec, order) in runs.items():
        case, elapsed, prof = run_explicit(args, prec, order, ids)
        n = case.ntot
        k = args.explicit_steps
        xk = {name: {"launches": prof[name][0], "ms_per_step": round(prof[name][1] / k, 4)}
              for name in KERNELS if prof[name][0]}
        out[label] = {"value": round(n * k / elapsed / 1e6, 3),
                      "ms_per_step": round(elapsed / k * 1e3, 4), "kernels": xk}
        log(f"[explicit {label}] {k} steps: {elapsed:.3f} s")
    return {"metric": "Mcell-steps/s",
            "workload": f"2D cylindrical eruption column {args.explicit_nx}x{args.explicit_nz}, "
                        f"uniform 2 m cells, 2 particle classes, sustained inlet vent, "
                        f"explicit three-stage Runge-Kutta (time_integration = 1)",
            "ntot": case.ntot, "steps": args.explicit_steps, "dt": case.dt, "forms": out}


def ensemble_summary(args):
    """The hazard ensemble: members of uneven size (nx drawn from [128, 512],
    nz = 2 nx, uniform 2 m cells), vent w, T and particle fraction drawn per
    member, explicit step in the FP32-flux form, all advanced by batched
    graph-replayed launches after the tile autotuner."""
    from pdac_amd import abi
    from pdac_amd.cases import eruption_column
    from pdac_amd.gpu import Ensemble, gpu_solver
    rng = np.random.default_rng(args.ensemble_seed)
    t0 = time.perf_counter()
    members = []
    for m in range(args.ensemble_members):
        nx = int(rng.integers(128, 513))
        case = eruption_column(nx=nx, nz=2 * nx, dr0=2.0, dr1=2.0, nsteps=1, dt=1.0e-4,
                               w_vent=float(rng.uniform(50.0, 200.0)),
                               t_vent=float(rng.uniform(800.0, 1200.0)),
                               eps_vent=float(rng.uniform(0.005, 0.03)), seed=1000 + m, vent="inlet")
        case.time_integration = 1
        members.append(gpu_solver(case, device=0, precision=abi.PREC_FP32FLUX))
    ncells = sum(m.case.ntot for m in members)
    log(f"[ensemble] {len(members)} members, {ncells} cells, set up in {time.perf_counter() - t0:.1f} s")
    e = Ensemble(members)
    chosen, times = e.tune(reps=3)
    log(f"[ensemble] tiles {chosen}")
    e.step(args.warmup)
    t1 = time.perf_counter()
    fails, cfl = e.step(args.ensemble_steps)
    elapsed = time.perf_counter() - t1
    log(f"[ensemble] {args.ensemble_steps} steps: {elapsed:.3f} s, max cfl {max(cfl):.3f}")
    e.close()
    for m in members:
        m.close()
    return {"metric": "Mcell-steps/s",
            "value": round(ncells * args.ensemble_steps / elapsed / 1e6, 3),
            "ms_per_step": round(elapsed / args.ensemble_steps * 1e3, 4),
            "workload": f"{args.ensemble_members} 2D cylindrical eruption columns, nx drawn from "
                        f"[128, 512] (seed {args.ensemble_seed}), nz = 2 nx, uniform 2 m cells, "
                        f"vent w U[50,200] m/s, T U[800,1200] K, particle fraction U[0.005,0.03], "
                        f"2 classes, sustained inlet vents, explicit Runge-Kutta, FP32-flux form, dt 1e-4 s",
            "ntot": ncells, "steps": args.ensemble_steps,
            "tiles": {k: f"{v[0]}x{v[1]}" for k, v in chosen.items()},
            "tuner_us": {k: [round(x, 2) for x in v] for k, v in times.items() if max(v) > 0},
            "max_eos_failures": max(fails), "max_cfl": round(max(cfl), 4)}


def column_summary(args):
    """The 2D cylindrical collapsing eruption column, 512 x 1024 cells (1 m
    at the axis stretching to 20 m), two classes, sustained 50 m inlet vent
    at w = 150 m/s, free in/outflow side and lid, semi-implicit, FP64 and
    FP32-flux forms from the same initial state."""
    from pdac_amd import abi
    from pdac_amd.cases import eruption_column
    from pdac_amd.gpu import gpu_solver
    case = eruption_column(nx=512, nz=1024, nsteps=1, vent="inlet", top=4)
    n, k = case.ntot, args.column_steps
    forms = {}
    for label, prec in (("fp64", abi.PREC_FP64), ("fp32flux", abi.PREC_FP32FLUX)):
        g = gpu_solver(case, device=0, precision=prec)
        g.step(args.column_warmup)
        g.lib.pdac_sync(g.ctx)
        nit0 = g.nit_total
        t0 = time.perf_counter()
        g.step(k)
        g.lib.pdac_sync(g.ctx)
        elapsed = time.perf_counter() - t0
        sweeps = g.nit_total - nit0
        g.close()
        forms[label] = {"value": round(n * k / elapsed / 1e6, 3), "ms_per_step": round(elapsed / k * 1e3, 4),
                        "sweeps_per_step": round(sweeps / k, 3)}
        log(f"[column {label}] {k} steps: {elapsed:.3f} s, {sweeps} sweeps")
    return {"metric": "Mcell-steps/s",
            "workload": "2D cylindrical eruption column 512x1024, 2 particle classes, sustained 50 m inlet vent "
                        "at 150 m/s, free in/outflow side and lid, semi-implicit",
            "ntot": n, "steps": k, "warmup": args.column_warmup, "dt": case.dt, "forms": forms}


def copy_calibration():
    """A plain device copy of known size (3 x 64 MiB) for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    n = 8 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 8, "copies": 3}), flush=True)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--spinup", type=int, default=150,
                    help="steps before the warm-up, so the timed window starts from a developed flow")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps each form runs in all; the forms' difference is measured after them")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=192)
    ap.add_argument("--kernel-timing", choices=["all", "dominant", "none"], default="dominant",
                    help="HIP events around every kernel, around the sweep kernel only (the dominant "
                         "one), or none")
    ap.add_argument("--forms", choices=["both", "fp64", "fp32flux"], default="both",
                    help="precision forms to run; with both, value is the FP32-flux form's")
    ap.add_argument("--cpu-steps", type=int, default=40,
                    help="steps of the CPU sample (0: skip the CPU baseline)")
    ap.add_argument("--column-steps", type=int, default=200,
                    help="steps of the 2D eruption column per form (0: skip it)")
    ap.add_argument("--column-warmup", type=int, default=20)
    ap.add_argument("--explicit-steps", type=int, default=100,
                    help="steps of the explicit-solver workload (0: skip it)")
    ap.add_argument("--explicit-nx", type=int, default=256)
    ap.add_argument("--explicit-nz", type=int, default=512)
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the hazard ensemble (0: skip it)")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-seed", type=int, default=2024)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 3 x 64 MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the whole domain lives on one MI355X "
              f"(--gpus {args.gpus} requested).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    from pdac_amd import abi

    ids = {name: i for i, name in enumerate(KERNELS)}
    t0 = time.perf_counter()
    case = make_case(args)
    case.flags()
    if getattr(case, "immb", 0):
        case.immersed()   # forcing points, filled cells, face masks (cached for the solvers)
    setup_s = time.perf_counter() - t0
    log(f"case of {case.ntot} cells set up on the host in {setup_s:.1f} s")
    runs = {}
    if args.forms in ("both", "fp32flux"):
        runs["fp32flux"] = run_form(args, case, abi.PREC_FP32FLUX, ids)
    if args.forms in ("both", "fp64"):
        runs["fp64"] = run_form(args, case, abi.PREC_FP64, ids)
    main_form = "fp32flux" if "fp32flux" in runs else "fp64"
    r = runs[main_form]
    n = r["n"]
    D, S, G = 3, case.nsolid, len(case.gas_type)
    elapsed, sweeps, prof = r["elapsed"], r["sweeps"], r["prof"]
    sweeps_prof = r["sweeps_prof"]
    value = n * args.steps / elapsed / 1e6
    w = words(D, S, G)
    wc = check_words(D, S)
    dom = max(KERNELS, key=lambda k: prof[k][1])
    launches, ms = prof[dom]
    if launches == 0:
        dom, ms = "none", 0.0
    per_launch_s = ms / 1e3 / max(launches, 1) if launches else float("nan")
    traffic = None
    if dom == "sweep":
        # one launch is one colour phase of a sweep.  It reads the residual
        # inputs of the cells whose inputs changed since their last check
        # (check words) and runs the in-cell iteration, which reads and
        # writes the reference sweep's words of a cell, for those not
        # converged at visit start; the unchanged cells it skips would give
        # the same result again.  Units counted on the device over the
        # event region (pdac_profile_checks / pdac_profile_visits).
        checks, visits = r["checks"], r["visits"]
        bytes_launch = 8.0 * ((checks - visits) * wc + visits * w["Is"]) / max(launches, 1)
        if SWEEP_TRAFFIC_PER_LAUNCH is not None:
            traffic = round(SWEEP_TRAFFIC_PER_LAUNCH)
        unit_note = (f"per launch {checks / max(launches, 1):.0f} cells checked ({wc} words each) of which "
                     f"{visits / max(launches, 1):.0f} visited ({w['Is']} words each)")
    elif dom == "iterate":
        bytes_launch = 8.0 * w["Is"] * n * sweeps_prof / max(launches, 1)
        unit_note = f"ntot cells x sweeps per launch, {w['Is']} words per cell and sweep"
    elif dom == "none":
        bytes_launch, unit_note = 0.0, "kernel timing off"
    else:
        grp = next(k for k, v in GROUPS.items() if dom in v)
        share = len(GROUPS[grp])
        bytes_launch = 8.0 * w[grp] * n / share
        unit_note = f"ntot cells, 1/{share} of subroutine group {grp}"
    achieved = bytes_launch / per_launch_s / 1e9
    kern = {k: {"launches": prof[k][0], "ms_total": round(prof[k][1], 3)} for k in KERNELS if prof[k][0]}
    out = {
        "host_setup_s": round(setup_s, 2),
        "metric": "Mcell-steps/s",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_form == "fp64" else "f64 with f32 flux arithmetic",
        "data": "synthetic",
        "config": {"workload": f"3D summit vent on a Gaussian cone with fractal roughness, immersed "
                               f"boundaries, {args.nx}x{args.ny}x{args.nz} cells of 10 m, 3 particle classes, "
                               f"semi-implicit",
                   "form": "FP32-flux" if main_form == "fp32flux" else "FP64 plain",
                   "ntot": n, "nsolid": S, "ngas": G, "rungekut": case.rungekut,
                   "dt": case.dt, "spinup": args.spinup, "sweeps_per_step": round(sweeps / args.steps, 3),
                   "iter_order": "red-black"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": bytes_launch, "mean_launch_us": round(per_launch_s * 1e6, 2),
                     "launches": launches, "units": unit_note},
        "event_region": {"steps": args.steps, "seconds": round(r["elapsed_prof"], 4),
                         "sweeps_per_step": round(sweeps_prof / args.steps, 3),
                         "timing": args.kernel_timing},
        "kernels": kern,
    }
    if len(runs) == 2:
        a, b = runs["fp32flux"], runs["fp64"]
        out["fp64_value"] = round(n * args.steps / b["elapsed"] / 1e6, 3)
        out["fp32flux_speedup"] = round(b["elapsed"] / a["elapsed"], 4)
        out["fp32flux_rel_l2"] = {
            "after_steps": min(a["steps_total"], b["steps_total"]),
            "p": rel_l2(a["final"]["p"], b["final"]["p"]),
            "eps_s": [rel_l2(a["final"]["eps_s"][s_], b["final"]["eps_s"][s_]) for s_ in range(S)]}
    if args.column_steps > 0:
        out["column_2d"] = column_summary(args)
    if args.explicit_steps > 0:
        out["explicit"] = explicit_summary(args, ids)
    if args.ensemble_members > 0:
        out["ensemble"] = ensemble_summary(args)
    if args.cpu_steps > 0:
        ccase, secs, nits, threads = cpu_sample(args)
        t = sum(secs)
        # per-step cost of the CPU step against its sweep count (least
        # squares over the sampled steps): fixed part and per-sweep part
        b, a = np.polyfit(np.asarray(nits, float), np.asarray(secs), 1)
        out["cpu_baseline"] = {
            "value": round(ccase.ntot * args.cpu_steps / t / 1e6, 5), "unit": "Mcell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"the same cone case on a 64x64x48 mesh of 10 m cells (cone scaled by 1/4), the "
                      f"first {args.cpu_steps} steps from its initial state, {t:.1f} s on {threads} host "
                      f"threads (OpenMP cell loops; boundary and snapshot copies serial), red-black "
                      f"order; sweeps per step {nits} (mean {np.mean(nits):.1f})",
            "ms_per_step_fixed": round(a * 1e3, 2), "ms_per_sweep": round(b * 1e3, 3)}
        log(f"cpu sample: {t:.1f} s for {args.cpu_steps} steps, sweeps {nits}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
