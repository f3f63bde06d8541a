#!/usr/bin/env python3
"""Throughput of the PDAC semi-implicit step on one MI355X.

Workload: the 2D cylindrical collapsing eruption column, 512 radial x 1024
vertical cells (1 m at the axis stretching to 20 m), 2 particle classes
(50 and 500 um), air + water vapour, stratified atmosphere, axial vent of
50 m radius at w = 150 m/s with seeded 2 % noise; plain FP64, rungekut = 1,
MUSCL + minmod.  Synthetic seeded initial state, resident in HBM before the
timed region.

One step is main_prog's step (prog.f:143-329): boundary, fieldn, tilde,
iter (red-black sweeps until every cell's mass residual is below
delg*rgp), ygas, htilde, ftem.  The metric is Mcell-steps/s over the whole
mesh including boundary cells.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line with a "roofline" object for the dominant kernel
(algorithmic bytes per launch over its mean HIP-event duration) and a
"cpu_baseline" object (the CPU restatement of the same step timed on a
bounded sample of the same workload on this host, one core).
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
# HBM bytes per launch of the sweep kernel (k_sweep, one colour phase), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE over every dispatch of the default
# workload (FP32-flux form), each scaled by what the counter reads for a copy
# of known size (--copy-calibration), averaged over the launches that did work
SWEEP_TRAFFIC_PER_LAUNCH = 16913454.0
# order of enum pdac_kernel (include/pdacgpu.h)
KERNELS = ["boundary", "fieldn", "momflux", "tilde", "predict", "gasflux", "snapshot",
           "sweep", "ygas_flux", "ygas_update", "htilde_flux", "htilde_update", "ftem", "fill",
           "sweep_check", "iterate", "xinit", "xflux", "xrhs", "xchain", "xmass", "xsolve"]


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


# kernels of each reference subroutine
GROUPS = {"F": ["fieldn"], "T": ["momflux", "tilde"], "I0": ["predict", "gasflux"],
          "Is": ["snapshot", "sweep_check", "sweep", "iterate"], "Y": ["ygas_flux", "ygas_update"],
          "H": ["htilde_flux", "htilde_update", "ftem"]}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_case(nx, nz):
    from pdac_amd.cases import eruption_column
    return eruption_column(nx=nx, nz=nz, nsteps=1)


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


def cpu_sample(case, nsteps):
    """The CPU restatement (oracle) on the same case, from the same state."""
    from oracle.binding import cpu_solver
    from pdac_amd import abi
    c = cpu_solver(case, iter_order=abi.ORDER_REDBLACK)
    t0 = time.perf_counter()
    nits = []
    for _ in range(nsteps):
        c.step(1)
        nits.append(c.nit_last)
    t = time.perf_counter() - t0
    c.close()
    return t, nits


def run_form(args, precision, ids):
    """Warm-up, the timed region and the event-timed region of one precision
    form on a fresh copy of the workload; returns timings and the final p and
    particle volume fractions."""
    import ctypes as ct

    from pdac_amd.gpu import gpu_solver
    case = make_case(args.nx, args.nz)
    n = case.ntot
    g = gpu_solver(case, device=0, precision=precision)
    t0 = time.perf_counter()
    if args.warmup:
        g.step(args.warmup)
    g.lib.pdac_sync(g.ctx)
    tag = "fp32-flux" if precision else "fp64"
    log(f"[{tag}] warmup {args.warmup} steps: {time.perf_counter() - t0:.2f} s, nit {g.nit_last}")
    # timed region: K steps, nothing recorded but the stream synchronisations
    nit_before = g.nit_total
    g.lib.pdac_sync(g.ctx)
    t0 = time.perf_counter()
    g.step(args.steps)
    g.lib.pdac_sync(g.ctx)
    elapsed = time.perf_counter() - t0
    sweeps = g.nit_total - nit_before
    log(f"[{tag}] timed {args.steps} steps: {elapsed:.3f} s, {sweeps} sweeps")
    # second region of the same length with HIP events around the kernels
    # (recording events costs host time per launch, so it is kept out of
    # the timed region above)
    if args.kernel_timing != "none":
        mask = 0xFFFFFFFF if args.kernel_timing == "all" else (1 << ids["sweep"]) | (1 << ids["iterate"])
        g.lib.pdac_profile_select(g.ctx, mask)
        g.lib.pdac_profile_enable(g.ctx, 1)
    visits0 = ct.c_int64()
    if g.lib.pdac_profile_visits(g.ctx, ct.byref(visits0)) != 0:
        raise RuntimeError(g.last_error())
    g.lib.pdac_profile_reset(g.ctx)
    nit_before = g.nit_total
    t1 = time.perf_counter()
    g.step(args.steps)
    g.lib.pdac_sync(g.ctx)
    elapsed_prof = time.perf_counter() - t1
    sweeps_prof = g.nit_total - nit_before
    prof = profile(g, ids)
    visits = ct.c_int64()
    if g.lib.pdac_profile_visits(g.ctx, ct.byref(visits)) != 0:
        raise RuntimeError(g.last_error())
    g.lib.pdac_profile_enable(g.ctx, 0)
    log(f"[{tag}] event-timed {args.steps} steps: {elapsed_prof:.3f} s, {sweeps_prof} sweeps")
    log(f"[{tag}] sweep work-list cells over the whole run: {visits0.value + visits.value}")
    rho = np.asarray(case.density[:case.nsolid], dtype=np.float64).reshape(-1, 1)
    final = {"p": g.get("p"), "eps_s": g.get("rlk").reshape(case.nsolid, -1) / rho}
    g.close()
    return {"case": case, "n": n, "elapsed": elapsed, "sweeps": sweeps, "elapsed_prof": elapsed_prof,
            "sweeps_prof": sweeps_prof, "prof": prof, "visits": visits.value, "final": final}


def make_explicit_case(nx, nz):
    """2D cylindrical eruption column on uniform 2 m cells advanced by the
    explicit solver (time_integration = 1, rungekut = 1: one three-stage
    Runge-Kutta update per step)."""
    from pdac_amd.cases import eruption_column
    case = eruption_column(nx=nx, nz=nz, dr0=2.0, dr1=2.0, nsteps=1, dt=1.0e-4)
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
                        f"uniform 2 m cells, 2 particle classes, explicit three-stage "
                        f"Runge-Kutta (time_integration = 1)",
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
                               eps_vent=float(rng.uniform(0.005, 0.03)), seed=1000 + m)
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
                        f"2 classes, explicit Runge-Kutta, FP32-flux form, dt 1e-4 s",
            "ntot": ncells, "steps": args.ensemble_steps,
            "tiles": {k: f"{v[0]}x{v[1]}" for k, v in chosen.items()},
            "tuner_us": {k: [round(x, 2) for x in v] for k, v in times.items() if max(v) > 0},
            "max_eos_failures": max(fails), "max_cfl": round(max(cfl), 4)}


def three_d_summary(args):
    """The 3D shape of the precision-form comparison: 256 x 256 x 192 cells of
    10 m, three particle classes, square vent on flat ground (the cone
    topography needs immersed boundaries, which this library does not have),
    semi-implicit, FP64 and FP32-flux forms from the same initial state."""
    from pdac_amd import abi
    from pdac_amd.cases import summit_vent_3d
    from pdac_amd.gpu import gpu_solver
    t0 = time.perf_counter()
    case = summit_vent_3d(nx=args.three_d_nx, ny=args.three_d_nx, nz=args.three_d_nz, nsteps=1)
    n, ns = case.ntot, case.nsolid
    log(f"[3d] case of {n} cells set up in {time.perf_counter() - t0:.1f} s")
    rho = np.asarray(case.density[:ns], dtype=np.float64).reshape(-1, 1)
    k = args.three_d_steps
    forms, final = {}, {}
    for label, prec in (("fp64", abi.PREC_FP64), ("fp32flux", abi.PREC_FP32FLUX)):
        t1 = time.perf_counter()
        g = gpu_solver(case, device=0, precision=prec)
        g.lib.pdac_sync(g.ctx)
        log(f"[3d {label}] uploaded in {time.perf_counter() - t1:.1f} s")
        if args.three_d_warmup:
            g.step(args.three_d_warmup)
        g.lib.pdac_sync(g.ctx)
        nit0 = g.nit_total
        t2 = time.perf_counter()
        g.step(k)
        g.lib.pdac_sync(g.ctx)
        elapsed = time.perf_counter() - t2
        sweeps = g.nit_total - nit0
        log(f"[3d {label}] {k} steps: {elapsed:.3f} s, {sweeps} sweeps")
        final[label] = {"p": g.get("p"), "eps_s": g.get("rlk").reshape(ns, -1) / rho}
        g.close()
        forms[label] = {"value": round(n * k / elapsed / 1e6, 3), "ms_per_step": round(elapsed / k * 1e3, 4),
                        "sweeps_per_step": round(sweeps / k, 3)}
    a, b = final["fp32flux"], final["fp64"]
    return {"metric": "Mcell-steps/s",
            "workload": f"3D Cartesian {args.three_d_nx}x{args.three_d_nx}x{args.three_d_nz} cells of 10 m, "
                        f"3 particle classes (30/250/1000 um), square 100 m vent on flat ground at "
                        f"w = 100 m/s, T = 1000 K, particle fraction 0.02; semi-implicit; no topography",
            "ntot": n, "steps": k, "warmup": args.three_d_warmup, "dt": case.dt, "forms": forms,
            "fp32flux_speedup": round(forms["fp32flux"]["value"] / forms["fp64"]["value"], 4),
            "fp32flux_rel_l2": {"after_steps": args.three_d_warmup + k,
                                "p": rel_l2(a["p"], b["p"]),
                                "eps_s": [rel_l2(a["eps_s"][s_], b["eps_s"][s_]) for s_ in range(ns)]}}


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
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--nz", type=int, default=1024)
    ap.add_argument("--kernel-timing", choices=["all", "dominant", "none"], default="dominant",
                    help="HIP events around every kernel, around the iteration's work-list "
                         "kernel only (the dominant one), or none")
    ap.add_argument("--forms", choices=["both", "fp64", "fp32flux"], default="both",
                    help="precision forms to run; with both, value is the FP32-flux form's")
    ap.add_argument("--cpu-steps", type=int, default=16,
                    help="steps of the CPU sample (0: skip the CPU baseline)")
    ap.add_argument("--explicit-steps", type=int, default=100,
                    help="steps of the explicit-solver workload (0: skip it)")
    ap.add_argument("--explicit-nx", type=int, default=256)
    ap.add_argument("--explicit-nz", type=int, default=512)
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the hazard ensemble (0: skip it)")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-seed", type=int, default=2024)
    ap.add_argument("--three-d-steps", type=int, default=20,
                    help="steps of the 3D workload per precision form (0: skip it)")
    ap.add_argument("--three-d-warmup", type=int, default=5)
    ap.add_argument("--three-d-nx", type=int, default=256)
    ap.add_argument("--three-d-nz", type=int, default=192)
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
    runs = {}
    if args.forms in ("both", "fp64"):
        runs["fp64"] = run_form(args, abi.PREC_FP64, ids)
    if args.forms in ("both", "fp32flux"):
        runs["fp32flux"] = run_form(args, abi.PREC_FP32FLUX, ids)
    main_form = "fp32flux" if "fp32flux" in runs else "fp64"
    r = runs[main_form]
    case, n = r["case"], r["n"]
    D, S, G = (3 if case.three_d else 2), case.nsolid, len(case.gas_type)
    elapsed, sweeps, prof, visits = r["elapsed"], r["sweeps"], r["prof"], r["visits"]
    sweeps_prof = r["sweeps_prof"]
    value = n * args.steps / elapsed / 1e6
    # dominant kernel and its roofline
    w = words(D, S, G)
    dom = max(KERNELS, key=lambda k: prof[k][1])
    launches, ms = prof[dom]
    if launches == 0:
        dom, ms = "none", 0.0
    per_launch_s = ms / 1e3 / max(launches, 1) if launches else float("nan")
    traffic = None
    worklist = None
    if dom == "sweep":
        # one launch is one colour phase of a sweep: the check of every cell of
        # the colour and the in-cell iteration of those not converged at visit
        # start.  Algorithmic bytes follow the reference's sweep, which reads
        # and writes the 62 sweep words of every cell; the kernel skips cells
        # whose inputs did not change since their last check (same result)
        cells = n / 2.0
        bytes_launch = 8.0 * w["Is"] * cells
        if SWEEP_TRAFFIC_PER_LAUNCH is not None:
            traffic = round(SWEEP_TRAFFIC_PER_LAUNCH)
        unit_note = (f"cells per launch = ntot/2 = {cells:.0f} (one colour phase of a sweep), "
                     f"{w['Is']} words each")
        wl = visits / max(launches, 1)
        worklist = {"cells_per_launch": round(wl, 1), "bytes_per_launch": 8.0 * w["Is"] * wl,
                    "gbs": round(8.0 * w["Is"] * wl / per_launch_s / 1e9, 2)}
    elif dom == "iterate":
        # one launch runs every sweep of one iter call
        bytes_launch = 8.0 * w["Is"] * n * sweeps_prof / max(launches, 1)
        unit_note = (f"ntot cells x sweeps per launch ({sweeps_prof / max(launches, 1):.2f}), "
                     f"{w['Is']} words per cell and sweep")
    elif dom == "none":
        bytes_launch, unit_note = 0.0, "kernel timing off"
    else:
        grp = next(k for k, v in GROUPS.items() if dom in v)
        share = len(GROUPS[grp])
        bytes_launch = 8.0 * w[grp] * n / share
        unit_note = f"ntot cells, 1/{share} of subroutine group {grp}"
    achieved = bytes_launch / per_launch_s / 1e9
    # whole-step algorithmic bytes
    steps_per_s = args.steps / elapsed
    b_step = 8.0 * (w["F"] + w["T"] + w["I0"] + w["Y"] + w["H"]) + 8.0 * w["Is"] * sweeps / args.steps
    step_gbs = b_step * n * steps_per_s / 1e9
    kern = {k: {"launches": prof[k][0], "ms_total": round(prof[k][1], 3)} for k in KERNELS if prof[k][0]}
    out = {
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
        "config": {"workload": f"2D cylindrical eruption column {args.nx}x{args.nz}, "
                               f"2 particle classes, semi-implicit",
                   "form": "FP32-flux" if main_form == "fp32flux" else "FP64 plain",
                   "ntot": n, "nsolid": S, "ngas": G, "rungekut": case.rungekut,
                   "dt": case.dt, "sweeps_per_step": round(sweeps / args.steps, 3),
                   "iter_order": "red-black"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": bytes_launch, "mean_launch_us": round(per_launch_s * 1e6, 2),
                     "launches": launches, "units": unit_note,
                     "worklist": worklist},
        "sweep_visits_per_step": round(visits / args.steps, 1),
        "event_region": {"steps": args.steps, "seconds": round(r["elapsed_prof"], 4),
                         "sweeps_per_step": round(sweeps_prof / args.steps, 3),
                         "timing": args.kernel_timing},
        "step_alg_gbs": round(step_gbs, 2),
        "step_alg_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "kernels": kern,
    }
    if len(runs) == 2:
        a, b = runs["fp32flux"], runs["fp64"]
        out["fp64_value"] = round(n * args.steps / b["elapsed"] / 1e6, 3)
        out["fp32flux_speedup"] = round(b["elapsed"] / a["elapsed"], 4)
        out["fp32flux_rel_l2"] = {
            "after_steps": args.warmup + 2 * args.steps,
            "p": rel_l2(a["final"]["p"], b["final"]["p"]),
            "eps_s": [rel_l2(a["final"]["eps_s"][s_], b["final"]["eps_s"][s_]) for s_ in range(S)]}
    if args.explicit_steps > 0:
        out["explicit"] = explicit_summary(args, ids)
    if args.ensemble_members > 0:
        out["ensemble"] = ensemble_summary(args)
    if args.three_d_steps > 0:
        out["three_d"] = three_d_summary(args)
    if args.cpu_steps > 0:
        t, nits = cpu_sample(make_case(args.nx, args.nz), args.cpu_steps)
        out["cpu_baseline"] = {
            "value": round(n * args.cpu_steps / t / 1e6, 5), "unit": "Mcell-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps of the same {args.nx}x{args.nz} case from its "
                      f"initial state (sweeps {nits}), {t:.1f} s on 1 host core"}
        log(f"cpu sample: {t:.1f} s for {args.cpu_steps} steps")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
