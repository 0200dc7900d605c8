#!/usr/bin/env python3
"""Benchmark of the MI355X IFE-PIC hot path on one GPU.

Workload (default): 512 x 512 Cartesian mesh, three immersed conductors at -1, 0.5, 0,
16,777,216 electrons + 16,777,216 Ar+ ions (64 per cell per species), FP32 gather /
push / MCC with fixed-point (int64) deposition, FP32-inner / FP64-refined Jacobi-PCG to
1e-8, PPR recovery, null-collision MCC against 300 K argon at 3.2e20 m^-3, periodic
cell sort. A step is one full explicit PIC step (MoveAll x2, GetParChg_2D,
IFE_Solve_2D, GetEField_SIDG_PPR, MCC, sort every 20 steps) with every input resident in
HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg3] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
METRIC = "particle-pushes/s + PCG ms/solve @1e-8 rel residual, learned vs plain, 1 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(name: str) -> dict:
    """HBM bytes per launch per kernel from the committed rocprofv3 PMC summary: lines
    `traffic_bytes_per_launch <kernel>: <bytes>`."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return {}
    return {m.group(1): float(m.group(2)) for m in
            re.finditer(r"traffic_bytes_per_launch\s+(\S+):\s*([0-9.eE+]+)", p.read_text())}


def cpu_baseline(pkg, pl, w, sample_particles: int) -> dict:
    """One full explicit step of the CPU oracle (scalar C, 1 core) on a bounded sample of the
    same workload state: MoveAll + lookup + deposition for `sample_particles` particles per
    species, and the field half of the step (RHS, MY_PCG to 1e-8 in FP64, PPR) on the full
    512^2 mesh. Scaled to particle-pushes/s of the whole workload."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    dom = w["domain"]
    hp = pl.setup_array("HP")
    xs = hp[0::2][:: dom.nny].copy()
    ys = hp[1::2][: dom.nny].copy()
    ei, ni = pl.setup_array("ELEMENT_INDEX"), pl.setup_array("NODE_INDEX")
    efx, efy = pl.field("EFX"), pl.field("EFY")
    coef = pl.field("COEF_PPR").reshape(-1, 6)
    norm = dom.normalisation()
    n_total = 0
    t_part = 0.0
    for isp, sp in enumerate((pkg.ELECTRON, pkg.ARGON_ION)):
        p = pl.particles(isp)
        n_total += len(p["x"])
        k = min(sample_particles, len(p["x"]))
        qm = (sp["charge_C"] / pkg.ELECTRON_CHARGE) / (sp["mass_kg"] / pkg.ELECTRON_MASS)
        t0 = time.perf_counter()
        x, y, vx, vy, vz, alive, cnt, err = orc.move_all(
            64, xs, ys, ei, ni, dom.object_array(), efx, efy, coef[:, 1:].ravel(), qm, dom.dt, norm["v_ref"],
            sp["mass_kg"], 2.0 * pkg.EVTOK, 1, 0, orc.tag(isp, orc.RNG_EMIT),
            p["x"][:k], p["y"][:k], p["vx"][:k], p["vy"][:k], p["vz"][:k], p["loc"][:k] - 1)
        keep = alive.astype(bool)
        loc, _ = orc.locate(xs, ys, x[keep], y[keep])
        orc.deposit(64, xs, ys, 0, x[keep], y[keep], vx[keep], vy[keep], vz[keep], loc, sp["mass_kg"], norm["v_ref"], True)
        t_part += (time.perf_counter() - t0) * (len(p["x"]) / k)
    # field half on the full mesh
    nt = pl.setup_array("NODE_TYPE").reshape(-1, 2)[:, 1]
    rho = pl.field("RHO")
    phi = pl.field("PHI")
    rp, col, val, diag = (pl.setup_array(n) for n in ("A_ROWPTR", "A_COL", "A_VAL", "A_DIAG"))
    un = pl.setup_array("UNKNOWN_NODE") - 1
    t0 = time.perf_counter()
    b = orc.rhs(xs, ys, 0, nt, pl.setup_array("DIRICHLET_PHI"), rho, len(diag))
    xsol, _, it, _ = orc.pcg(rp, col, val, diag, b, phi[un], 1e-8, 50000)
    phi2 = phi.copy()
    phi2[un] = xsol
    orc.ppr(xs, ys, pl.setup_array("POINT_PATCH"), phi2)
    t_field = time.perf_counter() - t0
    t_step = t_part + t_field
    return {"value": n_total / t_step, "unit": "particle-pushes/s", "cores": 1, "kind": "port",
            "sample": f"oracle.c scalar FP64 step: mover+lookup+deposit timed on {sample_particles} particles per "
                      f"species and scaled to {n_total}; RHS + MY_PCG (1e-8, {it} it) + PPR on the full mesh; "
                      f"{t_step:.2f} s per step"}


def solve_compare(pl, reps: int) -> dict:
    """Plain (mode 1) vs learned (mode 2) inner solver on the same RHS and warm start: the
    charge density deposited in one step and the potential of the step before it (what the
    solve inside a step sees)."""
    import time as _t
    pl.set_solver(1)
    phi0 = pl.field("PHI")
    pl.step(1)
    rho = pl.field("RHO")
    res = {}
    for name, mode in (("plain", 1), ("learned", 2)):
        pl.set_solver(mode)
        times, its = [], []
        for _ in range(reps):
            phi = phi0.copy()
            t0 = _t.perf_counter()
            pl.IFE_Solve_2D(Rho=rho, Phi=phi)
            times.append(_t.perf_counter() - t0)
            info = pl.last_solve()
            its.append(info["iterations"])
        res[name] = {"ms_per_solve": 1e3 * float(np.median(times)), "iterations": int(np.median(its)),
                     "relres": info["relres"], "outer": info["outer"]}
    pl.set_solver(1)
    return res


def learned_leg(pkg, pl, w, args) -> dict:
    """Harvest pairs from the plain solver during untimed steps, train the network on the GPU,
    then compare the plain and learned solves and time full steps with the learned solver."""
    import time as _t
    log("bench: learned preconditioner: harvest")
    pl.learn_setup(12, (16, 16), seed=7)
    pl.learn_harvest(True, 512, 1)
    pl.step(args.harvest_steps)
    pl.learn_harvest(False, 512, 1)
    info = pl.learn_info()
    log(f"bench: {info['dataset']} pairs; training {args.train_steps} steps")
    t0 = _t.perf_counter()
    l0, l1 = pl.learn_train(args.train_steps, batch=4, lr=2e-3, seed=11)
    t_train = _t.perf_counter() - t0
    log("bench: solve comparison")
    cmp_ = solve_compare(pl, args.solve_reps)
    log(f"bench: {cmp_}")
    # whole steps with the learned solver (same step contract as the plain run)
    pl.set_solver(2)
    pl.step(2)
    n0 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    t0 = _t.perf_counter()
    pl.step(args.steps)
    el = _t.perf_counter() - t0
    n1 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    ph = pl.phase_times()
    ks = pl.kernel_stats_full()
    fb = pl.learn_info()["fallbacks"]
    pl.set_solver(1)
    return {"pairs": info["dataset"], "train_steps": args.train_steps, "train_s": t_train,
            "loss_first_batch": l0, "loss_last_batch": l1, "solve": cmp_,
            "iteration_ratio": cmp_["plain"]["iterations"] / max(cmp_["learned"]["iterations"], 1),
            "speedup_per_solve": cmp_["plain"]["ms_per_solve"] / cmp_["learned"]["ms_per_solve"],
            "step_value": 0.5 * (n0 + n1) * args.steps / el, "step_ms": 1e3 * el / args.steps,
            "pcg_ms_per_solve": ph[3] / args.steps, "fcg_ms": ks["fcg_ms"], "fcg_iterations": ks["fcg_iterations"],
            "fallbacks": fb}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--harvest-steps", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=400)
    ap.add_argument("--solve-reps", type=int, default=5)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy 5 times (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": x.numel() * 4, "launches": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); the sharded design is "
              "described in DESIGN.md", file=sys.stderr)
        return 2

    import __graft_entry__ as ge
    pkg = ge.load_package()
    pl = pkg.PicLibrary(0)
    log(f"bench: setting up {args.workload}")
    t0 = time.perf_counter()
    w = pkg.build_workload(pl, args.workload, seed=2024)
    pl.initial_field()
    log(f"bench: set-up {time.perf_counter() - t0:.1f} s; warmup {args.warmup} steps")
    pl.step(args.warmup)
    n0 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    log(f"bench: timing {args.steps} steps, {n0} particles")
    t0 = time.perf_counter()
    pl.step(args.steps)
    elapsed = time.perf_counter() - t0
    phase = pl.phase_times()          # device ms summed over the timed steps
    n1 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    pushes = 0.5 * (n0 + n1) * args.steps
    value = pushes / elapsed
    solve = pl.last_solve()
    # plain PCG ms/solve at 1e-8 on the step's RHS (device time of the PCG phase)
    pcg_ms = phase[3] / args.steps
    ks = pl.kernel_stats()
    nnodes = (w["domain"].nnx) * (w["domain"].nny)
    # Rooflines of the two kernels that take the step's time, from HIP events in the timed
    # region. Algorithmic bytes per launch:
    #  push: 44 B per particle (read x, y, vx, vy, vz, cell; write x, y, vx, vy, cell_next, FP32)
    #  CG iteration: 72 B per node (read r, w, s, p, x and 8 scaled off-diagonals; write r, w, s, p, x)
    push_ms = ks["push_ms"] / max(ks["push_launches"], 1)
    push_bytes = 44.0 * (0.5 * (n0 + n1)) / 2
    cg_ms = ks["cg_ms"] / max(ks["cg_launches"], 1)
    cg_bytes = 72.0 * nnodes
    traffic = traffic_from_profiles("traffic_cfg3.md")
    roof_push = {"bound": "hbm", "achieved": push_bytes / (push_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "traffic": traffic.get("k_push_deposit"), "kernel": "k_push_deposit<float,1>",
                 "bytes_per_launch": push_bytes, "ms_per_launch": push_ms, "launches": ks["push_launches"],
                 "share_of_step": ks["push_ms"] / phase[7]}
    roof_cg = {"bound": "hbm", "achieved": cg_bytes / (cg_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "traffic": traffic.get("k_cg_iter"), "kernel": "k_cg_iter<1>",
               "bytes_per_launch": cg_bytes, "ms_per_launch": cg_ms, "launches": ks["cg_launches"],
               "share_of_step": ks["cg_ms"] / phase[7]}
    for r in (roof_push, roof_cg):
        r["frac"] = r["achieved"] / r["peak"]
    dominant, other = (roof_cg, roof_push) if ks["cg_ms"] >= ks["push_ms"] else (roof_push, roof_cg)
    out = {
        "metric": METRIC, "value": value, "unit": "particle-pushes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: 512x512 Cartesian, 3 conductors, 2x16.8M particles, "
                               "FP32 particles + FP32/FP64 PCG to 1e-8, MCC, sort/20",
                   "mesh": [w["domain"].nnx - 1, w["domain"].nny - 1], "particles_per_species": int(n0 // 2),
                   "parallelism": "single GPU"},
        "pcg_ms_per_solve": pcg_ms, "pcg_iterations": solve["iterations"], "pcg_outer": solve["outer"],
        "pcg_relres": solve["relres"],
        "phase_ms_per_step": {k: phase[i] / args.steps for i, k in enumerate(
            ["push_lookup_deposit", "deposit_finalize", "rhs", "pcg", "ppr", "mcc", "sort", "total"])},
        "roofline": dominant,
        "roofline_other": other,
    }
    if not args.no_learned:
        out["learned"] = learned_leg(pkg, pl, w, args)
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle, 1 core)")
        try:
            out["cpu_baseline"] = cpu_baseline(pkg, pl, w, args.cpu_sample)
        except Exception as e:  # the baseline is reported, not the product
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
