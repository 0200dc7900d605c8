#!/usr/bin/env python3
"""Benchmark of the MI355X IFE-PIC hot path on one GPU.

Workload (default): 512 x 512 Cartesian mesh, three immersed conductors at -1, 0.5, 0,
16.8M electrons + 16.8M Ar+ ions (71 per vacuum element per species) plus both species
injected through the left wall (InjectBeams_2D: Maxwellian flux drifting inward at 0.1 v_th,
at the plasma density), FP32 gather / push / MCC with fixed-point (int64) deposition,
mixed-precision field solve (FP32 inner, FP64 refinement) to a 1e-8 relative residual, PPR
recovery, null-collision MCC against 300 K argon at 3.2e20 m^-3, periodic cell sort. A step is
one full explicit PIC step (MoveAll x2, GetParChg_2D, IFE_Solve_2D, GetEField_SIDG_PPR,
InjectBeams_2D, MCC, sort every 20 steps) with every input resident in HBM. Prints one JSON
line.

The timed steps use the learned multigrid preconditioner (FCG inner solver, plain CG
fallback), trained on the GPU on pairs the plain solver harvests during the warm-up; the
same number of steps with the plain one-kernel CG is timed after them and reported under
"plain", with a per-solve comparison on the same RHS and warm start. --no-learned times the
plain solver only.

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


WORKLOAD_TEXT = {
    "cfg3": "512x512 Cartesian, 3 conductors, 2 species x 16.8M particles, left-wall injection (drifting "
            "Maxwellian, 0.1 v_th), FP32 particles + mixed-precision field solve to 1e-8, MCC, sort/20",
    "cfg2": "256x256 axisymmetric r-z, a conductor and a dielectric disk (eps 4, IFE interface elements), "
            "2 species x 3.9M particles, FP32 particles + mixed-precision field solve to 1e-8, MCC, sort/20",
    "cfg1": "64x64 Cartesian, one conductor, 2 species x 410k particles, FP64 particles and solve to 1e-10, MCC",
}


def traffic_from_profiles(name: str) -> dict:
    """HBM bytes per launch per kernel from the committed rocprofv3 PMC summary: lines
    `traffic_bytes_per_launch <kernel>: <bytes>`."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return {}
    return {m.group(1): float(m.group(2)) for m in
            re.finditer(r"traffic_bytes_per_launch\s+(\S+):\s*([0-9.eE+]+)", p.read_text())}


def cpu_threads() -> int:
    """Host threads for the multi-core CPU timing: the CPUs this process may run on, capped
    at 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _oracle_step(pkg, pl, w, orc, sample_particles: int, phi_prev) -> tuple:
    """One explicit step of the CPU oracle on the workload state: MoveAll + lookup +
    deposition for `sample_particles` particles per species (time scaled to the whole
    population), then RHS, MY_PCG to 1e-8 in FP64 and PPR on the full mesh. Runs with the
    oracle's current thread count. Returns (seconds per step, PCG iterations, particles)."""
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
    phi = np.asarray(phi_prev)          # warm start: the potential of the previous step, as in the step
    rp, col, val, diag = (pl.setup_array(n) for n in ("A_ROWPTR", "A_COL", "A_VAL", "A_DIAG"))
    un = pl.setup_array("UNKNOWN_NODE") - 1
    t0 = time.perf_counter()
    b = orc.rhs(xs, ys, 0, nt, pl.setup_array("DIRICHLET_PHI"), rho, len(diag))
    xsol, _, it, _ = orc.pcg(rp, col, val, diag, b, phi[un], 1e-8, 50000)
    phi2 = phi.copy()
    phi2[un] = xsol
    orc.ppr(xs, ys, pl.setup_array("POINT_PATCH"), phi2)
    t_field = time.perf_counter() - t0
    return t_part + t_field, it, n_total


def cpu_baseline(pkg, pl, w, sample_particles: int, phi_prev) -> dict:
    """The CPU oracle's full explicit step (C, FP64) on a bounded sample of the same
    workload state, timed twice: on one core in the reference's serial order, and on every
    host core of this process (OpenMP over particles, nodes and rows, at most 16 threads).
    The all-core figure is the reported baseline; the one-core figure rides along. Both are
    scaled to particle-pushes/s of the whole workload."""
    if w["domain"].delta != 0:
        return {"value": None, "note": "the oracle step timed here is the Cartesian one"}
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    nthr = cpu_threads()
    orc.set_threads(1)
    t1, it1, n_total = _oracle_step(pkg, pl, w, orc, sample_particles, phi_prev)
    try:
        orc.set_threads(nthr)
        tn, itn, _ = _oracle_step(pkg, pl, w, orc, sample_particles, phi_prev)
    finally:
        orc.set_threads(1)
    return {"value": n_total / tn, "unit": "particle-pushes/s", "cores": nthr, "kind": "port",
            "nproc": os.cpu_count(),
            "single_core": {"value": n_total / t1, "cores": 1, "s_per_step": t1, "pcg_iterations": it1},
            "sample": f"oracle.c FP64 step on {nthr} OpenMP threads ({tn:.2f} s per step; 1 core: {t1:.2f} s): "
                      f"mover+lookup+deposit timed on {sample_particles} particles per species and scaled to "
                      f"{n_total}; RHS + MY_PCG (1e-8, {itn} it) + PPR on the full mesh; MCC and the wall "
                      f"injection (~1e3 new particles per step) not included"}


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


def timed_steps(pl, steps: int) -> dict:
    """K explicit steps with every input resident in HBM; device phase times and kernel
    statistics from HIP events on the library's stream."""
    n0 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    t0 = time.perf_counter()
    pl.step(steps)
    elapsed = time.perf_counter() - t0
    n1 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    phase = pl.phase_times()
    ks = pl.kernel_stats_full()
    solve = pl.last_solve()
    return {"n0": n0, "n1": n1, "elapsed": elapsed, "value": 0.5 * (n0 + n1) * steps / elapsed,
            "ms_per_step": 1e3 * elapsed / steps, "phase": phase, "ks": ks, "solve": solve}


def rooflines(run: dict, steps: int, nnodes: int, traffic: dict, delta: int = 0) -> tuple:
    """Rooflines of the push kernel and of the plain CG iteration kernel (the two single
    kernels that can dominate a step). Algorithmic bytes per launch:
      push: 44 B per particle (read x, y, vx, vy, vz, cell; write x, y, vx, vy, cell_next, FP32)
      CG iteration: 72 B per node (read r, w, s, p, x and 8 scaled off-diagonals; write r, w, s, p, x)"""
    ks, phase = run["ks"], run["phase"]
    out = []
    if ks["push_launches"]:
        ms = ks["push_ms"] / ks["push_launches"]
        b = 44.0 * (0.5 * (run["n0"] + run["n1"])) / 2
        out.append((ks["push_ms"], {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "traffic": traffic.get("k_push_deposit"),
                                    # r-z: every slot moves through the general mover
                                    "kernel": "k_push_walls<float>" if delta == 1 else "k_push_deposit<float,1>",
                                    "bytes_per_launch": b, "ms_per_launch": ms,
                                    "launches": ks["push_launches"], "share_of_step": ks["push_ms"] / phase[7]}))
    if ks["cg_launches"]:
        ms = ks["cg_ms"] / ks["cg_launches"]
        b = 72.0 * nnodes
        out.append((ks["cg_ms"], {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "traffic": traffic.get("k_cg_iter"), "kernel": "k_cg_iter<1>",
                                  "bytes_per_launch": b, "ms_per_launch": ms, "launches": ks["cg_launches"],
                                  "share_of_step": ks["cg_ms"] / phase[7]}))
    for _, r in out:
        r["frac"] = r["achieved"] / r["peak"]
    out.sort(key=lambda t: -t[0])
    return [r for _, r in out]


PHASES = ["push_lookup_deposit", "deposit_finalize", "rhs", "pcg", "ppr", "mcc", "sort", "total"]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
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
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); a run over several "
              "GPUs would shard the mesh into bands of element columns, particles following their elements", file=sys.stderr)
        return 2

    import __graft_entry__ as ge
    pkg = ge.load_package()
    pl = pkg.PicLibrary(0)
    log(f"bench: setting up {args.workload}")
    t0 = time.perf_counter()
    w = pkg.build_workload(pl, args.workload, seed=2024)
    # the time loop's periodic cell sort, once at set-up: particles start in tile order and
    # the sort workspace exists before any timed step
    pl.sort_particles()
    pl.initial_field()
    dom = w["domain"]
    nnodes = dom.nnx * dom.nny
    # measured HBM traffic exists for the 512^2 workload only
    traffic = traffic_from_profiles("traffic_cfg3.md") if args.workload == "cfg3" else {}
    learned = not args.no_learned
    log(f"bench: set-up {time.perf_counter() - t0:.1f} s; warmup {args.warmup} steps (plain solver)")
    if learned:
        # the network is trained on pairs the plain solver harvests during the untimed warm-up
        pl.learn_setup(12, (16, 16), seed=7)
        pl.learn_harvest(True, 512, 1)
    pl.step(args.warmup)
    info = {}
    if learned:
        pl.learn_harvest(False, 512, 1)
        info = pl.learn_info()
        log(f"bench: {info['dataset']} harvested pairs; training {args.train_steps} steps on the GPU")
        t0 = time.perf_counter()
        l0, l1 = pl.learn_train(args.train_steps, batch=4, lr=2e-3, seed=11)
        t_train = time.perf_counter() - t0
        pl.set_solver(2)
        pl.step(2)                       # two more untimed steps, with the learned solver
    log(f"bench: timing {args.steps} steps ({'learned' if learned else 'plain'} solver)")
    main_run = timed_steps(pl, args.steps)
    fallbacks = pl.learn_info()["fallbacks"] if learned else 0
    plain_run = main_run
    if learned:
        pl.set_solver(1)
        pl.step(1)
        log(f"bench: timing {args.steps} steps (plain solver)")
        plain_run = timed_steps(pl, args.steps)
    roofs = rooflines(main_run, args.steps, nnodes, traffic, dom.delta)
    roofs_plain = rooflines(plain_run, args.steps, nnodes, traffic, dom.delta)
    other = [r for r in roofs_plain if r["kernel"] != roofs[0]["kernel"]]
    ph = main_run["phase"]
    out = {
        "metric": METRIC, "value": main_run["value"], "unit": "particle-pushes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup + (2 if learned else 0), "ms_per_step": main_run["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: " + WORKLOAD_TEXT.get(args.workload, "see pic-_amd workload()"),
                   "mesh": [dom.nnx - 1, dom.nny - 1], "particles_per_species": int(main_run["n0"] // 2),
                   "solver": "learned multigrid FCG (plain CG fallback)" if learned else "plain one-kernel CG",
                   "parallelism": "single GPU"},
        "pcg_ms_per_solve": ph[3] / args.steps, "pcg_iterations": main_run["solve"]["iterations"],
        "pcg_outer": main_run["solve"]["outer"], "pcg_relres": main_run["solve"]["relres"],
        "phase_ms_per_step": {k: ph[i] / args.steps for i, k in enumerate(PHASES)},
        "roofline": roofs[0],
        "roofline_other": other[0] if other else None,
    }
    if learned:
        cmp_ = solve_compare(pl, args.solve_reps)
        log(f"bench: solve comparison {cmp_}")
        out["plain"] = {"value": plain_run["value"], "ms_per_step": plain_run["ms_per_step"],
                        "pcg_ms_per_solve": plain_run["phase"][3] / args.steps,
                        "pcg_iterations": plain_run["solve"]["iterations"], "pcg_relres": plain_run["solve"]["relres"],
                        "phase_ms_per_step": {k: plain_run["phase"][i] / args.steps for i, k in enumerate(PHASES)}}
        out["learned"] = {"pairs": info["dataset"], "train_steps": args.train_steps, "train_s": t_train,
                          "loss_first_batch": l0, "loss_last_batch": l1, "fallbacks": fallbacks,
                          "fcg_ms_per_iteration": main_run["ks"]["fcg_ms"] / max(main_run["ks"]["fcg_iterations"], 1),
                          "solve_same_rhs": cmp_,
                          "iteration_ratio": cmp_["plain"]["iterations"] / max(cmp_["learned"]["iterations"], 1),
                          "speedup_per_solve": cmp_["plain"]["ms_per_solve"] / cmp_["learned"]["ms_per_solve"],
                          "step_speedup": plain_run["ms_per_step"] / main_run["ms_per_step"]}
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle, 1 core)")
        try:
            pl.set_solver(1)
            phi_prev = pl.field("PHI")
            pl.step(1)                   # the state the CPU step is timed on; its solve starts from phi_prev
            out["cpu_baseline"] = cpu_baseline(pkg, pl, w, args.cpu_sample, phi_prev)
        except Exception as e:  # the baseline is reported, not the product
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
