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
plain solver only. --workload cfg5 (1024^2, 2 x 33.1M particles) also fine-tunes the network
online inside the step every 500 steps (harvesting solve on the plain CG, then Adam steps on
the GPU): one such step is timed apart and its cost amortised over the period.

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
# the product's learned V-cycle: learned smoothers on the finest levels, minibatch share of
# harvested pairs
PRODUCT_LEARNED_LEVELS = 1
PRODUCT_SHARE = 0.25


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOAD_TEXT = {
    "cfg3": "512x512 Cartesian, 3 conductors, 2 species x 16.8M particles, left-wall injection (drifting "
            "Maxwellian, 0.1 v_th), FP32 particles + mixed-precision field solve to 1e-8, MCC, sort/20",
    "cfg2": "256x256 axisymmetric r-z, a conductor and a dielectric disk (eps 4, IFE interface elements), "
            "2 species x 3.9M particles, FP32 particles + mixed-precision field solve to 1e-8, MCC, sort/20",
    "cfg1": "64x64 Cartesian, one conductor, 2 species x 410k particles, FP64 particles and solve to 1e-10, MCC",
    "cfg5": "1024x1024 Cartesian, the 3 conductors of the 512^2 layout scaled, 2 species x 33.1M particles, "
            "left-wall injection, FP32 particles + learned-multigrid FCG to 1e-8 fine-tuned online every 500 "
            "steps on plain-solver pairs, MCC, sort/20",
}


def traffic_from_profiles(name: str) -> dict:
    """HBM bytes per launch per kernel from the committed rocprofv3 PMC summary: lines
    `traffic_bytes_per_launch <kernel>: <bytes>`."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return {}
    return {m.group(1): float(m.group(2)) for m in
            re.finditer(r"traffic_bytes_per_launch\s+(\S+):\s*([0-9.eE+]+)", p.read_text())}


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
    population), then RHS, MY_PCG to 1e-8 in FP64 and PPR on the full mesh, the wall
    injection of the workload and MCC (selection over every particle, the collisions of the
    selected ones). Runs with the oracle's current thread count. Returns (seconds per step,
    PCG iterations, particles)."""
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
    # wall injection (InjectBeams_2D) and MCC; the inputs are read from the device first (untimed)
    ns_sigma = 10001
    mcc_in = []
    for isp in (0, 1):
        q = pl.particles(isp)
        mcc_in.append((q["loc"] - 1, q["vx"], q["vy"], q["vz"], pl.mcc_table(isp, 3 if isp == 0 else 2)))
    t0 = time.perf_counter()
    for isp, sp in enumerate((pkg.ELECTRON, pkg.ARGON_ION)):
        if w.get("inject_drift_vth"):
            xm = sp["mass_kg"] / pkg.ELECTRON_MASS
            temp = w["Te"] if isp == 0 else w["Ti"]
            tmpj = temp * pkg.EVTOK / (dom.T_ref_eV * pkg.EVTOK)
            vt = np.sqrt(tmpj / xm)
            vd = [w["inject_drift_vth"] * vt, 0.0, 0.0]
            cc = orc.inject_consts(vd[0], vt, tmpj, xm, dom.xmin, dom.dt)
            weight = pkg.species_weight(dom, w["ppc"])
            expect = dom.n_ref * abs(cc[0]) * norm["v_ref"] * dom.dt * norm["time_ref"] * (dom.ymax - dom.ymin) * \
                norm["L_ref"] / weight
            tg = orc.tag(isp, orc.RNG_INJECT)
            k = orc.inject_count(expect, 1, 0, tg)
            if k > 0:
                orc.inject(64, k, 1 if isp == 0 else 0, cc, vd, dom.ymin, dom.ymax, 1, 0, tg, dom.xmin, dom.ymin,
                           1.0, 1.0, dom.nnx - 1, dom.nny - 1)
        nr, emin = (3, -3.0) if isp == 0 else (2, -1.0)
        cells, pvx, pvy, pvz, (prob, _, cr) = mcc_in[isp]
        thr = int(min(np.floor(cr * 2 ** 32), 2 ** 32 - 1))
        sel = orc.mcc_select(cells, thr, 1, 0, orc.tag(isp, orc.RNG_SELECT))
        rt, th = ([11, 12, 13], [0.0, 11.55, 15.76]) if isp == 0 else ([101, 103], [0.0, 0.0])
        for k in sel:
            orc.mcc_collide(1 if isp == 0 else 0, int(k), 1, 0, orc.tag(isp, orc.RNG_COLLIDE), norm["v_ref"],
                            sp["mass_kg"], pkg.ARGON_ION["mass_kg"], 300.0, nr, ns_sigma, emin,
                            (3.0 - emin) / (ns_sigma - 1), 1e3, prob, rt, th,
                            np.array([pvx[k], pvy[k], pvz[k]]))
    t_src = time.perf_counter() - t0
    return t_part + t_field + t_src, it, n_total


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
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "single_core": {"value": n_total / t1, "cores": 1, "s_per_step": t1, "pcg_iterations": it1},
            "sample": f"oracle.c FP64 step on {nthr} OpenMP threads ({tn:.2f} s per step; 1 core: {t1:.2f} s): "
                      f"mover+lookup+deposit timed on {sample_particles} particles per species and scaled to "
                      f"{n_total}; RHS + MY_PCG (1e-8, {itn} it) + PPR on the full mesh; wall injection and MCC "
                      f"(selection over every particle, collisions of the selected); a stated baseline, "
                      f"not a target: the GPU's roofline fractions measure the kernels"}


def solve_compare(pl, reps: int, warm: bool = False) -> dict:
    """Plain (mode 1) vs learned (mode 2) inner solver on the same RHS and warm start: the
    charge density deposited in one step and the potential of the step before it (what the
    solve inside a step sees). warm: also the learned solver with the learned warm start
    ("learned_warm"); ms_per_solve is the host wall time of the call, device_ms the solve's
    HIP-event time."""
    import time as _t
    pl.set_solver(1)
    phi0 = pl.field("PHI")
    pl.step(1)
    rho = pl.field("RHO")
    res, phis = {}, {}
    was_on = pl.warm_info()["enabled"] if warm else False
    legs = [("plain", 1, False), ("learned", 2, False)] + ([("learned_warm", 2, True)] if warm else [])
    for name, mode, wm_on in legs:
        pl.set_solver(mode)
        if warm:
            pl.warm_enable(wm_on)
        times, its, dev = [], [], []
        for _ in range(reps):
            phi = phi0.copy()
            t0 = _t.perf_counter()
            pl.IFE_Solve_2D(Rho=rho, Phi=phi)
            times.append(_t.perf_counter() - t0)
            dev.append(pl.last_solve_ms())
            info = pl.last_solve()
            its.append(info["iterations"])
        res[name] = {"ms_per_solve": 1e3 * float(np.median(times)), "device_ms": float(np.median(dev)),
                     "iterations": int(np.median(its)), "relres": info["relres"], "outer": info["outer"]}
        phis[name] = phi
    if warm:
        pl.warm_enable(was_on)
        res["phi_rel_l2_learned_warm_vs_plain"] = float(np.linalg.norm(phis["learned_warm"] - phis["plain"]) /
                                                        np.linalg.norm(phis["plain"]))
    pl.set_solver(1)
    res["phi_rel_l2_learned_vs_plain"] = float(np.linalg.norm(phis["learned"] - phis["plain"]) /
                                               np.linalg.norm(phis["plain"]))
    return res


def timed_steps(pl, steps: int) -> dict:
    """K explicit steps with every input resident in HBM; device phase times and kernel
    statistics from HIP events on the library's stream."""
    n0 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    step0 = pl.get_step()
    t0 = time.perf_counter()
    pl.step(steps)
    elapsed = time.perf_counter() - t0
    n1 = pl.counts(0)["nlive"] + pl.counts(1)["nlive"]
    phase = pl.phase_times()
    ks = pl.kernel_stats_full()
    solve = pl.last_solve()
    return {"n0": n0, "n1": n1, "step0": step0, "elapsed": elapsed, "value": 0.5 * (n0 + n1) * steps / elapsed,
            "ms_per_step": 1e3 * elapsed / steps, "phase": phase, "ks": ks, "solve": solve}


def push_bytes(delta: int) -> float:
    """Algorithmic HBM bytes per particle of the fused push (FP32 SoA state read and written once)."""
    return 56.0 if delta == 1 else 44.0


def rooflines(run: dict, steps: int, nnodes: int, traffic: dict, delta: int = 0, push_kernel: str = "") -> tuple:
    """Rooflines of the push kernel and of the plain CG iteration kernel (the two single
    kernels that can dominate a step). Algorithmic bytes per launch:
      push: 44 B per particle (read x, y, vx, vy, vz, cell; write x, y, vx, vy, cell_next, FP32);
            axisymmetric 56 B (the azimuth read and written, vz written too)
      CG iteration: 72 B per node (read r, w, s, p, x and 8 scaled off-diagonals; write r, w, s, p, x)"""
    ks, phase = run["ks"], run["phase"]
    out = []
    if ks["push_launches"]:
        ms = ks["push_ms"] / ks["push_launches"]
        b = push_bytes(delta) * (0.5 * (run["n0"] + run["n1"])) / 2
        out.append((ks["push_ms"], {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "traffic": traffic.get("k_push_deposit"),
                                    "kernel": push_kernel or "k_push_deposit<float,1,%d>" % delta,
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

# Algorithmic HBM bytes of one step, by phase (the figures behind "full_step"):
#   push + gather + lookup + deposit (fused)  44 B per particle (FP32 SoA state read and written once;
#                                             axisymmetric 56 B with the azimuth)
#   deposit finalisation                      per node: 3 int64 sums read and cleared per species, rho_s,
#                                             rho_p, Ek_tot, Ek_s written per species, rho and the volume
#   RHS (B rho - c)                           per unknown: 9 B entries (12 B) + rho + c + two rhs writes
#   PPR                                       per node: phi + class id + 6 coefficients + 5 FP64 and 5 FP32
#                                             gather copies + efx, efy in FP64 and FP32
#   MCC                                       the selection reads the cell of every slot (4 B) and writes its
#                                             mask bit; a collision reads and writes a particle (24 B)
#   cell sort (every sort_every steps)        52 B per particle (keys, gather of the slot arrays), amortised
#   solve, per iteration of the inner solver  plain one-kernel CG: 72 B per node; learned FCG: 240 B per node
#                                             (FCG vectors 88 B, V-cycle 152 B over its levels), plus 96 B per
#                                             node per FP64 refinement residual
FIN_B_NODE = 2 * (24 + 24 + 32) + 16
RHS_B_NODE = 9 * 12 + 32
PPR_B_NODE = 8 + 4 + 48 + 60 + 24
SORT_B_PART = 52.0
CG_B_NODE, FCG_B_NODE, RES64_B_NODE = 72.0, 240.0, 96.0


def full_step_roofline(run: dict, steps: int, nnodes: int, nunk: int, learned: bool, sort_every: int,
                       delta: int = 0) -> dict:
    """Summed algorithmic bytes per step / measured time per step, against the 8 TB/s HBM peak."""
    npart = 0.5 * (run["n0"] + run["n1"])
    ks, solve = run["ks"], run["solve"]
    if learned:
        iters = ks["fcg_iterations"] / steps
        solve_b = iters * FCG_B_NODE * nnodes + (solve["outer"] + 1) * RES64_B_NODE * nnodes
    else:
        iters = ks["cg_launches"] / steps
        solve_b = iters * CG_B_NODE * nnodes + (solve["outer"] + 1) * RES64_B_NODE * nnodes
    by = {"push_lookup_deposit": push_bytes(delta) * npart, "deposit_finalize": FIN_B_NODE * nnodes, "rhs": RHS_B_NODE * nunk,
          "pcg": solve_b, "ppr": PPR_B_NODE * nnodes, "mcc": 4.125 * npart,
          "sort": (SORT_B_PART * npart / sort_every) if sort_every else 0.0}
    total = sum(by.values())
    t = run["ms_per_step"] * 1e-3
    ach = total / t / 1e9
    return {"bound": "hbm", "bytes_per_step": total, "ms_per_step": run["ms_per_step"], "achieved": ach,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "floor_ms": total / (HBM_PEAK_GBS * 1e9) * 1e3, "solver_iterations_per_step": iters,
            "bytes_by_phase": by}


def fp64_error(pkg, pl, w, args, learned: bool) -> dict:
    """One step of the product configuration (FP32 particles, mixed-precision or learned
    solve) against the same step in FP64 (FP64 particles, FP64 PCG to 1e-12) from identical
    inputs: the particles after a cell sort (no dead slots, so both runs draw the same random
    streams), the potential, the step counter. Errors relative to the FP64 step."""
    dom = w["domain"]
    pl.sort_particles()
    step = pl.get_step()
    parts = [pl.particles(s) for s in (0, 1)]
    th = [pl.azimuth(s) for s in (0, 1)] if dom.delta == 1 else [None, None]
    phi0 = pl.field("PHI")
    pl.set_solver(2 if learned else 1)
    pl.step(1)
    nn = dom.nnx * dom.nny
    vol = pl.setup_array("CELL_VOLUME")[1::2]
    a = dict(phi=pl.field("PHI"), rho_s=pl.field("RHO_S")[:2 * nn], ek=pl.field("EK_TOT")[:2 * nn],
             n=[pl.counts(s)["nlive"] for s in (0, 1)], relres=pl.last_solve()["relres"])
    # the same state, FP64 throughout
    pl.setup(dom, bits=64, solver=0, tol=1e-12)
    weight = pkg.species_weight(dom, w["ppc"])
    for s, (sp, temp) in enumerate(((pkg.ELECTRON, w["Te"]), (pkg.ARGON_ION, w["Ti"]))):
        q = parts[s]
        cap = int(len(q["x"]) * 1.05) + 4096
        pl.set_species(s, sp["charge_C"], sp["mass_kg"], weight, temp * pkg.EVTOK, cap)
        pl.load_particles(s, q["x"], q["y"], q["vx"], q["vy"], q["vz"], q["loc"], th=th[s])
    pl.MCCBundleInit(pkg.ARGON_ION["mass_kg"], 300.0, 3.2e20, 0, 1)
    if w.get("inject_drift_vth"):
        v_ref = dom.normalisation()["v_ref"]
        for s, (sp, temp) in enumerate(((pkg.ELECTRON, w["Te"]), (pkg.ARGON_ION, w["Ti"]))):
            vt = np.sqrt(temp / dom.T_ref_eV / (sp["mass_kg"] / pkg.ELECTRON_MASS))
            pl.set_injection(s, dom.n_ref, temp, (w["inject_drift_vth"] * vt * v_ref, 0.0, 0.0), r_start=dom.xmin,
                             y_lo=dom.ymin, y_hi=dom.ymax)
    pl.set_run_params(20, 2024, 1)
    pl.set_field("PHI", phi0)
    pl.GetEField_SIDG_PPR()
    pl.set_step(step)
    pl.step(1)
    b = dict(phi=pl.field("PHI"), rho_s=pl.field("RHO_S")[:2 * nn], ek=pl.field("EK_TOT")[:2 * nn],
             n=[pl.counts(s)["nlive"] for s in (0, 1)])

    def rel(x, y):
        return float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))

    out = {"reference": "the same step from the same state with FP64 particles and FP64 PCG to 1e-12",
           "phi_max_rel": float(np.abs(a["phi"] - b["phi"]).max() / np.abs(b["phi"]).max()),
           "phi_rel_l2": rel(a["phi"], b["phi"]), "relres_mixed": a["relres"], "live_fp32": a["n"], "live_fp64": b["n"]}
    for s, name in ((0, "electrons"), (1, "ions")):
        sl = slice(s * nn, (s + 1) * nn)
        q_a, q_b = (a["rho_s"][sl] * vol).sum(), (b["rho_s"][sl] * vol).sum()
        e_a, e_b = (a["ek"][sl] * vol).sum(), (b["ek"][sl] * vol).sum()
        out[name] = {"rho_s_rel_l2": rel(a["rho_s"][sl], b["rho_s"][sl]),
                     "rho_s_max_rel": float(np.abs(a["rho_s"][sl] - b["rho_s"][sl]).max() / np.abs(b["rho_s"][sl]).max()),
                     "ek_rel_l2": rel(a["ek"][sl], b["ek"][sl]),
                     "total_charge_rel": float(abs(q_a - q_b) / abs(q_b)),
                     "total_energy_rel": float(abs(e_a - e_b) / abs(e_b))}
    return out


def cfg4_protocol(pkg, pl, args) -> dict:
    """The learned preconditioner on the stand-alone Poisson workload of the 512^2 mesh of the
    three-conductor configuration. Held-out charge fields: every step of an independent run
    of that configuration (its own seed). Every held-out field is solved cold (zero potential
    on the unknowns, Dirichlet values on the boundary) to the same 1e-8 relative residual
    (FP64 residual) by FP64 Jacobi-PCG (MY_PCG), the mixed-precision Jacobi CG, FP64
    multicolour IC(0)-PCG, and FCG with the untrained (damped-Jacobi) and trained V-cycles.
    Training variants: 2 or 4 learned smoother levels; minibatches with a share of 1/2, 1/4
    or 0 of pairs harvested from the plain solver over a run with the product seed (the rest
    synthetic pairs of every smoothness). Device time per solve from HIP events around the
    solver."""
    import time as _t
    nheld, npairs = args.cfg4_heldout, args.cfg4_pairs
    log(f"cfg4: held-out run, {nheld} steps (seed 77)")
    w = pkg.build_workload(pl, "cfg3", seed=77)
    dom = w["domain"]
    pl.sort_particles()
    pl.initial_field()
    held = []
    for _ in range(nheld):
        pl.step(1)
        held.append(pl.field("RHO"))
    phi0 = pl.setup_array("DIRICHLET_PHI").copy()

    def solve_all(mode, ref=None):
        it, ms, rr, err, phis = [], [], [], [], []
        pl.set_solver(mode)
        for k, rho in enumerate(held):
            phi = phi0.copy()
            pl.IFE_Solve_2D(Rho=rho, Phi=phi)
            info = pl.last_solve()
            it.append(info["iterations"]); rr.append(info["relres"]); ms.append(pl.last_solve_ms())
            if ref is None:
                phis.append(phi)
            else:
                err.append(float(np.linalg.norm(phi - ref[k]) / np.linalg.norm(ref[k])))
        pl.set_solver(1)
        it, ms = np.array(it), np.array(ms)
        r = {"iterations_mean": float(it.mean()), "iterations_median": float(np.median(it)),
             "ms_per_solve_mean": float(ms.mean()), "ms_per_solve_median": float(np.median(ms)),
             "relres_max": float(np.max(rr))}
        if ref is not None:
            r["phi_rel_l2_vs_jacobi_fp64_max"] = float(np.max(err))
        return r, (phis if ref is None else None), it

    methods = {}
    log("cfg4: baselines")
    methods["jacobi_pcg_fp64"], ref, _ = solve_all(0)
    methods["jacobi_cg_mixed"], _, _ = solve_all(1, ref)
    methods["ic0_pcg_fp64"], _, _ = solve_all(3, ref)
    log(f"cfg4: {methods}")
    variants = {}
    levels_list = [int(v) for v in args.cfg4_levels.split(",") if v.strip()]
    if 0 in levels_list:
        # the classical V-cycle: damped Jacobi on every level (no learned smoother to train)
        w = pkg.build_workload(pl, "cfg3", seed=2024)
        pl.sort_particles()
        pl.initial_field()
        pl.learn_setup(12, (), seed=7)
        variants["L0_jacobi_vcycle"], _, _ = solve_all(2, ref)
        log(f"cfg4: Jacobi V-cycle {variants['L0_jacobi_vcycle']}")
    for Ls in [v for v in levels_list if v > 0]:
        w = pkg.build_workload(pl, "cfg3", seed=2024)
        pl.sort_particles()
        pl.initial_field()
        pl.learn_setup(12, (16,) * Ls, seed=7)
        p_untrained = pl.learn_params().copy()
        pl.learn_harvest(True, npairs, 1)
        t0 = _t.perf_counter()
        steps = 0
        while pl.learn_info()["dataset"] < npairs and steps < args.cfg4_max_steps:
            pl.step(1)
            steps += 1
        pl.learn_harvest(False, npairs, 1)
        t_harvest = _t.perf_counter() - t0
        ds = pl.learn_info()["dataset"]
        fb0 = pl.learn_info()["fallbacks"]
        un, _, it_un = solve_all(2, ref)
        variants[f"L{Ls}_untrained"] = un
        log(f"cfg4: {Ls} learned levels, {ds} pairs from {steps} steps ({t_harvest:.1f} s); untrained {un}")
        neval = min(ds, 256)
        for share in (0.5, 0.25, 0.0) if Ls > 1 else (0.25, 0.0):
            pl.learn_set_params(p_untrained)
            pl.learn_data_mix(share)
            l_before = pl.learn_eval(neval)
            t0 = _t.perf_counter()
            lf, ll = pl.learn_train(args.cfg4_train, batch=8, lr=5e-3, seed=11)
            t_train = _t.perf_counter() - t0
            l_after = pl.learn_eval(neval)
            tr, _, it_tr = solve_all(2, ref)
            tr.update({"harvested_share": share, "pairs": ds, "harvest_steps": steps, "train_steps": args.cfg4_train,
                       "train_batch": 8, "train_s": t_train, "loss_harvested_before": l_before,
                       "loss_harvested_after": l_after, "loss_first_batch": lf, "loss_last_batch": ll,
                       "fields_fewer_iterations_than_untrained": int((it_tr < it_un).sum()),
                       "fields_more_iterations_than_untrained": int((it_tr > it_un).sum()),
                       "iteration_ratio_untrained_over_trained": float(it_un.mean() / it_tr.mean()),
                       "ms_ratio_untrained_over_trained": un["ms_per_solve_median"] / tr["ms_per_solve_median"]})
            variants[f"L{Ls}_share{share}"] = tr
            log(f"cfg4: L{Ls} share {share}: {tr}")
            if Ls == PRODUCT_LEARNED_LEVELS and share == PRODUCT_SHARE:
                p_product = pl.learn_params().copy()   # the product's V-cycle configuration
        variants[f"L{Ls}_fallbacks"] = pl.learn_info()["fallbacks"] - fb0
    # the learned warm start on top of the product's V-cycle (2 learned levels, share 1/4): trained
    # on the initial residuals of plain cold solves of the charge of a run with the product seed
    nw = args.cfg4_warm_pairs
    log(f"cfg4: warm start: {nw} plain cold solves for training pairs")
    w = pkg.build_workload(pl, "cfg3", seed=2024)
    pl.sort_particles()
    pl.initial_field()
    pl.learn_setup(12, (16,) * PRODUCT_LEARNED_LEVELS, seed=7)
    cap = 4 * nw
    t0 = _t.perf_counter()
    for _ in range(nw):
        pl.learn_harvest(False, cap, 1000)
        pl.step(1)
        rho = pl.field("RHO")
        pl.learn_harvest(True, cap, 1000)
        phi = phi0.copy()
        pl.IFE_Solve_2D(Rho=rho, Phi=phi)
    pl.learn_harvest(False, cap, 1000)
    t_harvest = _t.perf_counter() - t0
    pl.learn_set_params(p_product)
    pl.warm_setup(6, 16, seed=3)
    ninit = pl.warm_info()["initial_pairs"]
    l0, l0s = pl.warm_eval(min(ninit, 256))
    t0 = _t.perf_counter()
    lf, ll = pl.warm_train(args.cfg4_warm_train, batch=8, lr=2e-3, seed=5)
    t_train = _t.perf_counter() - t0
    l1, l1s = pl.warm_eval(min(ninit, 256))
    base, _, it_base = solve_all(2, ref)
    pl.warm_enable(True)
    wr, _, it_w = solve_all(2, ref)
    pl.warm_enable(False)
    p_w = pl.warm_params()
    wr.update({"training_pairs": ninit, "harvest_s": t_harvest, "train_steps": args.cfg4_warm_train, "train_batch": 8,
               "train_s": t_train, "parameters": int(p_w.size), "levels": pl.warm_layout()["H"], "channels": 16,
               "loss_train_pairs_before": l0, "loss_train_pairs_after": l1,
               "loss_after_line_search_before": l0s, "loss_after_line_search_after": l1s,
               "loss_first_batch": lf, "loss_last_batch": ll,
               "fields_fewer_iterations_than_without": int((it_w < it_base).sum()),
               "fields_more_iterations_than_without": int((it_w > it_base).sum()),
               "iteration_ratio_without_over_with": float(it_base.mean() / it_w.mean()),
               "ms_ratio_without_over_with": base["ms_per_solve_median"] / wr["ms_per_solve_median"]})
    variants[f"L{PRODUCT_LEARNED_LEVELS}_share{PRODUCT_SHARE}_rerun"] = base
    variants[f"L{PRODUCT_LEARNED_LEVELS}_share{PRODUCT_SHARE}_warm_start"] = wr
    log(f"cfg4: warm start: {wr}")
    best = min((k for k in variants if "share" in k), key=lambda k: variants[k]["ms_per_solve_mean"])
    return {"workload": "512x512 three-conductor mesh, stand-alone Poisson, cold start",
            "heldout_fields": len(held), "heldout_source": "every step of an independent run (seed 77)",
            "methods": methods, "learned_variants": variants, "best_learned": best, "mesh": [dom.nnx - 1, dom.nny - 1]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--train-steps", type=int, default=1000)
    ap.add_argument("--solve-reps", type=int, default=5)
    ap.add_argument("--learn-levels", type=int, default=PRODUCT_LEARNED_LEVELS,
                    help="learned smoother levels of the V-cycle (0..4)")
    ap.add_argument("--learn-share", type=float, default=PRODUCT_SHARE,
                    help="share of harvested pairs in a training minibatch")
    ap.add_argument("--no-warm-net", action="store_true", help="do not train the learned warm start")
    ap.add_argument("--warm-start", action="store_true",
                    help="apply the learned warm start inside the timed steps (default: compared on one RHS only)")
    ap.add_argument("--warm-train-steps", type=int, default=300)
    ap.add_argument("--no-fp64-error", action="store_true", help="skip the FP64 comparison step (it ends the run)")
    ap.add_argument("--finetune-every", type=int, default=None,
                    help="online fine-tuning period in steps (default: 500 for cfg5, off otherwise)")
    ap.add_argument("--finetune-train-steps", type=int, default=50)
    ap.add_argument("--cfg4-heldout", type=int, default=1000, help="cfg4: held-out charge fields")
    ap.add_argument("--cfg4-pairs", type=int, default=20000, help="cfg4: harvested training pairs")
    ap.add_argument("--cfg4-max-steps", type=int, default=2000, help="cfg4: at most this many harvest steps")
    ap.add_argument("--cfg4-train", type=int, default=4000, help="cfg4: training steps")
    ap.add_argument("--cfg4-levels", default="0,1,2,4", help="cfg4: learned smoother levels of the variants (0: Jacobi V-cycle)")
    ap.add_argument("--cfg4-warm-pairs", type=int, default=400, help="cfg4: plain cold solves harvested for the warm start")
    ap.add_argument("--cfg4-warm-train", type=int, default=2000, help="cfg4: warm-start training steps")
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
    if args.workload == "cfg4":
        r = cfg4_protocol(pkg, pl, args)
        out = {"metric": "PCG ms/solve @1e-8 rel residual, learned vs plain, 1 GPU", "unit": "ms/solve",
               "value": r["learned_variants"][r["best_learned"]]["ms_per_solve_mean"], "higher_is_better": False, "n_gpus": 1,
               "steps": r["heldout_fields"], "warmup": 0, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
               "data": "synthetic", "config": {"workload": "cfg4: " + r["workload"], "mesh": r["mesh"],
                                                "parallelism": "single GPU"}, "cfg4": r}
        print(json.dumps(out), flush=True)
        return 0
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
        pl.learn_setup(12, (16,) * args.learn_levels, seed=7)
        pl.learn_harvest(True, 512, 1)
    n_learned_warm = min(2, max(args.warmup - 1, 0)) if learned else 0
    pl.step(args.warmup - n_learned_warm)
    info = {}
    if learned:
        pl.learn_harvest(False, 512, 1)
        info = pl.learn_info()
        log(f"bench: {info['dataset']} harvested pairs; training {args.train_steps} steps on the GPU")
        t0 = time.perf_counter()
        pl.learn_data_mix(args.learn_share)
        l0, l1 = (pl.learn_train(args.train_steps, batch=8, lr=5e-3, seed=11) if args.learn_levels > 0
                  else (0.0, 0.0))
        t_train = time.perf_counter() - t0
        warm = None
        if not args.no_warm_net:
            # the learned warm start trains on the initial residuals of the warm-up solves
            pl.warm_setup(6, 16, seed=3)
            nwi = pl.warm_info()["initial_pairs"]
            t0 = time.perf_counter()
            wl0, wl1 = pl.warm_train(args.warm_train_steps, batch=8, lr=2e-3, seed=5)
            warm = {"initial_pairs": nwi, "train_steps": args.warm_train_steps, "train_s": time.perf_counter() - t0,
                    "loss_first_batch": wl0, "loss_last_batch": wl1, "in_time_loop": bool(args.warm_start)}
            pl.warm_enable(args.warm_start)
            log(f"bench: warm start trained: {warm}")
        pl.set_solver(2)
        if n_learned_warm:
            pl.step(n_learned_warm)      # the last warm-up steps run the learned solver
    ft_every = args.finetune_every if args.finetune_every is not None else (500 if args.workload == "cfg5" else 0)
    ft_every = ft_every if learned else 0
    if ft_every:
        # fine-tuning runs inside pic_step; with a period longer than the timed window the window
        # holds ordinary steps, and one fine-tune step is timed apart below and amortised
        pl.learn_finetune(ft_every, args.finetune_train_steps, 4, 1e-3, 21)
    log(f"bench: timing {args.steps} steps ({'learned' if learned else 'plain'} solver)")
    main_run = timed_steps(pl, args.steps)
    fallbacks = pl.learn_info()["fallbacks"] if learned else 0
    ft = None
    if ft_every:
        fires = any((main_run["step0"] + k) % ft_every == 0 and main_run["step0"] + k > 0 for k in range(args.steps))
        pl.learn_finetune(1, args.finetune_train_steps, 4, 1e-3, 21)
        one = timed_steps(pl, 1)
        pl.learn_finetune(ft_every, args.finetune_train_steps, 4, 1e-3, 21)
        fi = pl.learn_finetune_info()
        ft = {"every": ft_every, "train_steps_per_finetune": args.finetune_train_steps, "finetunes": fi["count"],
              "finetune_in_timed_window": fires, "finetune_step_ms": one["ms_per_step"],
              "train_ms_total": fi["train_ms"], "last_loss": fi["last_loss"],
              "harvest_solve_iterations": one["solve"]["iterations"],
              "amortised_ms_per_step": main_run["ms_per_step"] + (one["ms_per_step"] - main_run["ms_per_step"]) / ft_every,
              "fallbacks_after": pl.learn_info()["fallbacks"]}
        ft["amortised_value"] = main_run["value"] * main_run["ms_per_step"] / ft["amortised_ms_per_step"]
        log(f"bench: fine-tune step {ft}")
        pl.learn_finetune(0)
    plain_run = main_run
    if learned:
        pl.set_solver(1)
        pl.step(1)
        log(f"bench: timing {args.steps} steps (plain solver)")
        plain_run = timed_steps(pl, args.steps)
    # the moving push instantiation <R, DO_PUSH, RZ, FAST, slots per thread>: FP32 Cartesian
    # meshes of the common layout run the specialised kernel with 8 slots per thread (the sort
    # group is then 512 slots)
    if dom.delta == 1:
        push_kernel = "k_push_deposit<float,1,1,0,4>"
    else:
        push_kernel = "k_push_deposit<float,1,0,1,8>" if pl.sort_group() == 512 else "k_push_deposit<float,1,0,0,4>"
    roofs = rooflines(main_run, args.steps, nnodes, traffic, dom.delta, push_kernel)
    roofs_plain = rooflines(plain_run, args.steps, nnodes, traffic, dom.delta, push_kernel)
    other = [r for r in roofs_plain if r["kernel"] != roofs[0]["kernel"]]
    ph = main_run["phase"]
    out = {
        "metric": METRIC, "value": main_run["value"], "unit": "particle-pushes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": main_run["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: " + WORKLOAD_TEXT.get(args.workload, "see pic-_amd workload()"),
                   "mesh": [dom.nnx - 1, dom.nny - 1], "particles_per_species": int(main_run["n0"] // 2),
                   "solver": (f"learned multigrid FCG ({args.learn_levels} learned smoother level"
                              f"{'s' if args.learn_levels != 1 else ''}, plain CG fallback)") if learned
                             else "plain one-kernel CG",
                   "parallelism": "single GPU"},
        "pcg_ms_per_solve": ph[3] / args.steps, "pcg_iterations": main_run["solve"]["iterations"],
        "pcg_outer": main_run["solve"]["outer"], "pcg_relres": main_run["solve"]["relres"],
        "phase_ms_per_step": {k: ph[i] / args.steps for i, k in enumerate(PHASES)},
        "roofline": roofs[0],
        "roofline_other": other[0] if other else None,
        "full_step": full_step_roofline(main_run, args.steps, nnodes, pl.sizes()["nunknowns"], learned, 20,
                                        dom.delta),
    }
    if learned:
        cmp_ = solve_compare(pl, args.solve_reps, warm=warm is not None)
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
        if ft:
            out["learned"]["online_finetune"] = ft
        if warm is not None:
            w_ = dict(warm)
            lw, ln_ = cmp_["learned_warm"], cmp_["learned"]
            w_.update({"iterations_with": lw["iterations"], "iterations_without": ln_["iterations"],
                       "device_ms_with": lw["device_ms"], "device_ms_without": ln_["device_ms"],
                       "applied": pl.warm_info()["applied"]})
            out["learned"]["warm_start"] = w_
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle, 1 core)")
        try:
            pl.set_solver(1)
            phi_prev = pl.field("PHI")
            pl.step(1)                   # the state the CPU step is timed on; its solve starts from phi_prev
            out["cpu_baseline"] = cpu_baseline(pkg, pl, w, args.cpu_sample, phi_prev)
        except Exception as e:  # the baseline is reported, not the product
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    if not args.no_fp64_error and w["bits"] != 64:
        log("bench: error of the product step against an FP64 step from the same state")
        out["fp64_error"] = fp64_error(pkg, pl, w, args, learned)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
