"""The workload configurations on the GPU (pic_amd.workload: cfg1, cfg3 and its scaled
geometry, cfg2), against the oracle step by step and through size-independent properties.

* cfg1 at its full size (64^2, 100 per cell and species, FP64, MCC in 3.2e20 m^-3 argon,
  PCG to 1e-10) and the cfg3 geometry (three conductors at -1, 0.5, 0, left-wall injection,
  FP32 particles, mixed-precision solve to 1e-8) at 64^2, teacher-forced: every phase of each
  step (MoveAll per species, GetParChg_2D, IFE_Solve_2D, GetEField_SIDG_PPR, InjectBeams_2D,
  MCC; Main_IFE_Test_2.f90:329-344 with MCC_jw OneStep.F90:49-67) runs on the GPU's state
  before that phase, and the oracle restatement of the phase on the same state must agree:
  particles, elements, counters and node moments bit for bit, the potential within 1e-10
  of the FP64 restatement of MY_PCG (FP64 solve) or within the mixed solve's stated 1e-6
  (relative residual 1e-8), the PPR fields within 1e-12, MCC selections and event counts
  exactly and post-collision velocities within 1e-12 (libm ulps).
* cfg3 (512^2, 2 x 16.8M particles) and cfg2 (256^2 r-z, conductor + eps-4 disk) in the
  product configuration (FP32 particles, learned-multigrid FCG inside FP64 refinement,
  wall injection, MCC, periodic cell sort, pic_step), step by step: relative residual at
  most 1e-8, live counts n(k+1) = n(k) + injected + created - deleted, the deposited charge
  equal to the particles present (sum rho_s vol L^2 n_ref = Weight qs sum W, sum W = N), and
  no fallback to the plain solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


SPECIES = None


def _species(pkg):
    return [(pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"]), (pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"])]


MCC_SPEC = {0: dict(nr=3, emin=-3.0, rtype=[11, 12, 13], thr=[0.0, 11.55, 15.76]),
            1: dict(nr=2, emin=-1.0, rtype=[101, 103], thr=[0.0, 0.0])}


class Ctx:
    pass


def _context(pl, pkg, orc, w, seed, ppc=None):
    c = Ctx()
    c.ppc = ppc or w["ppc"]
    c.w, c.dom, c.seed = w, w["domain"], seed
    c.bits = w["bits"]
    c.ft = np.float64 if c.bits == 64 else np.float32
    hp = pl.setup_array("HP")
    c.xs = hp[0::2][:: c.dom.nny].copy()
    c.ys = hp[1::2][: c.dom.nny].copy()
    c.ei, c.ni = pl.setup_array("ELEMENT_INDEX"), pl.setup_array("NODE_INDEX")
    c.nt = pl.setup_array("NODE_TYPE").reshape(-1, 2)[:, 1]
    c.dphi = pl.setup_array("DIRICHLET_PHI")
    c.un = pl.setup_array("UNKNOWN_NODE") - 1
    c.rp, c.col, c.val = pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL")
    c.diag = pl.setup_array("A_DIAG")
    c.pp = pl.setup_array("POINT_PATCH")
    c.norm = c.dom.normalisation()
    c.nn = c.dom.nnx * c.dom.nny
    c.vol = orc.cell_volume(c.xs, c.ys, c.dom.delta)
    # Cell_Volume_zwz of the library equals the oracle's restatement, bit for bit
    assert np.array_equal(pl.setup_array("CELL_VOLUME")[1::2], c.vol)
    c.sp = _species(pkg)
    c.temps = (w["Te"], w["Ti"])
    c.weight = pkg.species_weight(c.dom, c.ppc)
    c.tables = {s: pl.mcc_table(s, MCC_SPEC[s]["nr"]) for s in (0, 1)}
    c.inj = {}
    return c


def _mover(orc, pkg, c, s, p, efx, efy, coef, step):
    chg, mass = c.sp[s]
    qm = (chg / pkg.ELECTRON_CHARGE) / (mass / pkg.ELECTRON_MASS)
    return orc.move_all(c.bits, c.xs, c.ys, c.ei, c.ni, c.dom.object_array(), efx.astype(c.ft), efy.astype(c.ft),
                        coef[:, 1:].astype(c.ft).ravel(), qm, c.dom.dt, c.norm["v_ref"], mass,
                        c.temps[s] * pkg.EVTOK, c.seed, step, orc.tag(s, orc.RNG_EMIT), p["x"], p["y"], p["vx"],
                        p["vy"], p["vz"], p["loc"] - 1)


def forced_step(pl, pkg, orc, c, solve_bound):
    """One teacher-forced step; returns a short summary."""
    step = pl.get_step()
    pre = [pl.particles(s) for s in (0, 1)]
    cnt0 = [pl.counters(s) for s in (0, 1)]
    efx, efy = pl.field("EFX"), pl.field("EFY")
    coef = pl.field("COEF_PPR").reshape(-1, 6)
    phi_prev = pl.field("PHI")
    ref = [_mover(orc, pkg, c, s, pre[s], efx, efy, coef, step) for s in (0, 1)]
    # ---- MoveAll per species, GetParChg_2D
    pl.MoveAll(0)
    pl.MoveAll(1)
    pl.GetParChg_2D()
    rho_or = np.zeros(c.nn)
    rho_s, rho_p, ek = pl.field("RHO_S"), pl.field("RHO_P"), pl.field("EK_TOT")
    moved = []
    for s in (0, 1):
        x, y, vx, vy, vz, alive, cnt, err = ref[s]
        assert err == 0
        keep = alive.astype(bool)
        p = pl.particles(s)
        assert len(p["x"]) == keep.sum(), (step, s)
        for k, a in zip(("x", "y", "vx", "vy", "vz"), (x, y, vx, vy, vz)):
            assert np.array_equal(p[k], a[keep].astype(np.float64)), (step, s, k)
        c1 = pl.counters(s)
        got = [c1[k] - cnt0[s][k] for k in ("emit", "right", "object", "reflect", "deleted")]
        assert got == [int(v) for v in (cnt[0], cnt[1], cnt[2], cnt[3], cnt[7])], (step, s)
        loc, bad = orc.locate(c.xs, c.ys, p["x"], p["y"])
        assert bad == 0 and np.array_equal(p["loc"] - 1, loc), (step, s)
        chg, mass = c.sp[s]
        wq, eq = orc.deposit(c.bits, c.xs, c.ys, c.dom.delta, p["x"].astype(c.ft), p["y"].astype(c.ft),
                             p["vx"].astype(c.ft), p["vy"].astype(c.ft), p["vz"].astype(c.ft), loc, mass,
                             c.norm["v_ref"], fix=True)
        rs, e = orc.normalize(c.dom.delta, c.weight, chg / pkg.ELECTRON_CHARGE, c.norm["L_ref"], c.dom.n_ref, c.vol,
                              wq, eq, rho_or)
        sl = slice(s * c.nn, (s + 1) * c.nn)
        assert np.array_equal(rho_p[sl], wq) and np.array_equal(rho_s[sl], rs) and np.array_equal(ek[sl], e), (step, s)
        moved.append(int(keep.sum()))
    rho = pl.field("RHO")
    assert np.array_equal(rho, rho_or)
    # ---- IFE_Solve_2D: the oracle's element-loop RHS and MY_PCG from the same warm start
    pl.IFE_Solve_2D()
    phi = pl.field("PHI")
    info = pl.last_solve()
    b = orc.rhs(c.xs, c.ys, c.dom.delta, c.nt, c.dphi, rho, len(c.un))
    tol_or = c.w["tol"] if solve_bound <= 1e-10 else 1e-13
    x, r, it_or, ierr = orc.pcg(c.rp, c.col, c.val, c.diag, b, phi_prev[c.un], tol_or, 50000)
    assert ierr == 0
    rel = np.linalg.norm(phi[c.un] - x) / np.linalg.norm(x)
    assert rel < solve_bound, (step, rel)
    assert info["relres"] <= c.w["tol"]
    if solve_bound <= 1e-10:
        assert abs(info["iterations"] - it_or) <= 1, (info, it_or)
    assert np.array_equal(phi[c.nt <= 0], c.dphi[c.nt <= 0])
    # ---- GetEField_SIDG_PPR on the GPU's potential
    pl.GetEField_SIDG_PPR()
    coef_or, efx_or, efy_or, _ = orc.ppr(c.xs, c.ys, c.pp, phi)
    scale = np.abs(coef_or).max()
    assert np.abs(pl.field("COEF_PPR") - coef_or).max() <= 1e-12 * scale
    assert np.abs(pl.field("EFX") - efx_or).max() <= 1e-12 * scale
    assert np.abs(pl.field("EFY") - efy_or).max() <= 1e-12 * scale
    # ---- InjectBeams_2D (it = the step counter, as inside pic_step)
    added = [0, 0]
    if c.inj:
        nbefore = [pl.counts(s)["nslots"] for s in (0, 1)]
        added = pl.InjectBeams_2D(step)[:2]
        for s in (0, 1):
            cc, vd, expect, y_lo, y_hi = c.inj[s]
            tg = orc.tag(s, orc.RNG_INJECT)
            assert added[s] == orc.inject_count(expect, c.seed, step, tg), (step, s)
            if added[s] == 0:
                continue
            xi, yi, vxi, vyi, vzi, ci = orc.inject(c.bits, added[s], 1 if s == 0 else 0, cc, vd, y_lo, y_hi, c.seed,
                                                   step, tg, c.dom.xmin, c.dom.ymin, 1.0, 1.0, c.dom.nnx - 1,
                                                   c.dom.nny - 1)
            p = pl.particles(s)
            n = added[s]
            assert pl.counts(s)["nslots"] == nbefore[s] + n
            for k, a in (("x", xi), ("y", yi), ("vx", vxi), ("vy", vyi), ("vz", vzi)):
                assert np.array_equal(p[k][-n:], a), (step, s, k)
            assert np.array_equal(p["loc"][-n:] - 1, ci)
    # ---- cell sort (a permutation that compacts the dead slots), then MCC on that state
    pl.sort_particles()
    before = [pl.particles(s) for s in (0, 1)]
    for s in (0, 1):
        assert pl.counts(s)["nslots"] == len(before[s]["x"])
    sel0 = [pl.counters(s) for s in (0, 1)]
    ev0 = [pl.mcc_events(s, MCC_SPEC[s]["nr"]) for s in (0, 1)]
    pl.MCC()
    assert pl.get_step() == step + 1
    after = [pl.particles(s) for s in (0, 1)]
    new_e, new_i, prim = [], [], []
    nsel = [0, 0]
    for s in (0, 1):
        sp = MCC_SPEC[s]
        prob, smax, cr = c.tables[s]
        thr = int(min(np.floor(cr * 2 ** 32), 2 ** 32 - 1))
        bb = before[s]
        sel = orc.mcc_select(bb["loc"] - 1, thr, c.seed, step, orc.tag(s, orc.RNG_SELECT))
        nsel[s] = len(sel)
        assert pl.counters(s)["mcc_selected"] - sel0[s]["mcc_selected"] == len(sel), (step, s)
        ev = np.zeros(sp["nr"], dtype=np.int64)
        v_exp = np.stack([bb["vx"], bb["vy"], bb["vz"]], 1).copy()
        for k in sel:
            idx, v, ne, ni = orc.mcc_collide(1 if s == 0 else 0, int(k), c.seed, step, orc.tag(s, orc.RNG_COLLIDE),
                                             c.norm["v_ref"], c.sp[s][1], c.sp[1][1], 300.0, sp["nr"], 10001,
                                             sp["emin"], (3.0 - sp["emin"]) / 10000, 1e3, prob, sp["rtype"], sp["thr"],
                                             v_exp[k])
            v_exp[k] = v
            if idx:
                ev[idx - 1] += 1
            if s == 0 and idx == 3:
                new_e.append(ne)
                new_i.append(ni)
                prim.append(k)
        assert np.array_equal(pl.mcc_events(s, sp["nr"]) - ev0[s], ev), (step, s)
        n0 = len(bb["x"])
        a = after[s]
        v_got = np.stack([a["vx"], a["vy"], a["vz"]], 1)[:n0]
        if c.bits == 64:
            np.testing.assert_allclose(v_got, v_exp, rtol=1e-12, atol=1e-15)
        else:
            np.testing.assert_allclose(v_got, v_exp.astype(np.float32).astype(np.float64), rtol=1e-6, atol=1e-12)
        assert np.array_equal(a["x"][:n0], bb["x"]) and np.array_equal(a["loc"][:n0], bb["loc"])
    for s, new in ((0, new_e), (1, new_i)):
        a = after[s]
        n0 = len(before[s]["x"])
        assert len(a["x"]) - n0 == len(new), (step, s)
        if new:
            rt = 1e-12 if c.bits == 64 else 1e-6
            np.testing.assert_allclose(np.stack([a["vx"], a["vy"], a["vz"]], 1)[n0:], np.array(new), rtol=rt,
                                       atol=1e-12)
            # the products sit where their primary electron was
            assert np.array_equal(a["x"][n0:], before[0]["x"][prim]) and np.array_equal(a["loc"][n0:],
                                                                                         before[0]["loc"][prim])
    return dict(step=step, live=[len(a["x"]) for a in after], moved=moved, iters=info["iterations"], solve_rel=rel,
                injected=added, selected=nsel, ionisations=len(new_e))


def test_cfg1_full_size_teacher_forced(gpu_lib, pkg, orc):
    """cfg1 (64^2, 100 ppc, FP64, MCC at 3.2e20 m^-3, PCG 1e-10): 4 steps, every phase against
    the oracle on the GPU's own state."""
    pl = gpu_lib
    say("case: cfg1 full size (64^2, 100 ppc, FP64), 4 teacher-forced steps")
    w = pkg.build_workload(pl, "cfg1", seed=1001, mcc=True)
    pl.set_run_params(0, 1001, 1)
    pl.set_step(0)
    pl.initial_field()
    c = _context(pl, pkg, orc, w, 1001)
    n = [pl.counts(s)["nlive"] for s in (0, 1)]
    assert 370000 < n[0] < 410000 and 370000 < n[1] < 410000
    for _ in range(4):
        say(f"  {forced_step(pl, pkg, orc, c, 1e-10)}")


def test_cfg3_geometry_teacher_forced(gpu_lib, pkg, orc):
    """The cfg3 geometry (three conductors at -1, 0.5, 0; injection of both species through the
    left wall) on 64^2 in the product precision (FP32 particles, mixed-precision solve to 1e-8),
    dense gas so that every MCC reaction occurs: 3 steps, every phase against the oracle."""
    pl = gpu_lib
    say("case: cfg3 geometry at 64^2 (FP32, mixed solve, injection, MCC), 3 teacher-forced steps")
    w = pkg.build_workload(pl, "cfg3s64", seed=3003, ppc=40, mcc=True, gas_density=2.0e24)
    dom = w["domain"]
    pl.set_run_params(0, 3003, 1)
    pl.set_step(0)
    pl.initial_field()
    c = _context(pl, pkg, orc, w, 3003, ppc=40)
    # the injection sources as build_workload registered them (pic.inp semantics, InjectBeams_2D):
    # drift given in m/s, normalised by the library (Input_2D.f90:489-494)
    norm = c.norm
    for s, (chg, mass) in enumerate(c.sp):
        xm = mass / pkg.ELECTRON_MASS
        tmpj = c.temps[s] * pkg.EVTOK / (dom.T_ref_eV * pkg.EVTOK)
        vt = np.sqrt(tmpj / xm)
        drift_ms = w["inject_drift_vth"] * np.sqrt(c.temps[s] / dom.T_ref_eV / xm) * norm["v_ref"]
        vd = [drift_ms / norm["v_ref"], 0.0, 0.0]
        cc = orc.inject_consts(vd[0], vt, tmpj, xm, dom.xmin, dom.dt)
        expect = 1.0 * dom.n_ref * abs(cc[0]) * norm["v_ref"] * dom.dt * norm["time_ref"] * (dom.ymax - dom.ymin) * \
            norm["L_ref"] / c.weight
        c.inj[s] = (cc, vd, expect, dom.ymin, dom.ymax)
    total = dict(injected=0, ionisations=0)
    for _ in range(3):
        r = forced_step(pl, pkg, orc, c, 1e-6)
        total["injected"] += sum(r["injected"])
        total["ionisations"] += r["ionisations"]
        say(f"  {r}")
    assert total["injected"] > 100 and total["ionisations"] > 0, total


def _product_run(pl, pkg, name, steps, train_steps):
    w = pkg.build_workload(pl, name, seed=2024)
    dom = w["domain"]
    pl.set_step(0)
    pl.initial_field()
    pl.learn_setup(12, (16,), 1)          # the product's V-cycle: one learned level
    pl.learn_harvest(True, 256, 1)
    pl.step(4)                               # plain solver: harvest pairs
    pl.learn_train(train_steps, 4, 1e-3, 1)
    pl.learn_harvest(False)
    pl.set_solver(2)
    vol = pl.setup_array("CELL_VOLUME")[1::2]
    norm = dom.normalisation()
    scale = norm["L_ref"] ** (2 if dom.delta == 0 else 3) * dom.n_ref
    weight = pkg.species_weight(dom, w["ppc"])
    qs = [pkg.ELECTRON["charge_C"] / pkg.ELECTRON_CHARGE, pkg.ARGON_ION["charge_C"] / pkg.ELECTRON_CHARGE]
    nn = dom.nnx * dom.nny
    fb0 = pl.learn_info()["fallbacks"]
    for k in range(steps):
        n0 = [pl.counts(s)["nlive"] for s in (0, 1)]
        c0 = [pl.counters(s) for s in (0, 1)]
        pl.step(1)
        info = pl.last_solve()
        assert info["relres"] <= w["tol"], (k, info)
        c1 = [pl.counters(s) for s in (0, 1)]
        d = [{key: c1[s][key] - c0[s][key] for key in c1[s]} for s in (0, 1)]
        pairs = d[0]["mcc_created"] // 2
        rho_s, rho_p = pl.field("RHO_S"), pl.field("RHO_P")
        for s in (0, 1):
            n1 = pl.counts(s)["nlive"]
            assert n1 == n0[s] + d[s]["injected"] + pairs - d[s]["deleted"], (k, s, n0[s], n1, d[s], pairs)
            sl = slice(s * nn, (s + 1) * nn)
            # the deposit counts the particles present after the move (before injection and MCC)
            wsum = rho_p[sl].sum()
            assert (rho_s[sl] * vol * scale).sum() == pytest.approx(weight * qs[s] * wsum, rel=1e-12)
            if dom.delta == 0:
                assert wsum == pytest.approx(n0[s] - d[s]["deleted"], rel=1e-12)
            else:
                assert 0.75 * (n0[s] - d[s]["deleted"]) <= wsum <= (n0[s] - d[s]["deleted"]) * (1 + 1e-12)
        if k % 3 == 0:
            say(f"  step {k}: live {[pl.counts(s)['nlive'] for s in (0, 1)]}, iterations {info['iterations']}, "
                f"relres {info['relres']:.2e}, injected {[d[s]['injected'] for s in (0, 1)]}, pairs {pairs}, "
                f"deleted {[d[s]['deleted'] for s in (0, 1)]}")
    assert pl.learn_info()["fallbacks"] == fb0
    return w


def test_cfg3_product_configuration_properties(gpu_lib, pkg):
    """cfg3 at full size (512^2, 2 x 16.8M particles), the product configuration, 12 steps
    (one cell sort at step 20 of the counter is not reached; the sort is exercised below)."""
    pl = gpu_lib
    say("case: cfg3 full size, product configuration, 12 steps of properties")
    _product_run(pl, pkg, "cfg3", 12, 200)
    pl.sort_particles()
    n = [pl.counts(s) for s in (0, 1)]
    assert all(c["nlive"] == c["nslots"] for c in n)
    assert n[0]["nlive"] > 16_000_000 and n[1]["nlive"] > 16_000_000


def test_online_finetune_inside_step(gpu_lib, pkg):
    """Online fine-tuning on the 128^2 cfg3 geometry: every 3rd step's solve runs the plain CG
    and harvests pairs, the network then trains 30 Adam steps; every step still meets the solve
    tolerance, no fallback, and the count, time and loss are reported."""
    pl = gpu_lib
    say("case: online fine-tune, cfg3s128, every 3 steps")
    w = pkg.build_workload(pl, "cfg3s128", seed=2024)
    pl.set_step(0)
    pl.initial_field()
    pl.learn_setup(12, (16,), 1)          # the product's V-cycle: one learned level
    pl.learn_harvest(True, 256, 1)
    pl.step(2)
    pl.learn_train(60, 4, 1e-3, 1)
    pl.learn_harvest(False)
    pl.set_solver(2)
    fb0 = pl.learn_info()["fallbacks"]
    pl.learn_finetune(3, 30, 4, 1e-3, 7)
    its = []
    for k in range(7):                       # counter 2..8 before each step: fine-tunes at 3 and 6
        pl.step(1)
        info = pl.last_solve()
        assert info["relres"] <= w["tol"], (k, info)
        its.append(info["iterations"])
    ft = pl.learn_finetune_info()
    pl.learn_finetune(0)
    pl.set_solver(1)
    say(f"  iterations {its}, fine-tune {ft}")
    assert ft["count"] == 2
    assert ft["train_ms"] > 0 and np.isfinite(ft["last_loss"]) and ft["last_loss"] > 0
    assert pl.learn_info()["fallbacks"] == fb0
    # the two fine-tune solves ran the plain inner CG: many more iterations than the learned ones
    assert its[1] > 3 * its[0] and its[4] > 3 * its[3], its


def _sigma_workload(E, electron):
    """The workloads' synthetic argon cross sections (m^2, E in eV; thresholds 11.55 and 15.76 eV),
    the gas the configurations define: electron elastic, excitation, ionisation; ion isotropic
    elastic and charge exchange."""
    E = np.asarray(E, dtype=float)
    if electron:
        el = 5.0e-21 + 1.5e-19 * np.sqrt(E / (E + 1.0)) * 12.0 / (E + 12.0)
        ex = np.where(E > 11.55, 1.2e-17 * np.log(np.maximum(E, 11.55) / 11.55) / (E * 11.55), 0.0)
        io = np.where(E > 15.76, 2.7e-17 * np.log(np.maximum(E, 15.76) / 15.76) / (E * 15.76), 0.0)
        return np.stack([el, ex, io], 1)
    iso = 1.5e-19 + 3.5e-19 / np.sqrt(1.0 + E)
    cx = (7.5e-10 - 0.4e-10 * np.log(E)) ** 2
    return np.stack([iso, cx], 1)


def test_mcc_tables_match_oracle(gpu_lib, pkg, orc):
    """MCCBundleInit (MCCInit.F90:14-152, SigmaNormalizedUpdate MCCSigma.f90:122-181): the
    library's cumulative, normalised probability tables on the 10001-point log grids, nu_max
    and CollisionRatio = 1 - exp(-nu_max dt) equal the oracle's restatement built from the same
    cross sections."""
    pl = gpu_lib
    say("case: MCC tables vs the oracle (electrons and ions)")
    w = pkg.build_workload(pl, "small24", seed=5, ppc=4, mcc=True)
    dom = w["domain"]
    norm = dom.normalisation()
    dt_si = dom.dt * norm["time_ref"]
    for s, electron in ((0, True), (1, False)):
        sp = MCC_SPEC[s]
        ns, emin = 10001, sp["emin"]
        eint = (3.0 - emin) / (ns - 1)
        E = 10.0 ** (emin + np.arange(ns) * eint)
        prob_or, smax_or = orc.mcc_table(sp["nr"], ns, emin, eint, _species(pkg)[s][1], pkg.ARGON_ION["mass_kg"],
                                         3.2e20, _sigma_workload(E, electron).ravel())
        prob, smax, cr = pl.mcc_table(s, sp["nr"])
        np.testing.assert_allclose(prob, prob_or, rtol=1e-13, atol=1e-300)
        assert smax == pytest.approx(smax_or, rel=1e-13)
        # (1 - exp(-x) at x ~ 1e-5 keeps about 11 digits)
        assert cr == pytest.approx(-np.expm1(-smax_or * dt_si), rel=1e-9)
        # cumulative over reactions, normalised to a maximum of 1
        t = prob.reshape(ns, sp["nr"])
        assert t.max() == 1.0 and np.all(np.diff(t, axis=1) >= 0)


def test_cfg2_product_configuration_properties(gpu_lib, pkg):
    """cfg2 at full size (256^2 r-z, conductor at -1 and an eps-4 dielectric disk, 2 x 3.9M),
    learned solver, MCC, sort: 24 steps (one cell sort inside, at counter 20)."""
    pl = gpu_lib
    say("case: cfg2 full size, product configuration, 24 steps of properties")
    _product_run(pl, pkg, "cfg2", 24, 200)


def test_kappa_and_polytropic_initial_loads(gpu_lib, pkg):
    """pic_load_distribution with the reference's MaxKappa laws (MCCInterface.f90:168-176): the
    loaded Vx follow VelKappaInit's kappa = 2 law and VelPolyInit's uniform gamma = 3 law (the
    laws the oracle's restatement is pinned to, tests/test_oracle_pins_cpu.py), Vy = Vz = 0,
    positions as the Maxwellian load's, and the particles step."""
    from scipy import stats
    from test_oracle_pins_cpu import kappa_load_cdf, load_scale
    pl = gpu_lib
    say("case: kappa and polytropic initial loads")
    w = pkg.workload("small32")
    dom = w["domain"]
    pl.setup(dom, bits=64, solver=0, tol=1e-10)
    weight = pkg.species_weight(dom, 20)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2.0 * pkg.EVTOK, 32 * 32 * 40)
    pl.set_species(1, pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], weight, 0.026 * pkg.EVTOK, 32 * 32 * 40)
    pl.load_distribution(0, 20, 2.0, 9, 0)
    pl.load_distribution(1, 20, 0.026, 10, 2)
    pe, pi = pl.particles(0), pl.particles(1)
    assert np.all(pe["vy"] == 0) and np.all(pe["vz"] == 0) and np.all(pi["vy"] == 0) and np.all(pi["vz"] == 0)
    se = load_scale(2.0 * pkg.EVTOK, pkg.ELECTRON["mass_kg"], 0)
    si = load_scale(0.026 * pkg.EVTOK, pkg.ARGON_ION["mass_kg"], 2)
    assert stats.kstest(pe["vx"], kappa_load_cdf(se)).pvalue > 1e-3
    assert stats.kstest(pi["vx"], stats.uniform(-np.sqrt(3) * si, 2 * np.sqrt(3) * si).cdf).pvalue > 1e-3
    # positions: the Maxwellian load of the same seed puts the particles at the same places
    pl.load_distribution(0, 20, 2.0, 9, 1)
    pm = pl.particles(0)
    assert np.array_equal(pm["x"], pe["x"]) and np.array_equal(pm["loc"], pe["loc"])
    pl.load_distribution(0, 20, 2.0, 9, 0)
    pl.set_run_params(0, 9, 0)
    pl.initial_field()
    pl.step(2)
    with pytest.raises(pkg.PicError):
        pl.load_distribution(0, 20, 2.0, 9, 3)
