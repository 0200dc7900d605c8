"""Error paths and call-sequence semantics of the C-ABI, each forced from the host so that
every kernel exits normally: non-finite particle state, PCG iteration limit, capacity
overflow, deposition range, inconsistent Location on load, the random-stream step
counter of the drop-in sequence, and a stop/restart that resumes bit for bit.

The reference STOPs or PAUSEs on these conditions (MY_PCG.f90:88-90 IERR, MCCInterface.f90
PAUSE in the boundary loop); the library returns a status code and keeps running."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _small(pl, pkg, case="small24", bits=64, ppc=8, mcc=False, seed=11, solver=None):
    w = pkg.build_workload(pl, case, seed=seed, ppc=ppc, mcc=mcc)
    if bits != w["bits"] or solver is not None:
        # rebuild at the asked precision (set_precision discards species of the other width)
        pl.setup(w["domain"], bits=bits, solver=w["solver"] if solver is None else solver, tol=w["tol"])
        dom = w["domain"]
        nel = (dom.nnx - 1) * (dom.nny - 1)
        cap = int(nel * ppc * 1.5) + 1024
        weight = pkg.species_weight(dom, ppc)
        pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, w["Te"] * pkg.EVTOK, cap)
        pl.set_species(1, pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], weight, w["Ti"] * pkg.EVTOK, cap)
        pl.load_maxwellian(0, ppc, w["Te"], seed)
        pl.load_maxwellian(1, ppc, w["Ti"], seed + 1)
        if mcc:
            pl.MCCBundleInit(pkg.ARGON_ION["mass_kg"], 300.0, 3.2e20, 0, 1)
        pl.set_run_params(20, seed, 1 if mcc else 0)
    return w


def test_nonfinite_particle_reported(gpu_lib, pkg):
    pl = gpu_lib
    say("case: non-finite particle -> PIC_ERR_NONFINITE")
    _small(pl, pkg)
    p = pl.particles(0)
    vx = p["vx"].copy()
    vx[7] = np.nan
    pl.load_particles(0, p["x"], p["y"], vx, p["vy"], p["vz"], p["loc"])
    pl.MoveAll(0)
    pl.MoveAll(1)
    with pytest.raises(pkg.PicError) as ei:
        pl.GetParChg_2D()
    assert ei.value.code == 7 and "non-finite" in str(ei.value)
    # the particle is deleted (with those that left through the walls) and the library keeps working
    c = pl.counts(0)
    assert c["nlive"] == len(vx) - c["deleted"] and c["deleted"] >= 1
    assert not np.isnan(pl.particles(0)["vx"]).any()
    pl.IFE_Solve_2D()
    pl.GetEField_SIDG_PPR()
    pl.step(1)


def test_pcg_iteration_limit_reported(gpu_lib, pkg):
    pl = gpu_lib
    for solver in (0, 1):
        say(f"case: PCG_MaxIT = 1 -> PIC_ERR_PCG_MAXITER (solver {solver})")
        _small(pl, pkg, solver=solver)
        pl.GetParChg_2D()
        pkg.check(pkg.lib().pic_set_pcg_params(1e-10, 1))
        with pytest.raises(pkg.PicError) as ei:
            pl.IFE_Solve_2D()
        assert ei.value.code == 5
        pkg.check(pkg.lib().pic_set_pcg_params(1e-10, 50000))
        pl.IFE_Solve_2D()
        assert pl.last_solve()["relres"] <= 1e-10


def test_capacity_overflow_reported(gpu_lib, pkg):
    pl = gpu_lib
    say("case: injection beyond the slot capacity -> PIC_ERR_CAPACITY")
    w = pkg.workload("small24")
    dom = w["domain"]
    pl.setup(dom, bits=64, solver=0, tol=1e-10)
    weight = pkg.species_weight(dom, 8)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2 * pkg.EVTOK, 24 * 24 * 8)
    pl.load_maxwellian(0, 8, 2.0, 3)
    n0 = pl.counts(0)["nslots"]
    v_ref = dom.normalisation()["v_ref"]
    pl.set_injection(0, 100 * dom.n_ref, 2.0, (0.1 * np.sqrt(2.0) * v_ref, 0.0, 0.0), 0.0, 0.0, 24.0)
    pl.set_run_params(0, 3, 0)
    with pytest.raises(pkg.PicError) as ei:
        for it in range(50):
            pl.InjectBeams_2D(it)
    assert ei.value.code == 4
    # nothing was written by the refused injection
    c = pl.counts(0)
    assert c["nslots"] <= 24 * 24 * 8 and c["nslots"] >= n0
    p = pl.particles(0)
    assert len(p["x"]) == c["nslots"] and np.all(np.isfinite(p["x"]))


def test_deposition_of_extreme_energies(gpu_lib, pkg, orc):
    """A kappa re-emission (vx = sqrt(c (1/R - 1)), MCCInterface.f90:397-410) gives an electron
    MeV to PeV energies when R is small. Weight x energy terms >= 2^11 eV are held exactly
    (whole-eV part summed apart from the 2^40-scale fraction), so such particles deposit bit
    for bit like the oracle; a term past 2^62 eV is refused with PIC_ERR_OVERFLOW."""
    pl = gpu_lib
    say("case: deposition of 4.5 MeV and 80 GeV electrons; refusal past 2^62 eV")
    w = _small(pl, pkg)
    dom = w["domain"]
    norm = dom.normalisation()
    hp = pl.setup_array("HP")
    xs, ys = hp[0::2][:: dom.nny].copy(), hp[1::2][: dom.nny].copy()
    vol = pl.setup_array("CELL_VOLUME")[1::2]
    nn = dom.nnx * dom.nny
    p = pl.particles(0)
    vx = p["vx"].copy()
    vx[3], vx[9] = 3.0e3, 4.0e5
    pl.load_particles(0, p["x"], p["y"], vx, p["vy"], p["vz"], p["loc"])
    pl.GetParChg_2D()               # lookup and deposit of the positions as they are
    mass = pkg.ELECTRON["mass_kg"]
    args = (64, xs, ys, 0, p["x"], p["y"], vx, p["vy"], p["vz"], p["loc"] - 1, mass, norm["v_ref"])
    wq, eq = orc.deposit(*args, fix=True)
    wf, ef = orc.deposit(*args, fix=False)
    assert eq.max() > 1e10
    np.testing.assert_allclose(eq, ef, rtol=1e-12, atol=1e-9)
    rs, e = orc.normalize(0, pkg.species_weight(dom, 8), -1.0, norm["L_ref"], dom.n_ref, vol, wq, eq, np.zeros(nn))
    assert np.array_equal(pl.field("EK_TOT")[:nn], e)
    assert np.array_equal(pl.field("RHO_P")[:nn], wq)
    vx[9] = 2.0e10                  # about 2e20 eV: every corner's W x E (W >= 1/4 on one) is past 2^62 eV
    pl.load_particles(0, p["x"], p["y"], vx, p["vy"], p["vz"], p["loc"])
    with pytest.raises(pkg.PicError) as ei:
        pl.GetParChg_2D()
    assert ei.value.code == 12


def test_load_rejects_inconsistent_location(gpu_lib, pkg):
    pl = gpu_lib
    say("case: Location inconsistent with the position -> PIC_ERR_POSITION, nothing loaded")
    _small(pl, pkg)
    p = pl.particles(0)
    nel = 24 * 24
    for bad_loc in (0, nel + 1, int(p["loc"][5]) + 1):
        loc = p["loc"].copy()
        loc[5] = bad_loc
        with pytest.raises(pkg.PicError) as ei:
            pl.load_particles(0, p["x"], p["y"], p["vx"], p["vy"], p["vz"], loc)
        assert ei.value.code == 8 and "particle 5" in str(ei.value)
        assert pl.counts(0)["nslots"] == 0
    pl.load_particles(0, p["x"], p["y"], p["vx"], p["vy"], p["vz"], p["loc"])
    assert pl.counts(0)["nlive"] == len(p["x"])
    # FP32 storage: the Location is checked against the rounded position
    _small(pl, pkg, bits=32)
    q = pl.particles(0)
    pl.load_particles(0, q["x"], q["y"], q["vx"], q["vy"], q["vz"], q["loc"])
    assert np.array_equal(pl.particles(0)["loc"], q["loc"])


def test_dropin_sequence_advances_the_step(gpu_lib, pkg, orc):
    """The drop-ins in the reference's order draw new random numbers every step: MCC ends a
    step (the counter keying every stream advances), and so does a second MoveAll of a
    species at the same counter value. Two back-to-back MCC calls select from different
    streams: each matches the oracle's selection at its own counter value."""
    pl = gpu_lib
    say("case: drop-in step counter")
    _small(pl, pkg, mcc=True, ppc=16)
    pl.MCCBundleInit(pkg.ARGON_ION["mass_kg"], 300.0, 3.2e25, 0, 1)     # dense gas: many selections
    pkg.check(pkg.lib().pic_set_run_params(0, 11, 1))
    pl.set_step(5)
    pl.GetParChg_2D()
    prob, smax, cr = pl.mcc_table(1, 2)
    thr = int(min(np.floor(cr * 2 ** 32), 2 ** 32 - 1))
    for k in range(2):
        before = pl.particles(1)
        sel0 = pl.counts(1)["mcc_selected"]
        pl.MCC()
        sel = orc.mcc_select(before["loc"] - 1, thr, 11, 5 + k, orc.tag(1, orc.RNG_SELECT))
        other = orc.mcc_select(before["loc"] - 1, thr, 11, 5 + (1 - k), orc.tag(1, orc.RNG_SELECT))
        assert pl.get_step() == 6 + k
        got = pl.counts(1)["mcc_selected"] - sel0
        assert got == len(sel) and len(sel) > 20 and not np.array_equal(sel, other)
    # a species moved twice without MCC in between starts a new step
    s0 = pl.get_step()
    pl.MoveAll(0)
    pl.MoveAll(1)
    pl.GetParChg_2D()
    assert pl.get_step() == s0
    pl.MoveAll(0)
    assert pl.get_step() == s0 + 1


@pytest.mark.parametrize("bits,solver", [(64, 0), (32, 1)])
def test_restart_resumes_bit_for_bit(gpu_lib, pkg, tmp_path, bits, solver):
    """Stop after the cell sort of step 20 (slots compacted), keep the particles, Phi
    (Dump_2D) and the step counter, set the run up again from nothing, restore, and run 6
    more steps: particles, moments and Phi equal those of the run that never stopped."""
    pl = gpu_lib
    say(f"case: stop/restart at step 20, bits {bits}, solver {solver}")
    _small(pl, pkg, "small32", bits=bits, solver=solver, ppc=8, mcc=True, seed=21)
    pkg.check(pkg.lib().pic_set_pcg_params(1e-10, 50000))
    pl.set_step(0)
    pl.initial_field()
    pl.step(20)
    snap = [pl.particles(s) for s in (0, 1)]
    pl.dump_2d(20, tmp_path)
    pl.step(6)
    ref = dict(p=[pl.particles(s) for s in (0, 1)], phi=pl.field("PHI"), rho=pl.field("RHO_S"),
               ek=pl.field("EK_TOT"))
    _small(pl, pkg, "small32", bits=bits, solver=solver, ppc=8, mcc=True, seed=21)
    pkg.check(pkg.lib().pic_set_pcg_params(1e-10, 50000))
    for s in (0, 1):
        q = snap[s]
        pl.load_particles(s, q["x"], q["y"], q["vx"], q["vy"], q["vz"], q["loc"])
    pl.restart_2d(20, tmp_path)
    pl.GetEField_SIDG_PPR()
    pl.set_step(20)
    pl.step(6)
    for s in (0, 1):
        a, b = ref["p"][s], pl.particles(s)
        for k in ("x", "y", "vx", "vy", "vz", "loc"):
            assert np.array_equal(a[k], b[k]), (s, k)
    assert np.array_equal(ref["phi"], pl.field("PHI"))
    assert np.array_equal(ref["rho"], pl.field("RHO_S"))
    assert np.array_equal(ref["ek"], pl.field("EK_TOT"))
