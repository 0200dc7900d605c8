"""GPU parity: the HIP library through its C-ABI against the CPU oracle on identical inputs.

Bars: element indices, counts, particle state of the mover and deposited node moments
bit-exact; PPR coefficients, RHS and stiffness operators within 1e-12 relative; PCG
solutions within 1e-10 relative of the FP64 oracle at FP64; MCC event counts exact and
post-collision velocities within 1e-12 relative (transcendentals differ by ulps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mesh_of(pl, dom):
    hp = pl.setup_array("HP")
    xs = hp[0::2][:: dom.nny].copy()
    ys = hp[1::2][: dom.nny].copy()
    return xs, ys, pl.setup_array("ELEMENT_INDEX"), pl.setup_array("NODE_INDEX")


@pytest.fixture(scope="module")
def small(gpu_lib, pkg):
    say("case: small24 setup")
    w = pkg.build_workload(gpu_lib, "small24", seed=11, ppc=16, mcc=True)
    gpu_lib.initial_field()
    return w


# ------------------------------------------------------------------ set-up arrays
def test_setup_matches_restatements(gpu_lib, pkg, orc, small):
    from patch_ref import element_and_node_classes, point_patches
    pl, dom = gpu_lib, small["domain"]
    say("case: setup arrays small24")
    xs, ys, ei, ni = mesh_of(pl, dom)
    circles = [(o.x, o.y, o.r) for o in dom.objects]
    ni_r, ei_r = element_and_node_classes(dom.nnx, dom.nny, xs, ys, circles)
    assert np.array_equal(ni, ni_r) and np.array_equal(ei, ei_r)
    pp_r, _ = point_patches(dom.nnx, dom.nny, ei_r)
    assert np.array_equal(pl.setup_array("POINT_PATCH"), pp_r)
    # stiffness on unknowns == oracle element-by-element operator
    nt = pl.setup_array("NODE_TYPE").reshape(-1, 2)[:, 1]
    rp, col, val = pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL")
    nu = len(rp) - 1
    rng = np.random.default_rng(0)
    xu = rng.normal(size=nu)
    y_lib = np.array([val[rp[i]:rp[i + 1]] @ xu[col[rp[i]:rp[i + 1]]] for i in range(nu)])
    y_or = orc.stiffness_apply(xs, ys, 0, nt, xu)
    np.testing.assert_allclose(y_lib, y_or, rtol=1e-12, atol=1e-12)
    # rhs = B rho - c == oracle element loops with the Dirichlet lift
    rho = rng.normal(size=dom.nnx * dom.nny)
    brp, bcol, bval = pl.setup_array("B_ROWPTR"), pl.setup_array("B_COL"), pl.setup_array("B_VAL")
    c = pl.setup_array("RHS_CONST")
    b_lib = np.array([bval[brp[i]:brp[i + 1]] @ rho[bcol[brp[i]:brp[i + 1]]] for i in range(nu)]) - c
    b_or = orc.rhs(xs, ys, 0, nt, pl.setup_array("DIRICHLET_PHI"), rho, nu)
    np.testing.assert_allclose(b_lib, b_or, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------ dielectric (IFE) set-up and solve
def _dielectric_system(pl, pkg, orc, case):
    cond_eps = case.endswith("cond")       # the conductor given the reference inputs' Eps = 1e4
    w = pkg.workload(case[:-4] if cond_eps else case)
    dom = w["domain"]
    if cond_eps:
        dom.objects[0].eps = 1.0e4
    pl.setup(dom, bits=64, solver=0, tol=1e-14)
    xs, ys, ei, ni = mesh_of(pl, dom)
    nt = pl.setup_array("NODE_TYPE").reshape(-1, 2)[:, 1]
    circles = [(o.x, o.y, o.r, o.eps, o.conductor) for o in dom.objects]
    A, c, B, types = orc.assemble_ife(xs, ys, dom.delta, circles, nt, pl.setup_array("DIRICHLET_PHI"))
    return dom, xs, ys, ni, nt, A, c, B, types


@pytest.mark.parametrize("case", ["diel24", "dielrz24", "diel24cond"])
def test_dielectric_setup_matches_oracle(gpu_lib, pkg, orc, case):
    """A dielectric disk (eps 4): its inside nodes are unknowns (particles are absorbed there,
    node_index -2), elements inside carry beta = eps, and the elements the circle cuts take the
    IFE element matrices; the stiffness, the Dirichlet lift and the RHS operator agree with the
    oracle's element-by-element assembly within 1e-12 relative (FP64, different quadratures)."""
    import scipy.sparse as sp
    pl = gpu_lib
    say(f"case: dielectric set-up {case}")
    dom, xs, ys, ni, nt, A, c, B, types = _dielectric_system(pl, pkg, orc, case)
    assert set(types) == {1, 2} and len(types) > (40 if case.endswith("cond") else 20)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    d = dom.objects[1]
    inside = ((X - d.x) ** 2 + (Y - d.y) ** 2 < d.r * d.r).ravel()
    assert np.all(ni[inside] == -2) and np.all(nt[inside] > 0)
    rp, col, val = pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL")
    nu = len(rp) - 1
    A_lib = sp.csr_matrix((val, col, rp), shape=(nu, nu))
    scale = abs(A).max()
    # 1e-12 of the largest entry; with eps = 1e4 the basis systems are conditioned like the
    # permittivity ratio and the two derivations agree to about 1e-16 * 1e4 * 1e2 = 1e-10
    tol = 1e-10 if case.endswith("cond") else 1e-12
    assert abs(A_lib - A).max() <= tol * scale
    np.testing.assert_allclose(pl.setup_array("RHS_CONST"), c, rtol=0, atol=tol * max(1.0, abs(c).max()))
    brp, bcol, bval = pl.setup_array("B_ROWPTR"), pl.setup_array("B_COL"), pl.setup_array("B_VAL")
    B_lib = sp.csr_matrix((bval, bcol, brp), shape=(nu, dom.nnx * dom.nny))
    assert abs(B_lib - B).max() <= tol * abs(B).max()
    # the jump changes the operator: without the IFE element the cut rows differ
    A_q1, _, _, _ = orc.assemble_ife(xs, ys, dom.delta, [(o.x, o.y, o.r, o.eps, o.conductor) for o in dom.objects],
                                     nt, pl.setup_array("DIRICHLET_PHI"), mistake="q1_cut")
    assert abs(A_q1 - A).max() > 0.1 * scale


@pytest.mark.parametrize("case", ["diel24", "dielrz24"])
def test_dielectric_solve_matches_oracle(gpu_lib, pkg, orc, case):
    """IFE_Solve_2D on the dielectric domain: the FP64 solution within 1e-10 relative of a
    direct sparse solve of the oracle's system on the same charge."""
    import scipy.sparse.linalg as spl
    pl = gpu_lib
    say(f"case: dielectric solve {case}")
    dom, xs, ys, ni, nt, A, c, B, types = _dielectric_system(pl, pkg, orc, case)
    rng = np.random.default_rng(5)
    rho = rng.normal(size=dom.nnx * dom.nny)
    pl.set_field("RHO", rho)
    pl.IFE_Solve_2D()
    phi = pl.field("PHI")
    u = spl.spsolve(A.tocsc(), B @ rho - c)
    ref = pl.setup_array("DIRICHLET_PHI").copy()
    ref[nt > 0] = u[nt[nt > 0] - 1]
    np.testing.assert_allclose(phi, ref, rtol=0, atol=1e-10 * abs(ref).max())


def test_dielectric_limits_are_refused(gpu_lib, pkg):
    """Configurations without a bilinear IFE element are refused at set-up with
    PIC_ERR_UNSUPPORTED (host-side checks): a dielectric box, a dielectric on non-square
    elements, an element touching two objects."""
    import ctypes as C
    say("case: dielectric limits")
    base = pkg.workload("diel24")["domain"]
    c = base.to_c()
    c.objects[1].shape = 3
    c.objects[1].locations[0][0], c.objects[1].locations[0][1] = 12.0, 8.0
    c.objects[1].locations[1][0], c.objects[1].locations[1][1] = 18.0, 14.0
    assert pkg.lib().pic_setup_domain(C.byref(c)) == 9
    c = base.to_c()
    c.ymax = 36.0                                   # hy = 1.5 hx
    assert pkg.lib().pic_setup_domain(C.byref(c)) == 9
    c = base.to_c()
    c.objects[0].locations[0][0] = 10.9             # the conductor now overlaps the disk's elements
    c.objects[0].dimensions[0] = 4.0
    assert pkg.lib().pic_setup_domain(C.byref(c)) == 9
    # the unmodified domain is accepted afterwards
    assert pkg.lib().pic_setup_domain(C.byref(base.to_c())) == 0


def test_empty_species_steps(gpu_lib, pkg):
    """A species with no particles (ions not loaded) steps, sorts and deposits: its moments
    are zero and the electrons' step is unchanged by its presence."""
    pl = gpu_lib
    say("case: empty species small24")
    outs = []
    for with_ions in (False, True):
        w = pkg.workload("small24")
        dom = w["domain"]
        pl.setup(dom, bits=64, solver=0, tol=1e-12)
        nel = (dom.nnx - 1) * (dom.nny - 1)
        weight = pkg.species_weight(dom, 8)
        pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2 * pkg.EVTOK, nel * 16)
        if with_ions:
            pl.set_species(1, pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], weight, 0.026 * pkg.EVTOK,
                           nel * 16)
        pl.load_maxwellian(0, 8, 2.0, 5)
        pl.set_run_params(2, 77, 0)
        pl.sort_particles()
        pl.initial_field()
        pl.step(3)
        if with_ions:
            assert pl.counts(1)["nlive"] == 0
        outs.append((pl.particles(0)["x"].copy(), pl.field("RHO").copy(), pl.field("PHI").copy()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_axisymmetric_load_follows_volume(gpu_lib, pkg):
    """r-z loads: element (i, j) receives floor((i+1) q_j) - floor(i q_j) particles, q_j = ppc ney
    (y_{j+1}^2 - y_j^2) / (ymax^2 - ymin^2), so rows away from the object hold counts in
    proportion to their volume (uniform density)."""
    pl = gpu_lib
    say("case: r-z load rz24")
    dom = pkg.workload("rz24")["domain"]
    pl.setup(dom, bits=64, solver=0, tol=1e-10)
    nel = (dom.nnx - 1) * (dom.nny - 1)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], pkg.species_weight(dom, 8), 2 * pkg.EVTOK,
                   nel * 16)
    pl.load_maxwellian(0, 8, 2.0, 5)
    y = pl.particles(0)["y"]
    n, ppc = dom.nny - 1, 8
    o = dom.objects[0]
    rows = np.floor(y).astype(int)
    for j in range(n):
        if j + 1 > o.y - o.r - 1 and j < o.y + o.r + 1:
            continue
        q = ppc * n * ((j + 1) ** 2 - j ** 2) / float(n * n)
        assert np.count_nonzero(rows == j) == sum(int(np.floor((i + 1) * q) - np.floor(i * q)) for i in range(n)), j
    # total close to ppc per element
    assert abs(len(y) / nel - ppc) < 1.0


# ------------------------------------------------------------------ mover
@pytest.mark.parametrize("bits,case", [(64, "small24"), (32, "small24"), (64, "skew24"), (32, "skew24")])
def test_moveall_bit_exact(gpu_lib, pkg, orc, bits, case):
    """skew24: origin (0.3, -0.1) and spacings 0.7 x 0.9, so the fused pass locates in FP64,
    divides by the spacing in the weights and rounds its FP32 node lines from the FP64 ones."""
    pl = gpu_lib
    say(f"case: MoveAll bit-exact bits={bits} {case}")
    w = pkg.workload(case)
    dom = w["domain"]
    pl.setup(dom, bits=bits, solver=0, tol=1e-12)
    weight = pkg.species_weight(dom, 16)
    nel = (dom.nnx - 1) * (dom.nny - 1)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2 * pkg.EVTOK, nel * 32)
    pl.set_species(1, pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], weight, 0.026 * pkg.EVTOK, nel * 32)
    pl.load_maxwellian(0, 16, 2.0, 5)
    pl.load_maxwellian(1, 16, 0.026, 6)
    pl.set_run_params(0, 77, 0)
    pl.set_step(3)
    pl.initial_field()
    xs, ys, ei, ni = mesh_of(pl, dom)
    coef = pl.field("COEF_PPR").reshape(-1, 6)
    efx, efy = pl.field("EFX"), pl.field("EFY")
    # make the field strong enough to drive particles into walls and the object
    efx *= 50.0
    efy *= 50.0
    pl.set_field("EFX", efx)
    pl.set_field("EFY", efy)
    coef[:, 1:] *= 50.0
    pl.set_field("COEF_PPR", coef.ravel())
    norm = dom.normalisation()
    ft = np.float64 if bits == 64 else np.float32
    for isp, (chg, mass, T) in enumerate([(pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], 2.0),
                                          (pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], 0.026)]):
        p0 = pl.particles(isp)
        qm = (chg / pkg.ELECTRON_CHARGE) / (mass / pkg.ELECTRON_MASS)
        res = orc.move_all(bits, xs, ys, ei, ni, dom.object_array(), efx.astype(ft), efy.astype(ft),
                           coef[:, 1:].astype(ft).ravel(), qm, dom.dt, norm["v_ref"], mass, T * pkg.EVTOK, 77, 3,
                           orc.tag(isp, orc.RNG_EMIT), p0["x"], p0["y"], p0["vx"], p0["vy"], p0["vz"], p0["loc"] - 1)
        x, y, vx, vy, vz, alive, cnt, err = res
        assert err == 0
        pl.MoveAll(isp)
        pl.GetParChg_2D()
        p1 = pl.particles(isp)
        keep = alive.astype(bool)
        assert len(p1["x"]) == keep.sum()
        for k, a in zip(("x", "y", "vx", "vy", "vz"), (x, y, vx, vy, vz)):
            assert np.array_equal(p1[k], a[keep].astype(np.float64)), k
        c = pl.counts(isp)
        assert (c["emit"], c["right"], c["object"], c["reflect"], c["deleted"]) == \
               tuple(int(v) for v in (cnt[0], cnt[1], cnt[2], cnt[3], cnt[7]))
        loc, bad = orc.locate(xs, ys, x[keep].astype(np.float64), y[keep].astype(np.float64))
        assert bad == 0 and np.array_equal(p1["loc"] - 1, loc)
        say(f"  species {isp}: {keep.sum()} live, emitted {cnt[0]}, absorbed {cnt[1] + cnt[2]}, reflected {cnt[3]}")


def test_lookup_on_grid_lines_fp32(gpu_lib, pkg, orc):
    """FP32 particles at rest exactly on element edges, one float step either side of them,
    mid-cell and on the top and right walls (the top-edge clamps of the lookup): after the
    fused pass (which locates in-domain FP32 coordinates without the grid-line arithmetic on
    this mesh) every element equals the oracle's restatement of the reference's lookup, and
    the particles are where the oracle's MoveAll leaves them."""
    pl = gpu_lib
    say("case: lookup on grid lines, FP32, small24")
    dom = pkg.workload("small24")["domain"]
    pl.setup(dom, bits=32, solver=0, tol=1e-12)
    xs, ys, ei, ni = mesh_of(pl, dom)
    f = np.float32
    def cand(lines):
        out = [lines[-1]]
        for v in lines[1:-1]:
            v32 = f(v)
            out += [v32, np.nextafter(v32, f(-np.inf)), np.nextafter(v32, f(np.inf)), v32 + f(0.5)]
        return np.unique(np.array(out, dtype=f))
    cx, cy = cand(xs), cand(ys)
    X, Y = np.meshgrid(cx, cy, indexing="ij")
    x = X.ravel().astype(np.float64)
    y = Y.ravel().astype(np.float64)
    n = len(x)
    weight = pkg.species_weight(dom, 16)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2 * pkg.EVTOK, n + 64)
    zero = np.zeros(n)
    pl.load_particles(0, x, y, zero, zero, zero)
    pl.set_run_params(0, 77, 0)
    pl.set_step(1)
    nn = dom.nnx * dom.nny
    pl.set_field("EFX", np.zeros(nn))
    pl.set_field("EFY", np.zeros(nn))
    pl.set_field("COEF_PPR", np.zeros(6 * nn))
    p0 = pl.particles(0)
    norm = dom.normalisation()
    qm = (pkg.ELECTRON["charge_C"] / pkg.ELECTRON_CHARGE) / (pkg.ELECTRON["mass_kg"] / pkg.ELECTRON_MASS)
    z32 = np.zeros(nn, dtype=f)
    res = orc.move_all(32, xs, ys, ei, ni, dom.object_array(), z32, z32, np.zeros(5 * nn, dtype=f), qm, dom.dt,
                       norm["v_ref"], pkg.ELECTRON["mass_kg"], 2 * pkg.EVTOK, 77, 1, orc.tag(0, orc.RNG_EMIT),
                       p0["x"], p0["y"], p0["vx"], p0["vy"], p0["vz"], p0["loc"] - 1)
    xo, yo, vxo, vyo, vzo, alive, cnt, err = res
    assert err == 0
    pl.MoveAll(0)
    pl.GetParChg_2D()
    p1 = pl.particles(0)
    keep = alive.astype(bool)
    assert len(p1["x"]) == keep.sum()
    assert np.array_equal(p1["x"], xo[keep].astype(np.float64)) and np.array_equal(p1["y"], yo[keep].astype(np.float64))
    loc, bad = orc.locate(xs, ys, xo[keep].astype(np.float64), yo[keep].astype(np.float64))
    assert bad == 0 and np.array_equal(p1["loc"] - 1, loc)
    say(f"  {n} particles, {keep.sum()} live")


@pytest.mark.parametrize("bits", [64, 32])
def test_moveall_axisymmetric_bit_exact(gpu_lib, pkg, orc, bits):
    """The r-z mover (MoveOne.f90:94-104, :221-236, :246-263; x axial, y radial, velocities
    in a fixed Cartesian frame, the azimuth PO%Z moved with them) against the oracle, bit for
    bit, azimuth included, with the field strong enough to reach the axial walls, the outer
    radius (absorbing in the build; the reference has no branch there) and the conductor. The
    particles are reloaded with their azimuths through load_particles(th=...)."""
    pl = gpu_lib
    say(f"case: MoveAll axisymmetric bits={bits}")
    w = pkg.workload("rz24")
    dom = w["domain"]
    pl.setup(dom, bits=bits, solver=0, tol=1e-12)
    weight = pkg.species_weight(dom, 16)
    nel = (dom.nnx - 1) * (dom.nny - 1)
    species = [(pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], 2.0),
               (pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], 0.026)]
    for isp, (chg, mass, T) in enumerate(species):
        pl.set_species(isp, chg, mass, weight, T * pkg.EVTOK, nel * 32)
        pl.load_maxwellian(isp, 16, T, 5 + isp)
        p, th = pl.particles(isp), pl.azimuth(isp)
        assert len(th) == len(p["x"]) and np.all((th >= 0) & (th < 2 * np.pi))
        pl.load_particles(isp, p["x"], p["y"], p["vx"], p["vy"], p["vz"], loc=p["loc"], th=th)
        assert np.array_equal(pl.azimuth(isp), th)
    pl.set_run_params(0, 77, 0)
    pl.set_step(3)
    pl.initial_field()
    xs, ys, ei, ni = mesh_of(pl, dom)
    coef = pl.field("COEF_PPR").reshape(-1, 6)
    efx, efy = pl.field("EFX") * 20.0, pl.field("EFY") * 20.0
    pl.set_field("EFX", efx)
    pl.set_field("EFY", efy)
    coef[:, 1:] *= 20.0
    pl.set_field("COEF_PPR", coef.ravel())
    norm = dom.normalisation()
    ft = np.float64 if bits == 64 else np.float32
    for isp, (chg, mass, T) in enumerate(species):
        p0, th0 = pl.particles(isp), pl.azimuth(isp)
        qm = (chg / pkg.ELECTRON_CHARGE) / (mass / pkg.ELECTRON_MASS)
        x, y, vx, vy, vz, alive, cnt, err, th = orc.move_all(
            bits, xs, ys, ei, ni, dom.object_array(), efx.astype(ft), efy.astype(ft), coef[:, 1:].astype(ft).ravel(),
            qm, dom.dt, norm["v_ref"], mass, T * pkg.EVTOK, 77, 3, orc.tag(isp, orc.RNG_EMIT),
            p0["x"], p0["y"], p0["vx"], p0["vy"], p0["vz"], p0["loc"] - 1, delta=1, th=th0)
        assert err == 0
        pl.MoveAll(isp)
        pl.GetParChg_2D()
        p1, th1 = pl.particles(isp), pl.azimuth(isp)
        keep = alive.astype(bool)
        assert len(p1["x"]) == keep.sum()
        for k, a in zip(("x", "y", "vx", "vy", "vz"), (x, y, vx, vy, vz)):
            assert np.array_equal(p1[k], a[keep].astype(np.float64)), k
        assert np.array_equal(th1, th[keep].astype(np.float64))
        c = pl.counts(isp)
        assert (c["emit"], c["right"], c["object"], c["reflect"], c["deleted"]) == \
               tuple(int(v) for v in (cnt[0], cnt[1], cnt[2], cnt[3], cnt[7]))
        loc, bad = orc.locate(xs, ys, x[keep].astype(np.float64), y[keep].astype(np.float64))
        assert bad == 0 and np.array_equal(p1["loc"] - 1, loc)
        moved = np.abs(th[keep] - th0[keep]) > 0
        say(f"  species {isp}: {keep.sum()} live, emitted {cnt[0]}, absorbed {cnt[1] + cnt[2]}, "
            f"azimuth changed for {moved.sum()}")
        assert moved.any()


# ------------------------------------------------------------------ deposition
@pytest.mark.parametrize("bits,case", [(64, "small24"), (32, "small24"), (64, "rz24"), (32, "rz24"),
                                       (64, "skew24"), (32, "skew24")])
def test_getparchg_moments(gpu_lib, pkg, orc, bits, case):
    pl = gpu_lib
    say(f"case: GetParChg moments bits={bits} {case}")
    w = pkg.build_workload(pl, case, seed=21, ppc=16, mcc=False)
    dom = w["domain"]
    pl.setup(dom, bits=bits, solver=0, tol=1e-12)
    weight = pkg.species_weight(dom, 16)
    nel = (dom.nnx - 1) * (dom.nny - 1)
    pl.set_species(0, pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"], weight, 2 * pkg.EVTOK, nel * 20)
    pl.set_species(1, pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"], weight, 0.026 * pkg.EVTOK, nel * 20)
    pl.load_maxwellian(0, 16, 2.0, 21)
    pl.load_maxwellian(1, 16, 0.026, 22)
    pl.GetParChg_2D()
    xs, ys, ei, ni = mesh_of(pl, dom)
    norm = dom.normalisation()
    vol = pl.setup_array("CELL_VOLUME")[1::2]
    nn = dom.nnx * dom.nny
    rho_or = np.zeros(nn)
    rho_s, rho_p, ek = pl.field("RHO_S"), pl.field("RHO_P"), pl.field("EK_TOT")
    for isp, (chg, mass) in enumerate([(pkg.ELECTRON["charge_C"], pkg.ELECTRON["mass_kg"]),
                                       (pkg.ARGON_ION["charge_C"], pkg.ARGON_ION["mass_kg"])]):
        p = pl.particles(isp)
        loc, bad = orc.locate(xs, ys, p["x"], p["y"])
        assert bad == 0 and np.array_equal(loc, p["loc"] - 1)
        ft = np.float64 if bits == 64 else np.float32
        args = (bits, xs, ys, dom.delta, p["x"].astype(ft), p["y"].astype(ft), p["vx"].astype(ft), p["vy"].astype(ft),
                p["vz"].astype(ft), loc, mass, norm["v_ref"])
        wq, eq = orc.deposit(*args, fix=True)
        wf, ef = orc.deposit(*args, fix=False)
        assert np.array_equal(rho_p[isp * nn:(isp + 1) * nn], wq)            # bit-exact
        # vs the FP64 sums: each fixed-point quantum is within 2^-41 of its term, and a node sums at
        # most 4 x (particles of its four elements) terms
        nterm = 4 * np.bincount(loc, minlength=(dom.nnx - 1) * (dom.nny - 1)).max()
        np.testing.assert_allclose(rho_p[isp * nn:(isp + 1) * nn], wf, rtol=0, atol=nterm * 2.0 ** -41)
        rs, e = orc.normalize(dom.delta, weight, chg / pkg.ELECTRON_CHARGE, norm["L_ref"], dom.n_ref, vol, wq, eq,
                              rho_or)
        assert np.array_equal(rho_s[isp * nn:(isp + 1) * nn], rs)
        assert np.array_equal(ek[isp * nn:(isp + 1) * nn], e)
        if dom.delta == 0:
            assert wq.sum() == pytest.approx(len(p["x"]), rel=1e-12)
        else:
            # the axis row's lower-node weights carry beta = 0.75 (GetParChg_2D.f90:803-821):
            # a particle's weights sum to beta (R2^2 - R^2)/den + (R^2 - R1^2)/den there, 1 elsewhere
            cj = loc % (dom.nny - 1)
            r1, r2 = dom.ymin + ys[cj], dom.ymin + ys[cj + 1]
            beta = np.where(np.abs(ys[cj] - dom.ymin) < 1e-12, 0.75, 1.0)
            den = r2 * r2 - r1 * r1
            expect = beta * (r2 * r2 - p["y"] ** 2) / den + (p["y"] ** 2 - r1 * r1) / den
            assert (beta < 1).any()
            assert wq.sum() == pytest.approx(expect.sum(), rel=1e-12)
    assert np.array_equal(pl.field("RHO"), rho_or)


# ------------------------------------------------------------------ PPR
def test_ppr_matches_oracle(gpu_lib, pkg, orc, small):
    pl, dom = gpu_lib, small["domain"]
    say("case: PPR small24")
    pkg.build_workload(pl, "small24", seed=11, ppc=16, mcc=False)
    xs, ys, ei, ni = mesh_of(pl, dom)
    rng = np.random.default_rng(8)
    phi = rng.normal(size=dom.nnx * dom.nny)
    pl.set_field("PHI", phi)
    pl.GetEField_SIDG_PPR()
    coef_or, efx_or, efy_or, ns = orc.ppr(xs, ys, pl.setup_array("POINT_PATCH"), phi)
    assert ns == 0
    np.testing.assert_allclose(pl.field("COEF_PPR"), coef_or, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pl.field("EFX"), efx_or, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pl.field("EFY"), efy_or, rtol=1e-12, atol=1e-12)
    # Get_ParticleEField_PPR drop-in
    e0 = rng.integers(0, (dom.nnx - 1) * (dom.nny - 1), 500)
    ci, cj = np.divmod(e0, dom.nny - 1)
    px, py = xs[ci] + rng.random(500), ys[cj] + rng.random(500)
    ex, ey = pl.Get_ParticleEField_PPR(px, py, e0 + 1)
    ex_o, ey_o = orc.ppr_gather(xs, ys, ei, ni, pl.field("COEF_PPR"), px, py, e0)
    np.testing.assert_allclose(ex, ex_o, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(ey, ey_o, rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ PCG
def test_pcg_fp64_matches_my_pcg(gpu_lib, pkg, orc):
    pl = gpu_lib
    say("case: PCG fp64 small40")
    w = pkg.workload("small40")
    dom = w["domain"]
    pl.setup(dom, bits=64, solver=0, tol=1e-12)
    rng = np.random.default_rng(12)
    nn = dom.nnx * dom.nny
    rho = rng.normal(size=nn)
    phi = pl.setup_array("DIRICHLET_PHI").copy()
    pl.IFE_Solve_2D(Rho=rho, Phi=phi)
    info = pl.last_solve()
    rp, col, val = pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL")
    diag = pl.setup_array("A_DIAG")
    b = pl.field("RHS")
    un = pl.setup_array("UNKNOWN_NODE") - 1
    x_or, r_or, it_or, ierr = orc.pcg(rp, col, val, diag, b, np.zeros(len(b)), 1e-12, 50000)
    assert ierr == 0
    assert abs(info["iterations"] - (it_or + 1)) <= 1
    np.testing.assert_allclose(phi[un], x_or, rtol=1e-10, atol=1e-10 * np.abs(x_or).max())
    assert info["relres"] < 1e-11
    # MY_JPCG_Solver drop-in on the reference's SPARSE layout
    mat, nz = pkg.sparse_from_csr(rp, col, val)
    R = b.copy()
    X = np.zeros(len(b))
    it = pl.MY_JPCG_Solver(mat, R, X, diag.copy(), nz, len(b))
    assert abs(it - it_or) <= 1
    np.testing.assert_allclose(X, x_or, rtol=1e-10, atol=1e-10 * np.abs(x_or).max())


def test_pcg_mixed_precision_refinement(gpu_lib, pkg):
    pl = gpu_lib
    say("case: PCG mixed small64")
    w = pkg.workload("small64")
    dom = w["domain"]
    nn = dom.nnx * dom.nny
    rng = np.random.default_rng(3)
    rho = rng.normal(size=nn)
    pl.setup(dom, bits=32, solver=1, tol=1e-8)
    phi32 = pl.setup_array("DIRICHLET_PHI").copy()
    pl.IFE_Solve_2D(Rho=rho, Phi=phi32)
    info = pl.last_solve()
    assert info["relres"] <= 1e-8 and info["outer"] >= 1
    pl.setup(dom, bits=64, solver=0, tol=1e-13)
    phi64 = pl.setup_array("DIRICHLET_PHI").copy()
    pl.IFE_Solve_2D(Rho=rho, Phi=phi64)
    # stated bound for the mixed solve: ||phi32 - phi64|| / ||phi64|| <= 1e-6 at a 1e-8 residual
    assert np.linalg.norm(phi32 - phi64) / np.linalg.norm(phi64) < 1e-6
    say(f"  mixed: {info}")


# ------------------------------------------------------------------ MCC
def test_mcc_selection_and_events_match(gpu_lib, pkg, orc):
    pl = gpu_lib
    say("case: MCC small24 (dense gas)")
    w = pkg.build_workload(pl, "small24", seed=31, ppc=16, mcc=True, gas_density=3.2e25)
    pl.set_run_params(0, 31, 1)
    pl.set_step(5)
    pl.GetParChg_2D()
    norm = w["domain"].normalisation()
    before = [pl.particles(s) for s in (0, 1)]
    tabs = []
    for s, nr in ((0, 3), (1, 2)):
        prob, smax, cr = pl.mcc_table(s, nr)
        tabs.append((prob, cr))
    pl.MCC()
    after = [pl.particles(s) for s in (0, 1)]
    masses = (pkg.ELECTRON["mass_kg"], pkg.ARGON_ION["mass_kg"])
    specs = {0: dict(nr=3, emin=-3.0, rtype=[11, 12, 13], thr=[0.0, 11.55, 15.76]),
             1: dict(nr=2, emin=-1.0, rtype=[101, 103], thr=[0.0, 0.0])}
    new_e, new_i = [], []
    for s in (0, 1):
        prob, cr = tabs[s]
        thr = np.uint32(min(np.floor(cr * 2 ** 32), 2 ** 32 - 1))
        b = before[s]
        sel = orc.mcc_select(b["loc"] - 1, int(thr), 31, 5, orc.tag(s, orc.RNG_SELECT))
        assert pl.counts(s)["mcc_selected"] == len(sel) and len(sel) > 20
        sp = specs[s]
        ev = np.zeros(sp["nr"], dtype=np.int64)
        v_exp = np.stack([b["vx"], b["vy"], b["vz"]], 1).copy()
        for k in sel:
            idx, v, ne, ni = orc.mcc_collide(1 if s == 0 else 0, int(k), 31, 5, orc.tag(s, orc.RNG_COLLIDE),
                                             norm["v_ref"], masses[s], masses[1], 300.0, sp["nr"], 10001, sp["emin"],
                                             (3.0 - sp["emin"]) / 10000, 1e3, prob, sp["rtype"], sp["thr"], v_exp[k])
            v_exp[k] = v
            if idx:
                ev[idx - 1] += 1
            if s == 0 and idx == 3:
                new_e.append(ne)
                new_i.append(ni)
        assert np.array_equal(pl.mcc_events(s, sp["nr"]), ev), (s, ev)
        a = after[s]
        n0 = len(b["x"])
        np.testing.assert_allclose(np.stack([a["vx"], a["vy"], a["vz"]], 1)[:n0], v_exp, rtol=1e-12, atol=1e-15)
        say(f"  species {s}: selected {len(sel)}, events {ev.tolist()}")
    ne_lib = np.stack([after[0]["vx"], after[0]["vy"], after[0]["vz"]], 1)[len(before[0]["x"]):]
    ni_lib = np.stack([after[1]["vx"], after[1]["vy"], after[1]["vz"]], 1)[len(before[1]["x"]):]
    assert len(ne_lib) == len(new_e) and len(ni_lib) == len(new_i)
    if new_e:
        np.testing.assert_allclose(ne_lib, np.array(new_e), rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(ni_lib, np.array(new_i), rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------------ drop-ins
def test_positioning_dropins(gpu_lib, pkg, orc):
    pl = gpu_lib
    say("case: positioning drop-ins")
    w = pkg.workload("small24")
    dom = w["domain"]
    pl.setup(dom, bits=64, solver=0)
    xs = np.arange(dom.nnx, dtype=float)
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(2000) * 24, np.arange(25.0), np.full(25, 7.0)])
    y = np.concatenate([rng.random(2000) * 24, np.full(25, 3.0), np.arange(25.0)])
    e_lib = pl.locate(x, y)
    e_or, bad = orc.locate(xs, xs, x, y)
    assert bad == 0 and np.array_equal(e_lib - 1, e_or)
    inner = (x > 0.5) & (x < 23.5) & (y > 0.5) & (y < 23.5) & (np.mod(x, 1) > 1e-9) & (np.mod(y, 1) > 1e-9)
    ei = pl.InjectPositioning(x[inner], y[inner])
    assert np.array_equal(ei - 1, e_or[inner])
    ne, tf = pl.ParticlePositioning(x, y, e_lib)
    assert np.array_equal(ne, e_lib)
    on_inner_edge = ((np.mod(x, 1) == 0) & (x > 0) & (x < 24)) | ((np.mod(y, 1) == 0) & (y > 0) & (y < 24))
    corner = (np.mod(x, 1) == 0) & (np.mod(y, 1) == 0)
    assert np.array_equal(tf.astype(bool), on_inner_edge & ~corner)
    em = pl.generate_Elementmap(0)
    assert np.array_equal(em, (np.arange(24)[None, :] + 1) + np.arange(24)[:, None] * 24)


# ------------------------------------------------------------------ full step determinism
@pytest.mark.parametrize("case", ["small32", "rz32", "diel32", "dielrz32", "skew32"])
def test_full_step_deterministic_and_conserving(gpu_lib, pkg, case):
    """Seven full steps (mover, deposit, solve, PPR, MCC with ionisation products, a cell sort
    every 3 steps) twice from the same state: identical bits, the live count consistent; on
    the axisymmetric domain the azimuths travel with their particles through the sort and the
    MCC products."""
    pl = gpu_lib
    say(f"case: full step determinism {case}")
    outs = []
    for rep in range(2):
        w = pkg.build_workload(pl, case, seed=41, ppc=16, mcc=True, gas_density=3.2e24)
        pl.set_run_params(3, 41, 1)
        pl.set_step(0)
        pl.initial_field()
        pl.step(7)
        o = [pl.field("PHI"), pl.field("RHO"), pl.particles(0)["x"], pl.particles(1)["vx"]]
        if w["domain"].delta == 1:
            o += [pl.azimuth(0), pl.azimuth(1)]
            assert len(o[-2]) == len(o[2]) and np.all(np.isfinite(o[-2]))
        outs.append(o)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    c0 = pl.counts(0)
    assert c0["nlive"] == len(outs[1][2])
    if case.startswith("rz"):
        assert c0["mcc_created"] > 0


# ------------------------------------------------------------------ wall injection
@pytest.mark.parametrize("bits,slab", [(64, "full"), (32, "full"), (32, "top"), (64, "top")])
def test_injection_matches_oracle(gpu_lib, pkg, orc, bits, slab):
    """InjectBeams_2D: the count, and every appended particle's position, velocity and element,
    bit-exact against the oracle's restatement on the same streams (FP64 and FP32 particles).
    The "top" slab is 1e-3 high under the mesh's upper edge, where about one FP32 Y in 2000
    rounds onto the edge and is stored as the float below it."""
    pl = gpu_lib
    say(f"case: InjectBeams_2D small32 bits={bits} slab={slab}")
    w = pkg.workload("small32")
    dom = w["domain"]
    pl.setup(dom, bits=bits, solver=0, tol=1e-10)
    nel = (dom.nnx - 1) * (dom.nny - 1)
    weight = pkg.species_weight(dom, 8)
    masses = {0: pkg.ELECTRON["mass_kg"], 1: pkg.ARGON_ION["mass_kg"]}
    temps = {0: 2.0, 1: 0.026}
    pl.set_species(0, pkg.ELECTRON["charge_C"], masses[0], weight, temps[0] * pkg.EVTOK, nel * 16)
    pl.set_species(1, pkg.ARGON_ION["charge_C"], masses[1], weight, temps[1] * pkg.EVTOK, nel * 16)
    pl.load_maxwellian(0, 8, temps[0], 5)
    pl.load_maxwellian(1, 8, temps[1], 6)
    pl.set_run_params(20, 77, 0)
    norm = dom.normalisation()
    y_hi = float(dom.ymax)
    y_lo = 0.0 if slab == "full" else y_hi - 1e-3
    ylen = y_hi - y_lo
    src = {}
    for s in (0, 1):
        xm = masses[s] / pkg.ELECTRON_MASS
        tmpj = temps[s] * pkg.EVTOK / (dom.T_ref_eV * pkg.EVTOK)
        vt = np.sqrt(tmpj / xm)
        drift = (0.1 * vt * norm["v_ref"] * (1 if s == 0 else 20), 0.02 * norm["v_ref"], -0.01 * norm["v_ref"])
        vd = [d / norm["v_ref"] for d in drift]
        c = orc.inject_consts(vd[0], vt, tmpj, xm, 0.0, dom.dt)
        per_dens = abs(c[0]) * norm["v_ref"] * dom.dt * norm["time_ref"] * ylen * norm["L_ref"] / weight
        dens = (2000.0 if s == 0 else 400.0) / per_dens       # about 2000 electrons, 400 ions per call
        pl.set_injection(s, dens, temps[s], drift, r_start=0.0, y_lo=y_lo, y_hi=y_hi)
        expect = (dens / dom.n_ref) * dom.n_ref * abs(c[0]) * norm["v_ref"] * dom.dt * norm["time_ref"] * ylen * \
            norm["L_ref"] / weight
        src[s] = (c, vd, expect)
    before = [pl.counts(s)["nslots"] for s in (0, 1)]
    added = pl.InjectBeams_2D(7)
    say(f"  added {added[:2]}")
    assert added[0] > 1000 and added[1] > 100
    for s in (0, 1):
        c, vd, expect = src[s]
        tg = orc.tag(s, orc.RNG_INJECT)
        assert added[s] == orc.inject_count(expect, 77, 7, tg)
        x, y, vx, vy, vz, cell = orc.inject(bits, added[s], 1 if s == 0 else 0, c, vd, y_lo, y_hi, 77, 7, tg,
                                            dom.xmin, dom.ymin, 1.0, 1.0, dom.nnx - 1, dom.nny - 1)
        p = pl.particles(s)
        assert pl.counts(s)["nslots"] == before[s] + added[s]
        n = added[s]
        for k, ref in (("x", x), ("y", y), ("vx", vx), ("vy", vy), ("vz", vz)):
            assert np.array_equal(p[k][-n:], ref), (s, k)
        assert np.array_equal(p["loc"][-n:] - 1, cell), s
        assert np.all(cell >= 0) and np.all(y < y_hi)
    # the injected particles move and deposit like any other in the next step
    pl.initial_field()
    pl.step(1)
    assert pl.counts(0)["nlive"] > 0


@pytest.mark.parametrize("bits", [64, 32])
def test_injection_axisymmetric_matches_oracle(gpu_lib, pkg, orc, bits):
    """InjectBeams_2D on the axisymmetric domain (InjectBeams_2D.f90:59-61): the count with the
    disk area PI (y_hi - y_lo)^2 L_ref^2 (y_lo = 0: the axis), and every particle's position,
    radius, azimuth (PO%Z), velocity and element bit-exact against the oracle."""
    pl = gpu_lib
    say(f"case: InjectBeams_2D rz32 bits={bits}")
    w = pkg.workload("rz32")
    dom = w["domain"]
    pl.setup(dom, bits=bits, solver=0, tol=1e-10)
    nel = (dom.nnx - 1) * (dom.nny - 1)
    weight = pkg.species_weight(dom, 8)
    masses = {0: pkg.ELECTRON["mass_kg"], 1: pkg.ARGON_ION["mass_kg"]}
    temps = {0: 2.0, 1: 0.026}
    pl.set_species(0, pkg.ELECTRON["charge_C"], masses[0], weight, temps[0] * pkg.EVTOK, nel * 16)
    pl.set_species(1, pkg.ARGON_ION["charge_C"], masses[1], weight, temps[1] * pkg.EVTOK, nel * 16)
    pl.load_maxwellian(0, 8, temps[0], 5)
    pl.load_maxwellian(1, 8, temps[1], 6)
    pl.set_run_params(20, 77, 0)
    norm = dom.normalisation()
    y_lo, y_hi = 0.0, float(dom.ymax)
    src = {}
    for s in (0, 1):
        xm = masses[s] / pkg.ELECTRON_MASS
        tmpj = temps[s] * pkg.EVTOK / (dom.T_ref_eV * pkg.EVTOK)
        vt = np.sqrt(tmpj / xm)
        drift = (0.1 * vt * norm["v_ref"] * (1 if s == 0 else 20), 0.02 * norm["v_ref"], -0.01 * norm["v_ref"])
        vd = [d / norm["v_ref"] for d in drift]
        c = orc.inject_consts(vd[0], vt, tmpj, xm, 0.0, dom.dt)
        dy = y_hi - y_lo
        per_dens = abs(c[0]) * norm["v_ref"] * dom.dt * norm["time_ref"] * np.pi * (dy * dy) * \
            (norm["L_ref"] * norm["L_ref"]) / weight
        dens = (3000.0 if s == 0 else 500.0) / per_dens
        pl.set_injection(s, dens, temps[s], drift, r_start=0.0, y_lo=y_lo, y_hi=y_hi)
        expect = (dens / dom.n_ref) * dom.n_ref * abs(c[0]) * norm["v_ref"] * dom.dt * norm["time_ref"] * np.pi * \
            (dy * dy) * (norm["L_ref"] * norm["L_ref"]) / weight
        src[s] = (c, vd, expect)
    before = [pl.counts(s)["nslots"] for s in (0, 1)]
    th_before = [pl.azimuth(s) for s in (0, 1)]
    added = pl.InjectBeams_2D(5, 1)
    say(f"  added {added[:2]}")
    assert added[0] > 2000 and added[1] > 300
    for s in (0, 1):
        c, vd, expect = src[s]
        tg = orc.tag(s, orc.RNG_INJECT)
        assert added[s] == orc.inject_count(expect, 77, 5, tg)
        x, y, vx, vy, vz, cell, th = orc.inject(bits, added[s], 1 if s == 0 else 0, c, vd, y_lo, y_hi, 77, 5, tg,
                                                dom.xmin, dom.ymin, 1.0, 1.0, dom.nnx - 1, dom.nny - 1, delta=1)
        p = pl.particles(s)
        n = added[s]
        assert pl.counts(s)["nslots"] == before[s] + n
        for k, ref in (("x", x), ("y", y), ("vx", vx), ("vy", vy), ("vz", vz)):
            assert np.array_equal(p[k][-n:], ref), (s, k)
        assert np.array_equal(p["loc"][-n:] - 1, cell)
        az = pl.azimuth(s)
        assert np.array_equal(az[:-n], th_before[s]) and np.array_equal(az[-n:], th)
        assert np.all(cell >= 0) and np.all(y < y_hi) and np.all((th >= 0) & (th < 2 * np.pi))
    pl.initial_field()
    pl.step(1)


# ------------------------------------------------------------------ cell sort
def _sort_keys(loc0, nex, ney, tile=16):
    """The tile-major element key of the cell sort: 16 x 16-element tiles visited column by
    column in serpentine order, elements column-major inside a tile."""
    ci, cj = loc0 // ney, loc0 % ney
    nty = (ney + tile - 1) // tile
    ti, tj = ci // tile, cj // tile
    tpos = np.where(ti & 1, nty - 1 - tj, tj)
    return (ti * nty + tpos) * tile * tile + (ci % tile) * tile + cj % tile


def _transposed(n, group):
    """Slot d holds the particle at sorted position p(d): inside a full group of 64 K slots the
    sorted position K l + k sits at slot 64 k + l; a partial last group keeps the order."""
    d = np.arange(n)
    nfull = n // group * group
    g0 = d & ~(group - 1)
    r = d - g0
    return np.where(d < nfull, g0 + (r & 63) * (group // 64) + (r >> 6), d)


@pytest.mark.parametrize("case", ["small32", "rz32", "cfg3s64"])
def test_cell_sort_is_the_stable_tile_order(gpu_lib, pkg, case):
    """The hand-written radix sort: after steps that delete particles and append MCC products,
    pic_sort_particles leaves exactly the live particles, in the stable order of their tile-major
    element keys (ties in slot order), laid out group-transposed; every array (and the azimuth)
    moves with its particle; a permutation: nothing lost or duplicated."""
    pl = gpu_lib
    say(f"case: cell sort {case}")
    w = pkg.build_workload(pl, case, seed=51, ppc=12, mcc=True, gas_density=3.2e24)
    dom = w["domain"]
    pl.set_run_params(0, 51, 1)
    pl.set_step(0)
    pl.initial_field()
    pl.step(5)
    pres = [pl.particles(s) for s in (0, 1)]
    ths = [pl.azimuth(s) if dom.delta == 1 else None for s in (0, 1)]
    c0 = pl.counts(0)
    assert c0["nslots"] > c0["nlive"]                # dead slots to drop
    # 8 slots per thread for the FP32 Cartesian case (h = 1 from the origin), else 4
    group = pl.sort_group()
    assert group == (512 if (w["bits"] == 32 and dom.delta == 0) else 256)
    pl.sort_particles()
    for s in (0, 1):
        pre, th = pres[s], ths[s]
        keys = _sort_keys(pre["loc"] - 1, dom.nnx - 1, dom.nny - 1)
        order = np.argsort(keys, kind="stable")
        pos = order[_transposed(len(order), group)]
        post = pl.particles(s)
        assert pl.counts(s)["nslots"] == len(pre["x"]) == pl.counts(s)["nlive"]
        for k in ("x", "y", "vx", "vy", "vz", "loc"):
            assert np.array_equal(post[k], pre[k][pos]), (s, k)
        if th is not None:
            assert np.array_equal(pl.azimuth(s), th[pos])
        # keys ascending in sorted order
        assert np.all(np.diff(keys[order]) >= 0)
