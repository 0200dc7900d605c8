"""Pins of oracle functions that the golden vectors leave open, each against what the physics
or the mathematics fixes independently of the restated code, and each with a planted
mistake (oracle.planted, or a variant input) that the check catches:

* PostCollisionVelocity + CosKaiVahedi (TypeMCC.f90:101-133, MCCChi.f90:13-25): elastic
  scattering keeps the centre-of-mass relative speed; the scattering angle follows the
  Vahedi-Surendra differential cross section (Comput. Phys. Commun. 87 (1995) 179),
  integrated numerically here, or is isotropic for the ion reaction 101;
* SelectProbility (TypeMCC.f90:155-198): interpolation linear in log10 E on the table's
  grid, the first row below it, the E/Emax tail above it, zero below thresholds;
* VelMaxInit (ParticleOne.f90:135-152): the rejection sampler yields E/kT ~ Gamma(3/2),
  the 3-D Maxwell-Boltzmann energy law, in isotropic directions;
* Cell_Volume_zwz (SetupGrids_2D_QLL.f90:196-395): each node's volume is its dual cell,
  revolved about the axis in r-z, and the volumes add up to the domain;
* the deposit normalisation (GetParChg_2D.f90:894-908): sum rho_s vol L^2 n_ref (L^3 in r-z)
  returns Weight qs times the deposited weight;
* axisymmetric element matrices: second-order convergence on a manufactured solution, and
  the r-z IFE element on a dielectric sphere in a uniform field (closed form);
* Get_PointPatch next to an object (Get_PointPatch.f90:140-190, 379-590): patch lists
  derived by hand."""
import numpy as np
import pytest
from scipy import integrate, stats

PI_LY = 3.14159265358979


# ------------------------------------------------------------------ MCC collision kinematics
def _rel_after(orc, g, v, energy, mr, iso, seed, slot):
    vo = orc.post_collision(g, v, energy, mr, iso, seed, slot)
    # the projectile moves by MassRatio (g' - g), g' the centre-of-mass relative velocity after
    return g + (vo - v) / mr


@pytest.mark.parametrize("iso", [0, 1])
def test_scattering_keeps_relative_speed(orc, iso):
    rng = np.random.default_rng(2)
    for k in range(300):
        g = rng.normal(size=3) * 10 ** rng.uniform(3, 7)
        v = rng.normal(size=3) * 1e5
        gp = _rel_after(orc, g, v, 10 ** rng.uniform(-2, 3), 0.37, iso, 5, k)
        assert np.linalg.norm(gp) == pytest.approx(np.linalg.norm(g), rel=1e-11)
    # g along x only (Gper = 0): the new velocity has speed |g| in a random direction
    vo = orc.post_collision(np.array([3.0e5, 0, 0]), np.zeros(3), 5.0, 0.5, iso, 5, 1)
    assert np.linalg.norm(vo) == pytest.approx(3.0e5, rel=1e-12)


def test_scattering_planted_mistake_is_caught(orc):
    g, v = np.array([1.0e5, -2.0e5, 3.0e5]), np.zeros(3)
    with orc.planted(1):
        bad = [abs(np.linalg.norm(_rel_after(orc, g, v, 10.0, 0.5, 0, 5, k)) / np.linalg.norm(g) - 1) for k in range(50)]
    assert max(bad) > 1e-3


def _cos_chi(orc, energy, iso, n, seed):
    g = np.array([2.0e5, -1.0e5, 1.5e5])
    c = np.empty(n)
    for k in range(n):
        gp = _rel_after(orc, g, np.zeros(3), energy, 0.5, iso, seed, k)
        c[k] = gp @ g / (g @ g)
    return c


def _vahedi_cdf(energy):
    """CDF of cos(chi) from the Vahedi-Surendra differential cross section
    sigma(E, chi) ~ 1 / (1 + E sin^2(chi/2)), in the variable cos(chi) (numerical quadrature)."""
    f = lambda c: 1.0 / (1.0 + energy * (1.0 - c) / 2.0)
    total = integrate.quad(f, -1.0, 1.0, epsabs=0, epsrel=1e-12)[0]
    return np.vectorize(lambda c: integrate.quad(f, -1.0, c, epsabs=0, epsrel=1e-12)[0] / total)


@pytest.mark.parametrize("energy", [1.0, 10.0, 100.0])
def test_elastic_angle_follows_vahedi(orc, energy):
    c = _cos_chi(orc, energy, 0, 6000, 17)
    assert stats.kstest(c, _vahedi_cdf(energy)).pvalue > 1e-3
    if energy == 100.0:
        # an isotropic draw is told apart (forward peaked at 100 eV)
        iso = _cos_chi(orc, energy, 1, 6000, 17)
        assert stats.kstest(iso, _vahedi_cdf(energy)).pvalue < 1e-6


def test_ion_isotropic_angle_is_uniform(orc):
    c = _cos_chi(orc, 0.5, 1, 6000, 23)
    assert stats.kstest(c, stats.uniform(-1, 2).cdf).pvalue > 1e-3


# ------------------------------------------------------------------ SelectProbility
def _linear_table(nr, ns):
    i = np.arange(ns)[:, None]
    a = np.array([0.05, 0.2, 0.3])[:nr]
    b = np.array([0.004, -0.001, 0.002])[:nr]
    return (a + b * i).ravel()                     # Probility(1:nr, i) linear in the row index


def test_select_probability_interpolation(orc):
    nr, ns, emin, eint = 3, 21, -1.0, 0.2
    emax = 10 ** (emin + eint * (ns - 1))
    prob = _linear_table(nr, ns)
    thr = np.zeros(nr)
    rows = prob.reshape(ns, nr)
    rng = np.random.default_rng(3)
    for lx in rng.uniform(0.0, ns - 1.0, 200):   # off the grid: exact linear interpolation in log10 E
        P = orc.select_prob(10 ** (emin + lx * eint), prob, thr, nr, ns, emin, eint, emax)
        np.testing.assert_allclose(P, rows[0] + (rows[1] - rows[0]) * lx, rtol=1e-11, atol=1e-14)
    for i in (1, 7, 20):                          # on the grid: the row itself
        P = orc.select_prob(10 ** (emin + i * eint), prob, thr, nr, ns, emin, eint, emax)
        np.testing.assert_allclose(P, rows[i], rtol=1e-9)
    np.testing.assert_array_equal(orc.select_prob(10 ** (emin - 0.5), prob, thr, nr, ns, emin, eint, emax), rows[0])
    E = 3.0 * emax                                # the tail: last row times E / EnergyMax
    np.testing.assert_allclose(orc.select_prob(E, prob, thr, nr, ns, emin, eint, emax), rows[-1] * 3.0, rtol=1e-14)
    thr = np.array([0.0, 2.0, 5.0])               # below a threshold the reaction is off
    P = orc.select_prob(3.0, prob, thr, nr, ns, emin, eint, emax)
    assert P[0] > 0 and P[1] > 0 and P[2] == 0.0
    with orc.planted(2):
        P = orc.select_prob(10 ** (emin + 4.5 * eint), prob, np.zeros(nr), nr, ns, emin, eint, emax)
    assert not np.allclose(P, rows[0] + (rows[1] - rows[0]) * 4.5, rtol=1e-6)


# ------------------------------------------------------------------ VelMaxInit
def test_target_velocity_is_maxwellian(orc):
    mass, T, n = 6.6335e-26, 300.0, 40000
    kT = 1.3807e-23 * T
    v = orc.vel_maxwell(7, n, mass, T)
    e = 0.5 * mass * (v ** 2).sum(1) / kT
    assert stats.kstest(e, stats.gamma(1.5).cdf).pvalue > 1e-3
    assert e.mean() == pytest.approx(1.5, rel=0.02)
    assert stats.kstest(v[:, 0] / np.linalg.norm(v, axis=1), stats.uniform(-1, 2).cdf).pvalue > 1e-3
    with orc.planted(3):
        vb = orc.vel_maxwell(7, n, mass, T)
    eb = 0.5 * mass * (vb ** 2).sum(1) / kT
    assert stats.kstest(eb, stats.gamma(1.5).cdf).pvalue < 1e-6


# ------------------------------------------------------------------ cell volumes and normalisation
def _dual_volume(xs, ys, delta):
    """The dual cell of node (i, j): [x_i - hx/2, x_i + hx/2] x [y_j - hy/2, y_j + hy/2] within
    the domain; its area, or in r-z the volume it sweeps about the axis (Pi_LY as the reference)."""
    out = np.zeros(len(xs) * len(ys))
    for i, x in enumerate(xs):
        xl = max(xs[0], x - (xs[1] - xs[0]) / 2)
        xr = min(xs[-1], x + (xs[1] - xs[0]) / 2)
        for j, y in enumerate(ys):
            yl = max(ys[0], y - (ys[1] - ys[0]) / 2)
            yr = min(ys[-1], y + (ys[1] - ys[0]) / 2)
            out[i * len(ys) + j] = (xr - xl) * ((yr - yl) if delta == 0 else PI_LY * (yr * yr - yl * yl))
    return out


@pytest.mark.parametrize("delta,ymin", [(0, -1.5), (1, 0.0), (1, 2.0)])
def test_cell_volumes_are_dual_cells(orc, delta, ymin):
    xs = 0.3 + 0.5 * np.arange(8)
    ys = ymin + 0.25 * np.arange(6)
    vol = orc.cell_volume(xs, ys, delta)
    np.testing.assert_allclose(vol, _dual_volume(xs, ys, delta), rtol=1e-13)
    total = (xs[-1] - xs[0]) * ((ys[-1] - ys[0]) if delta == 0 else PI_LY * (ys[-1] ** 2 - ys[0] ** 2))
    assert vol.sum() == pytest.approx(total, rel=1e-13)
    if delta == 1:
        with orc.planted(4):
            bad = orc.cell_volume(xs, ys, delta)
        assert not np.allclose(bad, _dual_volume(xs, ys, delta), rtol=1e-6)


@pytest.mark.parametrize("delta", [0, 1])
def test_normalisation_returns_the_deposited_charge(orc, delta):
    xs = np.arange(9, dtype=float)
    ys = np.arange(7, dtype=float)
    rng = np.random.default_rng(5)
    m = 3000
    x, y = rng.random(m) * 8, rng.random(m) * 6
    v = rng.normal(size=(3, m))
    cell, _ = orc.locate(xs, ys, x, y)
    w, e = orc.deposit(64, xs, ys, delta, x, y, *v, cell, 9.1095e-31, 4.19e5, fix=True)
    vol = orc.cell_volume(xs, ys, delta)
    weight, qs, L, nref = 3.7e5, -1.0, 2.35e-7, 1e21
    rho = np.zeros_like(w)
    rs, ek = orc.normalize(delta, weight, qs, L, nref, vol, w, e, rho)
    scale = L ** (2 if delta == 0 else 3) * nref
    assert (rs * vol * scale).sum() == pytest.approx(weight * qs * w.sum(), rel=1e-12)
    assert (ek * vol * scale).sum() == pytest.approx(weight * e.sum(), rel=1e-12)
    np.testing.assert_array_equal(rho, rs)
    if delta == 0:
        assert w.sum() == pytest.approx(m, rel=1e-12)


# ------------------------------------------------------------------ axisymmetric element matrices
def _rz_manufactured(orc, n, delta_used=1):
    """-(1/r) d/dr (r du/dr) - d2u/dx2 = f on [0, 2] x [0, 1] (x axial, r radial) with
    u = (1 - r^2) sin(pi x / 2): Dirichlet at x = 0, 2 and r = 1, natural on the axis.
    Returns the maximum nodal error."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    xs, ys = np.linspace(0, 2, 2 * n + 1), np.linspace(0, 1, n + 1)
    nnx, nny = len(xs), len(ys)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    u = ((1 - Y ** 2) * np.sin(np.pi * X / 2)).ravel()
    f = (np.sin(np.pi * X / 2) * (4.0 + (np.pi / 2) ** 2 * (1 - Y ** 2))).ravel()
    dirich = ((X == 0) | (X == 2) | (Y == 1)).ravel()
    rows, cols, kv, mv = [], [], [], []
    for ci in range(nnx - 1):
        for cj in range(nny - 1):
            K, M = orc.element_matrices(delta_used, xs[ci], ys[cj], xs[ci + 1] - xs[ci], ys[cj + 1] - ys[cj])
            nd = [ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1]
            for a in range(4):
                for b in range(4):
                    rows.append(nd[a]); cols.append(nd[b]); kv.append(K[a, b]); mv.append(M[a, b])
    N = nnx * nny
    K = sp.csr_matrix((kv, (rows, cols)), shape=(N, N))
    M = sp.csr_matrix((mv, (rows, cols)), shape=(N, N))
    free = ~dirich
    b = M @ f - K[:, dirich] @ u[dirich]
    sol = spl.spsolve(K[free][:, free].tocsc(), b[free])
    return np.abs(sol - u[free]).max()


def test_axisymmetric_stiffness_converges_second_order(orc):
    err = [_rz_manufactured(orc, n) for n in (8, 16, 32)]
    assert err[1] < err[0] / 3.2 and err[2] < err[1] / 3.2, err
    # planted: the Cartesian matrices (no r weight) solve a different equation
    assert _rz_manufactured(orc, 32, delta_used=0) > 10 * err[2]


def _sphere_problem(orc, n, eps, r0=0.5031, mistake=None):
    """A dielectric sphere (relative permittivity eps, radius r0, centred on the axis) in the
    uniform axial field E = -grad(-x): phi = -3/(eps + 2) x inside, -x + (eps - 1)/(eps + 2)
    r0^3 x / rho^3 outside (rho the distance from the centre). Domain [-1, 1] x [0, 1] in (x, r),
    Dirichlet data from phi at x = +-1 and r = 1, natural on the axis. Nodal RMS error."""
    import scipy.sparse.linalg as spl
    xs, ys = np.linspace(-1.0, 1.0, 2 * n + 1), np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    rho = np.hypot(X, Y).ravel()
    xr = X.ravel()
    k = (eps - 1) / (eps + 2)
    outside = -xr + k * r0 ** 3 * xr / np.where(rho < r0, 1.0, rho) ** 3
    exact = np.where(rho < r0, -3.0 / (eps + 2) * xr, outside)
    bnd = ((X == -1) | (X == 1) | (Y == 1)).ravel()
    nt2 = np.zeros(len(rho), dtype=np.int32)
    nt2[~bnd] = np.arange(1, (~bnd).sum() + 1)
    nt2[bnd] = -1
    A, c, B, types = orc.assemble_ife(xs, ys, 1, [(0.0, 0.0, r0, eps, False)], nt2, np.where(bnd, exact, 0.0),
                                      mistake=mistake)
    err = spl.spsolve(A.tocsc(), -c) - exact[~bnd]
    return np.sqrt(np.mean(err ** 2)), types


@pytest.mark.parametrize("eps", [4.0, 0.25])
def test_axisymmetric_ife_dielectric_sphere(orc, eps):
    """The nodal RMS error falls on every halving of h (second order overall: 16 -> 128 gains
    about 20x), like the Cartesian circular-interface problem; plain Q1 on the cut elements or
    the permittivities swapped there leave an error many times larger."""
    rms = []
    for n in (16, 32, 64, 128):
        r, types = _sphere_problem(orc, n, eps)
        assert set(types) <= {1, 2} and types
        rms.append(r)
    assert all(rms[i + 1] < rms[i] / 1.8 for i in range(3)), rms
    assert rms[3] < rms[0] / 15, rms
    if eps == 4.0:
        for mistake in ("q1_cut", "swap"):
            assert _sphere_problem(orc, 64, eps, mistake=mistake)[0] > 10 * rms[2], mistake


# ------------------------------------------------------------------ patches next to an object
def _patches(n, circles, treat_interface_as_vacuum=False):
    from patch_ref import element_and_node_classes, point_patches
    xs = np.arange(n + 1, dtype=float)
    ni, ei = element_and_node_classes(n + 1, n + 1, xs, xs, circles)
    if treat_interface_as_vacuum:
        ei = np.where(ei > 0, -1, ei).astype(np.int32)
    pp, ptype = point_patches(n + 1, n + 1, ei)
    nn = (n + 1) ** 2
    pp = pp.reshape(50, nn).T                         # point_patch(nnodes, 50), Fortran order
    nid = lambda i, j: i * (n + 1) + j + 1            # 1-based node id of node (i, j)
    get = lambda i, j: sorted(pp[nid(i, j) - 1, 1:1 + pp[nid(i, j) - 1, 0]].tolist())
    block = lambda i0, i1, j0, j1: sorted(nid(i, j) for i in range(i0, i1 + 1) for j in range(j0, j1 + 1))
    return get, block, ptype, nid


def test_patch_lists_next_to_an_object_hand_derived():
    """8 x 8 unit mesh, circle at (4, 4) of radius 1.5: nodes (3..5, 3..5) inside, elements
    (3..4, 3..4) inside (-2), the ring of 12 around them cut (interface ids > 0).
    * node (2, 4), vacuum, left of the object: elements (1,3), (1,4) vacuum, (2,3), (2,4) cut ->
      type 3 with the two vacuum elements (6 nodes, i 1..2, j 3..5); fewer than 12 nodes, so the
      boundary treatment unites the unextended patches of those 6 nodes: (1,3), (1,4), (1,5)
      are 3 x 3 interior blocks, (2,3), (2,5) are type 3 pairs of the vacuum column -> i 0..2,
      j 2..6, 15 nodes;
    * node (3, 4), inside on the object's edge: elements (3,3), (3,4) inside, (2,3), (2,4) cut ->
      type 3 with the two inside elements (i 3..4, j 3..5); the union of its nodes' patches is
      the 3 x 3 block i, j 3..5;
    * radius 0.9: only node (4, 4) inside and all four of its elements cut: no element
      qualifies, the patch is empty (the reference's recovery of such nodes is commented out,
      GetEfield_SIDG_PPR.f90:255-296), so its coefficients stay 0."""
    get, block, ptype, nid = _patches(8, [(4.0, 4.0, 1.5)])
    assert ptype[nid(2, 4) - 1] == 3 and ptype[nid(3, 4) - 1] == 3 and ptype[nid(4, 4) - 1] == 1
    assert get(2, 4) == block(0, 2, 2, 6)
    assert get(3, 4) == block(3, 5, 3, 5)
    assert get(4, 4) == block(3, 5, 3, 5)
    assert get(1, 4) == block(0, 2, 3, 5)          # interior vacuum node: its own 3 x 3
    get2, _, ptype2, nid2 = _patches(8, [(4.0, 4.0, 0.9)])
    assert ptype2[nid2(4, 4) - 1] == 3 and get2(4, 4) == []
    # planted: interface elements taken as vacuum give the plain 3 x 3 patches instead
    getb, blockb, _, _ = _patches(8, [(4.0, 4.0, 1.5)], treat_interface_as_vacuum=True)
    assert getb(2, 4) == blockb(1, 3, 3, 5) != get(2, 4)


# ------------------------------------------------------------------ axisymmetric wall injection
def test_axisymmetric_injection_fills_the_disk_uniformly(orc):
    """InjectBeams_2D on an axisymmetric domain (PosInit, ParticleOne.f90:95-98): the injected
    particles are uniform over the annulus y_lo <= r <= y_hi in area, i.e. r^2 is uniform, and
    the azimuth is uniform on [0, 2 pi); planted: r drawn uniform in r is told apart."""
    c = orc.inject_consts(0.05, 0.5, 0.25, 1.0, 0.0, 0.05)
    y_lo, y_hi, n = 0.5, 6.0, 20000
    out = orc.inject(64, n, 1, c, [0.05, 0.0, 0.0], y_lo, y_hi, 9, 3, orc.tag(0, orc.RNG_INJECT), 0.0, 0.0, 1.0,
                     1.0, 8, 8, delta=1)
    r, th = out[1], out[6]
    u = (r ** 2 - y_lo ** 2) / (y_hi ** 2 - y_lo ** 2)
    assert stats.kstest(u, stats.uniform(0, 1).cdf).pvalue > 1e-3
    assert stats.kstest(th, stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3
    assert r.min() >= y_lo and r.max() < y_hi
    with orc.planted(5):
        rb = orc.inject(64, n, 1, c, [0.05, 0.0, 0.0], y_lo, y_hi, 9, 3, orc.tag(0, orc.RNG_INJECT), 0.0, 0.0, 1.0,
                        1.0, 8, 8, delta=1)[1]
    ub = (rb ** 2 - y_lo ** 2) / (y_hi ** 2 - y_lo ** 2)
    assert stats.kstest(ub, stats.uniform(0, 1).cdf).pvalue < 1e-6


# ------------------------------------------------------------------ kappa and polytropic initial loads
def kappa_load_cdf(s, kappa=2.0):
    """CDF of the kappa load's Vx: density (1 + V^2 / ((2 kappa - 3) s^2))^-kappa on |V| <= 10 s
    (numerical quadrature)."""
    f = lambda V: (1.0 + V * V / ((2 * kappa - 3) * s * s)) ** (-kappa)
    tot = integrate.quad(f, -10 * s, 10 * s, epsabs=0, epsrel=1e-12, limit=200)[0]
    return np.vectorize(lambda V: integrate.quad(f, -10 * s, min(max(V, -10 * s), 10 * s), epsabs=0, epsrel=1e-12,
                                                 limit=200)[0] / tot)


def load_scale(T, M, kind):
    """s = sqrt(Sigma / Sigma_e) of ParticleOne.f90:167-168 (kappa) and :216-217 (polytropic)."""
    kB, Te, Me = 1.3807e-23, 11605.0, 9.1095e-31
    return np.sqrt((2.0 * kB * T / M) / (2.0 * kB * Te / Me))


def test_kappa_and_polytropic_loads(orc):
    """VelKappaInit draws Vx from the kappa = 2 law (a scaled Student-t with 3 degrees of
    freedom, cut at 10 s) and VelPolyInit (gamma = 3) uniform on +-sqrt(3) s, Vy = Vz = 0;
    planted: the acceptance density with half its exponent is told apart."""
    for T, M in ((23210.0, 9.1095e-31), (301.7, 6.69e-26)):
        s = load_scale(T, M, 0)
        v = orc.load_velocity(0, 5, 20000, T, M)
        assert np.all(v[:, 1:] == 0.0)
        assert stats.kstest(v[:, 0], kappa_load_cdf(s)).pvalue > 1e-3
        p = orc.load_velocity(2, 5, 20000, T, M)[:, 0]
        assert stats.kstest(p, stats.uniform(-np.sqrt(3) * s, 2 * np.sqrt(3) * s).cdf).pvalue > 1e-3
    with orc.planted(6):
        vb = orc.load_velocity(0, 5, 20000, 23210.0, 9.1095e-31)[:, 0]
    assert stats.kstest(vb, kappa_load_cdf(load_scale(23210.0, 9.1095e-31, 0))).pvalue < 1e-6
