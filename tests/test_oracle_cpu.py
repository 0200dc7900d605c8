"""CPU tests: pin the oracle against published vectors, the reference's own fixture
files (copied as data under tests/golden/) and what the mathematics fixes; check the
host logic of the package and that the C-ABI library exports every declared symbol."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ Philox
def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32_10
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, out in kat:
        assert tuple(int(v) for v in orc.philox(ctr, key)) == out


# ------------------------------------------------------------------ normalisation
def test_normalisation_matches_reference_output(pkg):
    """OUTPUT/physics_parameter.inp:40-50 of the reference (n_ref=1e21, T_ref=1 eV)."""
    gold = json.loads((GOLD / "physics_parameter.json").read_text())
    d = pkg.Domain(3, 3, n_ref=1e21, T_ref_eV=1.0)
    n = d.normalisation()
    for k in ("L_ref", "time_ref", "v_ref", "phi_ref"):
        assert n[k] == pytest.approx(gold[k], rel=1e-14), k


# ------------------------------------------------------------------ BRINV
def test_brinv_matches_numpy_inverse(orc):
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.normal(size=(6, 6))
        inv, ok = orc.brinv(a)
        assert ok == 1
        np.testing.assert_allclose(inv, np.linalg.inv(a), rtol=1e-10, atol=1e-12)
    # a matrix that needs row and column pivoting (zero leading diagonal)
    a = np.eye(6)[[3, 0, 5, 1, 4, 2]] * np.arange(1, 7)
    inv, ok = orc.brinv(a)
    np.testing.assert_allclose(inv @ a, np.eye(6), atol=1e-14)
    # singular
    _, ok = orc.brinv(np.ones((6, 6)))
    assert ok == 0


# ------------------------------------------------------------------ element matrices
def test_element_matrices_closed_form(orc):
    hx, hy = 1.0, 0.5
    K, M = orc.element_matrices(0, 0.0, 0.0, hx, hy)
    Mx = np.array([[2, -2, -1, 1], [-2, 2, 1, -1], [-1, 1, 2, -2], [1, -1, -2, 2]]) / 6.0
    My = np.array([[2, 1, -1, -2], [1, 2, -2, -1], [-1, -2, 2, 1], [-2, -1, 1, 2]]) / 6.0
    np.testing.assert_allclose(K, hy / hx * Mx + hx / hy * My, rtol=1e-14, atol=1e-15)
    P = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) * hx * hy / 36.0
    np.testing.assert_allclose(M, P, rtol=1e-14)
    # rows of K sum to zero (constants are in the kernel); M sums to the area
    assert np.abs(K.sum(1)).max() < 1e-14 and M.sum() == pytest.approx(hx * hy)
    # axisymmetric: M integrates r over the element
    K1, M1 = orc.element_matrices(1, 0.0, 2.0, 1.0, 1.0)
    assert M1.sum() == pytest.approx(2.5)


# ------------------------------------------------------------------ patches and PPR
def _mesh(n, circles=()):
    from patch_ref import element_and_node_classes, point_patches
    xs = np.arange(n + 1, dtype=np.float64)
    ys = np.arange(n + 1, dtype=np.float64)
    ni, ei = element_and_node_classes(n + 1, n + 1, xs, ys, circles)
    pp, ptype = point_patches(n + 1, n + 1, ei)
    return xs, ys, ni, ei, pp, ptype


def test_patch_sizes_hand_derived():
    n = 6
    xs, ys, ni, ei, pp, ptype = _mesh(n)
    nn = (n + 1) ** 2
    cnt = pp[:nn]
    nid = lambda i, j: i * (n + 1) + j
    assert cnt[nid(0, 0)] == 9            # corner: 3x3 block at the corner
    assert cnt[nid(0, 3)] == 15           # left edge: 3 columns x 5 rows
    assert cnt[nid(0, 1)] == 12           # next to a corner: 3 x 4
    assert cnt[nid(3, 3)] == 9            # interior: 3x3
    assert cnt[nid(3, n)] == 15           # top edge


def test_ppr_reproduces_quadratics_exactly(orc):
    """Polynomial preservation: for phi quadratic, every patch fit is exact, so coef_ppr holds
    the Taylor coefficients at each node and Get_ParticleEField_PPR returns grad phi."""
    n = 8
    xs, ys, ni, ei, pp, ptype = _mesh(n, circles=[(4.0, 4.0, 1.6)])
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    a = [0.3, -1.2, 0.7, 0.25, -0.4, 0.15]
    phi = (a[0] + a[1] * X + a[2] * Y + a[3] * X * X + a[4] * X * Y + a[5] * Y * Y).ravel()
    coef, efx, efy, ns = orc.ppr(xs, ys, pp, phi)
    assert ns == 0
    coef = coef.reshape(-1, 6)
    Xf, Yf = X.ravel(), Y.ravel()
    has = pp[: len(Xf)] > 0
    np.testing.assert_allclose(coef[has, 1], (a[1] + 2 * a[3] * Xf + a[4] * Yf)[has], atol=1e-11)
    np.testing.assert_allclose(coef[has, 2], (a[2] + a[4] * Xf + 2 * a[5] * Yf)[has], atol=1e-11)
    np.testing.assert_allclose(coef[has, 3], a[3], atol=1e-11)
    np.testing.assert_allclose(coef[has, 4], a[4], atol=1e-11)
    np.testing.assert_allclose(coef[has, 5], a[5], atol=1e-11)
    np.testing.assert_allclose(efx[has], -coef[has, 1])
    # point gather on vacuum elements returns grad phi exactly
    rng = np.random.default_rng(3)
    e0 = np.flatnonzero(ei == -1)[:50]
    ci, cj = np.divmod(e0, n)
    px, py = ci + rng.random(len(e0)), cj + rng.random(len(e0))
    ex, ey = orc.ppr_gather(xs, ys, ei, ni, coef.ravel(), px, py, e0)
    np.testing.assert_allclose(ex, a[1] + 2 * a[3] * px + a[4] * py, atol=1e-10)
    np.testing.assert_allclose(ey, a[2] + a[4] * px + 2 * a[5] * py, atol=1e-10)


def test_ppr_interior_closed_form(orc):
    """On a uniform interior 3x3 patch the least-squares gradient is the averaged central
    difference E_x = -(1/(6h)) sum_eta (phi[i+1,j+eta] - phi[i-1,j+eta])."""
    n = 6
    xs, ys, ni, ei, pp, _ = _mesh(n)
    rng = np.random.default_rng(5)
    phi = rng.normal(size=(n + 1) ** 2)
    coef, efx, efy, _ = orc.ppr(xs, ys, pp, phi)
    P = phi.reshape(n + 1, n + 1)
    i, j = 3, 2
    ex = -(P[i + 1, j - 1:j + 2] - P[i - 1, j - 1:j + 2]).sum() / 6.0
    ey = -(P[i - 1:i + 2, j + 1] - P[i - 1:i + 2, j - 1]).sum() / 6.0
    assert efx[i * (n + 1) + j] == pytest.approx(ex, rel=1e-12)
    assert efy[i * (n + 1) + j] == pytest.approx(ey, rel=1e-12)


# ------------------------------------------------------------------ PCG
def _poisson_system(orc, n):
    xs = np.arange(n + 1, dtype=float)
    ys = xs.copy()
    nnx = n + 1
    nt2 = np.zeros(nnx * nnx, dtype=np.int32)
    u = 0
    for i in range(nnx):
        for j in range(nnx):
            if i in (0, n) or j in (0, n):
                nt2[i * nnx + j] = -1
            else:
                u += 1
                nt2[i * nnx + j] = u
    # assemble A from the oracle's element-by-element operator (columns = unit vectors)
    A = np.zeros((u, u))
    for k in range(u):
        e = np.zeros(u)
        e[k] = 1.0
        A[:, k] = orc.stiffness_apply(xs, ys, 0, nt2, e)
    return xs, ys, nt2, A


def test_pcg_matches_direct_solve(orc):
    xs, ys, nt2, A = _poisson_system(orc, 12)
    n = A.shape[0]
    rows, cols = np.nonzero(np.abs(A) > 0)
    rowptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    val = A[rows, cols]
    b = np.random.default_rng(2).normal(size=n)
    x, r, it, ierr = orc.pcg(rowptr, cols, val, np.diag(A).copy(), b, np.zeros(n), 1e-13, 1000)
    assert ierr == 0 and it > 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-11)
    # the returned residual is the recursively updated one and is small
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-13
    # IERR = 1 at MAXITER, ITER left unset
    _, _, it2, ierr2 = orc.pcg(rowptr, cols, val, np.diag(A).copy(), b, np.zeros(n), 1e-30, 3)
    assert ierr2 == 1 and it2 == -1


def test_stiffness_interior_stencil(orc):
    xs, ys, nt2, A = _poisson_system(orc, 6)
    # Q1 Laplacian on squares: 8/3 on the diagonal, -1/3 to all eight neighbours
    row = A[A.shape[0] // 2]
    assert row.max() == pytest.approx(8.0 / 3.0)
    assert sorted(np.round(row[row != 0] * 3).astype(int).tolist()) == [-1] * 8 + [8]


def test_rhs_linear_and_mass_consistent(orc):
    n = 8
    xs = np.arange(n + 1, dtype=float)
    nnx = n + 1
    nt2 = np.array([(-1 if (i in (0, n) or j in (0, n)) else 1) for i in range(nnx) for j in range(nnx)], dtype=np.int32)
    nt2[nt2 > 0] = np.arange(1, (nt2 > 0).sum() + 1)
    nu = int(nt2.max())
    rng = np.random.default_rng(4)
    r1, r2 = rng.normal(size=nnx * nnx), rng.normal(size=nnx * nnx)
    z = np.zeros(nnx * nnx)
    b1 = orc.rhs(xs, xs, 0, nt2, z, r1, nu)
    b2 = orc.rhs(xs, xs, 0, nt2, z, r2, nu)
    b12 = orc.rhs(xs, xs, 0, nt2, z, 2.0 * r1 - r2, nu)
    np.testing.assert_allclose(b12, 2.0 * b1 - b2, atol=1e-12)
    # constant rho = 1: interior rows integrate N_i = h^2
    b1 = orc.rhs(xs, xs, 0, nt2, z, np.ones(nnx * nnx), nu)
    np.testing.assert_allclose(b1, 1.0, rtol=1e-14)


# ------------------------------------------------------------------ push / boundaries
def _uniform_field_case(orc, bits, n_part=200, E=(0.3, -0.2), seed=11):
    n = 8
    xs, ys, ni, ei, pp, _ = _mesh(n)
    nn = (n + 1) ** 2
    efx, efy = np.full(nn, E[0]), np.full(nn, E[1])
    coef5 = np.zeros(nn * 5)
    rng = np.random.default_rng(seed)
    x = 2.0 + 4.0 * rng.random(n_part)
    y = 2.0 + 4.0 * rng.random(n_part)
    vx, vy, vz = rng.normal(size=(3, n_part)) * 0.5
    cell, bad = orc.locate(xs, ys, x, y)
    assert bad == 0
    return xs, ys, ni, ei, efx, efy, coef5, x, y, vx, vy, vz, cell


@pytest.mark.parametrize("bits", [64, 32])
def test_push_uniform_field_closed_form(orc, bits):
    xs, ys, ni, ei, efx, efy, coef5, x, y, vx, vy, vz, cell = _uniform_field_case(orc, bits)
    qm, dt = -1.0, 0.05
    ft = np.float64 if bits == 64 else np.float32
    out = orc.move_all(bits, xs, ys, ei, ni, np.zeros((0, 6)), efx, efy, coef5, qm, dt, 4e5, 9.1e-31, 2 * 11605.0,
                       1, 0, 0, x, y, vx, vy, vz, cell)
    X, Y, VX, VY, VZ, alive, cnt, err = out
    assert err == 0 and alive.all() and cnt.sum() == 0
    # v += qm E dt and x += v dt in the working precision, with the weights summing to one
    ev = ft(qm) * ft(0.3) * ft(dt)
    np.testing.assert_allclose(VX, vx.astype(ft) + ev, rtol=0, atol=4 * np.finfo(ft).eps)
    np.testing.assert_allclose(X, x.astype(ft) + ft(dt) * VX, rtol=0, atol=16 * np.finfo(ft).eps)
    np.testing.assert_array_equal(VZ, vz.astype(ft))


def test_push_walls(orc):
    n = 8
    xs, ys, ni, ei, pp, _ = _mesh(n)
    nn = (n + 1) ** 2
    z = np.zeros(nn)
    # particles heading to the right wall, the bottom wall, the top wall and the left wall
    x = np.array([7.9, 4.0, 4.0, 0.05])
    y = np.array([4.0, 0.02, 7.97, 4.0])
    vx = np.array([4.0, 0.0, 0.0, -2.0])
    vy = np.array([0.0, -1.0, 1.0, 0.5])
    vz = np.array([0.0, 0.0, 0.0, 0.3])
    cell, _ = orc.locate(xs, ys, x, y)
    v_ref = 419396.657327663
    X, Y, VX, VY, VZ, alive, cnt, err = orc.move_all(64, xs, ys, ei, ni, np.zeros((0, 6)), z, z, np.zeros(5 * nn), -1.0,
                                                     0.05, v_ref, 9.1095e-31, 2 * 11605.0, 3, 0, orc.tag(0, 2),
                                                     x, y, vx, vy, vz, cell)
    assert err == 0
    assert alive.tolist() == [0, 1, 1, 1]                       # right wall absorbs (X = -2000, deleted)
    assert X[0] == -2000.0 and cnt[1] == 1 and cnt[7] == 1
    # bottom: specular reflection, y re-moved from dymin + 10E-6 with the remaining time
    t_rem = (0.02 - 0.05 - 0.0) / -1.0
    assert VY[1] == 1.0 and Y[1] == pytest.approx(float(np.float32(1e-5)) + t_rem * 1.0, abs=1e-15)
    assert VY[2] == -1.0 and cnt[3] == 2
    # left wall: kappa re-emission, Vy and Vz divided by v_ref again (VelRes of all three)
    assert VX[3] > 0 and cnt[0] == 1
    assert VY[3] == pytest.approx(0.5 / v_ref) and VZ[3] == pytest.approx(0.3 / v_ref)


def test_push_object_absorbs(orc):
    n = 8
    xs, ys, ni, ei, pp, _ = _mesh(n, circles=[(4.0, 4.0, 1.5)])
    nn = (n + 1) ** 2
    z = np.zeros(nn)
    x, y = np.array([2.0, 2.0]), np.array([4.0, 7.0])
    vx, vy, vz = np.array([30.0, 30.0]), np.zeros(2), np.zeros(2)
    cell, _ = orc.locate(xs, ys, x, y)
    obj = np.array([[1, 4.0, 4.0, 0, 0, 2.25]])
    out = orc.move_all(64, xs, ys, ei, ni, obj, z, z, np.zeros(5 * nn), -1.0, 0.05, 4e5, 9.1e-31, 2 * 11605.0,
                       1, 0, 0, x, y, vx, vy, vz, cell)
    assert out[5].tolist() == [0, 1] and out[6][2] == 1


# ------------------------------------------------------------------ lookup
def test_locate_brute_force(orc):
    n = 5
    xs = np.arange(n + 1, dtype=float) * 2.0     # hx = 2: power of two
    ys = np.arange(n + 1, dtype=float) * 2.0
    rng = np.random.default_rng(9)
    x = rng.random(500) * 10.0
    y = rng.random(500) * 10.0
    # exact grid points and the far walls
    x = np.concatenate([x, [0.0, 2.0, 4.0, 10.0, 10.0, 3.0]])
    y = np.concatenate([y, [0.0, 2.0, 5.0, 10.0, 3.0, 10.0]])
    e, bad = orc.locate(xs, ys, x, y)
    assert bad == 0
    i = np.minimum(np.floor(x / 2.0).astype(int), n - 1)
    j = np.minimum(np.floor(y / 2.0).astype(int), n - 1)
    np.testing.assert_array_equal(e, j + i * n)


# ------------------------------------------------------------------ deposition
@pytest.mark.parametrize("bits", [64, 32])
def test_deposit_conservation_and_fixed_point(orc, bits):
    n = 6
    xs = np.arange(n + 1, dtype=float)
    rng = np.random.default_rng(4)
    m = 3000
    x, y = rng.random(m) * n, rng.random(m) * n
    v = rng.normal(size=(3, m))
    cell, _ = orc.locate(xs, xs, x, y)
    wq, eq = orc.deposit(bits, xs, xs, 0, x, y, *v, cell, 9.1095e-31, 4.19e5, fix=True)
    wf, ef = orc.deposit(bits, xs, xs, 0, x, y, *v, cell, 9.1095e-31, 4.19e5, fix=False)
    assert wq.sum() == pytest.approx(m, rel=1e-12)                    # weights sum to one per particle
    np.testing.assert_allclose(wq, wf, rtol=1e-12, atol=1e-11)       # fixed point == FP64 sums
    np.testing.assert_allclose(eq, ef, rtol=1e-11, atol=1e-10)
    # order independence of the fixed-point sums
    perm = rng.permutation(m)
    wp, ep = orc.deposit(bits, xs, xs, 0, x[perm], y[perm], *(v[:, perm]), cell[perm], 9.1095e-31, 4.19e5, fix=True)
    assert np.array_equal(wp, wq) and np.array_equal(ep, eq)


def test_deposit_particle_on_node(orc):
    xs = np.arange(5, dtype=float)
    w, e = orc.deposit(64, xs, xs, 0, np.array([2.0]), np.array([3.0]), np.zeros(1), np.zeros(1), np.zeros(1),
                       np.array([2 * 4 + 3 - 0]), 1.0, 1.0, fix=True)
    # the particle sits on the lower-left node of element (i=3, j=4) in 1-based terms
    assert w[2 * 5 + 3] == 1.0 and w.sum() == 1.0


def test_deposit_extreme_energies_exact(orc):
    """The heavy tail of the kappa re-emission (energies 0.5 kT (1/R - 1), up to ~1e16 eV for
    the smallest R): terms of weight x energy >= 2^11 eV keep their whole-eV part apart, so
    the fixed-point node sums equal the FP64 sums to rounding and stay order independent
    (one int64 at 2^40 holds only 2^23 eV)."""
    xs = np.arange(5, dtype=float)
    rng = np.random.default_rng(9)
    m = 400
    x, y = rng.random(m) * 4, rng.random(m) * 4
    v = rng.normal(size=(3, m))
    v[0, [3, 50, 51, 200]] = [3.0e3, 4.0e5, 2.0e7, 9.0e7]      # 4.5 MeV ... 2.6e15 eV
    cell, _ = orc.locate(xs, xs, x, y)
    wq, eq = orc.deposit(64, xs, xs, 0, x, y, *v, cell, 9.1095e-31, 4.19e5, fix=True)
    wf, ef = orc.deposit(64, xs, xs, 0, x, y, *v, cell, 9.1095e-31, 4.19e5, fix=False)
    assert ef.max() > 2.0 ** 40
    np.testing.assert_allclose(eq, ef, rtol=1e-13, atol=1e-10)
    perm = rng.permutation(m)
    wp, ep = orc.deposit(64, xs, xs, 0, x[perm], y[perm], *(v[:, perm]), cell[perm], 9.1095e-31, 4.19e5, fix=True)
    assert np.array_equal(ep, eq)
    # planted: one 2^40-scale word per term (no whole-eV part) wraps on these energies
    q = np.array([int(np.rint(t * 2.0 ** 40)) & (2 ** 64 - 1) for t in [ef.max()]])
    wrapped = (q[0] - 2 ** 64 if q[0] >= 2 ** 63 else q[0]) * 2.0 ** -40
    assert not np.isclose(wrapped, ef.max(), rtol=1e-6)


# ------------------------------------------------------------------ MCC
def test_mcc_table_cumulative(orc):
    ns, nr = 101, 3
    sig = np.abs(np.random.default_rng(0).normal(size=(ns, nr))) * 1e-20
    prob, smax = orc.mcc_table(nr, ns, -3.0, 0.06, 9.1e-31, 6.6e-26, 3e20, sig.ravel())
    P = prob.reshape(ns, nr)
    assert P.max() == pytest.approx(1.0) and np.all(np.diff(P, axis=1) >= 0)
    E = 10 ** (-3.0 + 0.06 * np.arange(ns))
    miu = 9.1e-31 * 6.6e-26 / (9.1e-31 + 6.6e-26)
    nu = (sig.sum(1) * np.sqrt(2 * E * 1.6022e-19 / miu) * 3e20)
    assert smax == pytest.approx(nu.max(), rel=1e-12)
    np.testing.assert_allclose(P[:, -1], nu / smax, rtol=1e-12)


def test_mcc_charge_exchange_and_excitation_kinematics(orc):
    # one-reaction tables forcing each event
    ns = 11
    prob = np.ones(ns)
    v = np.array([1e-3, -2e-3, 0.5e-3])
    # charge exchange: the ion takes the gas target velocity
    idx, vout, _, _ = orc.mcc_collide(0, 5, 9, 1, orc.tag(1, 4), 4e5, 6.6e-26, 6.6e-26, 300.0, 1, ns, -1.0, 0.4, 1e3,
                                      prob, [103], [0.0], v)
    assert idx == 1 and np.linalg.norm(vout * 4e5) < 2000.0   # thermal speed of 300 K argon
    # excitation of a fast electron: |v| drops by the threshold energy in the centre of mass
    ve = np.array([8.0, 0.0, 0.0])   # normalised (~32 eV)
    idx, vout, _, _ = orc.mcc_collide(1, 7, 9, 1, orc.tag(0, 4), 4.19396657e5, 9.1095e-31, 6.69e-26, 300.0, 1, ns,
                                      -3.0, 0.6, 1e3, prob, [12], [11.55], ve)
    e_in = 0.5 * 9.1095e-31 * (8.0 * 4.19396657e5) ** 2 / 1.6022e-19
    e_out = 0.5 * 9.1095e-31 * np.sum((vout * 4.19396657e5) ** 2) / 1.6022e-19
    assert e_in - e_out == pytest.approx(11.55, abs=0.05)
    # ionisation creates a secondary electron whose energy completes the budget
    idx, vout, ne, ni = orc.mcc_collide(1, 8, 9, 1, orc.tag(0, 4), 4.19396657e5, 9.1095e-31, 6.69e-26, 300.0, 1, ns,
                                        -3.0, 0.6, 1e3, prob, [13], [15.76], ve)
    e_out = 0.5 * 9.1095e-31 * np.sum((vout * 4.19396657e5) ** 2) / 1.6022e-19
    e_sec = 0.5 * 9.1095e-31 * np.sum((ne * 4.19396657e5) ** 2) / 1.6022e-19
    assert e_in - e_out - e_sec == pytest.approx(15.76, abs=0.05)
    assert np.linalg.norm(ni * 4.19396657e5) < 2000.0


def test_mcc_selection_rate(orc):
    n = 200000
    cell = np.zeros(n, dtype=np.int32)
    thr = int(0.01 * 2 ** 32)
    sel = orc.mcc_select(cell, thr, 5, 3, orc.tag(0, 3))
    assert abs(len(sel) - 2000) < 5 * np.sqrt(2000)
    assert np.all(np.diff(sel) > 0)


def test_mcc_selection_stream(orc):
    """Slot s draws word s % 4 of the Philox block with counter (s // 4, step, tag, 0):
    brute force over the blocks, with dead slots (cell < 0) never selected."""
    n = 1003
    rng = np.random.default_rng(9)
    cell = rng.integers(0, 50, n).astype(np.int32)
    cell[rng.random(n) < 0.2] = -1
    seed, step, tg = 0x0123456789ABCDEF, 7, orc.tag(1, orc.RNG_SELECT)
    thr = int(0.3 * 2 ** 32)
    words = np.concatenate([orc.philox([g, step, tg, 0], [seed & 0xFFFFFFFF, seed >> 32])
                            for g in range((n + 3) // 4)])[:n]
    expect = np.nonzero((words < thr) & (cell >= 0))[0]
    sel = orc.mcc_select(cell, thr, seed, step, tg)
    assert np.array_equal(sel, expect) and 200 < len(sel) < 300
    # every live slot at the top threshold, none at zero
    assert np.array_equal(orc.mcc_select(cell, 2 ** 32 - 1, seed, step, tg), np.nonzero(cell >= 0)[0])
    assert len(orc.mcc_select(cell, 0, seed, step, tg)) == 0


# ------------------------------------------------------------------ library exports
def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(str(pkg.LIB_PATH))
    syms = pkg.header_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python prototypes cover the header exactly
    assert sorted(pkg._PROTOS) == syms


def test_workload_definitions(pkg):
    from patch_ref import element_and_node_classes
    w = pkg.workload("cfg3")
    d = w["domain"]
    assert d.nnx == 513 and len(d.objects) == 3 and w["ppc"] == 71
    # 16.8M particles per species: ppc per element outside the conductors (the interface
    # elements keep the particles that fall outside the objects)
    xs = np.arange(d.nnx, dtype=float)
    _, ei = element_and_node_classes(d.nnx, d.nny, xs, xs, [(o.x, o.y, o.r) for o in d.objects])
    assert (ei == -1).sum() * w["ppc"] < 16777216 < (ei != -2).sum() * w["ppc"]
    w1 = pkg.workload("cfg1")
    assert (w1["domain"].nnx - 1) ** 2 * w1["ppc"] == 409600
    # cfg5: the cfg3 layout scaled to 1024^2, about 32M particles per species (element centres
    # outside the conductors x ppc)
    w5 = pkg.workload("cfg5")
    d5 = w5["domain"]
    c = np.arange(d5.nnx - 1) + 0.5
    X, Y = np.meshgrid(c, c, indexing="ij")
    out = np.ones_like(X, dtype=bool)
    for o in d5.objects:
        out &= (X - o.x) ** 2 + (Y - o.y) ** 2 >= o.r ** 2
        assert o.r / d5.xmax == pytest.approx([q.r for q in d.objects][d5.objects.index(o)] / d.xmax)
    assert 32.0e6 < out.sum() * w5["ppc"] < 34.0e6
    s64 = pkg.workload("cfg3s64")["domain"]
    assert s64.nnx == 65 and [o.phi for o in s64.objects] == [-1.0, 0.5, 0.0]


@pytest.mark.parametrize("bits", [64, 32])
def test_push_corner_crossing_settles(orc, bits):
    """A particle crossing the left and the bottom wall in one move: the left-wall
    re-emission divides Vy by v_ref (VelRes of all three components) and the bottom
    back-up time then exceeds the move; the position is mirrored and the loop ends."""
    n = 8
    xs, ys, ni, ei, pp, _ = _mesh(n)
    nn = (n + 1) ** 2
    z = np.zeros(nn)
    x, y = np.array([0.02, 0.03, 7.99]), np.array([0.01, 7.98, 0.02])
    vx, vy, vz = np.array([-2.0, -1.5, 0.5]), np.array([-3.0, 2.5, -2.0]), np.zeros(3)
    cell, _ = orc.locate(xs, ys, x, y)
    X, Y, VX, VY, VZ, alive, cnt, err = orc.move_all(bits, xs, ys, ei, ni, np.zeros((0, 6)), z, z, np.zeros(5 * nn),
                                                     -1.0, 0.05, 419396.657327663, 9.1095e-31, 2 * 11605.0, 3, 0,
                                                     orc.tag(0, 2), x, y, vx, vy, vz, cell)
    assert err == 0
    live = alive.astype(bool)
    assert np.all((X[live] > 0) & (X[live] <= n) & (Y[live] >= 0) & (Y[live] <= n))
    assert cnt[0] >= 2 and cnt[3] >= 2


@pytest.mark.parametrize("bits", [64, 32])
def test_push_wall_fuzz_always_settles(orc, bits):
    """Random particles near the walls and corners with large velocities, repeated steps:
    the boundary loop always ends inside the domain (or the particle is absorbed)."""
    n = 8
    xs, ys, ni, ei, pp, _ = _mesh(n)
    nn = (n + 1) ** 2
    rng = np.random.default_rng(17)
    m = 20000
    x = np.where(rng.random(m) < 0.5, rng.random(m) * 0.3, n - rng.random(m) * 0.3)
    y = np.where(rng.random(m) < 0.5, rng.random(m) * 0.3, n - rng.random(m) * 0.3)
    x = np.clip(x, 1e-3, n - 1e-3)
    v = rng.normal(size=(3, m)) * rng.choice([0.1, 2.0, 40.0], size=m)
    ex, ey = rng.normal(size=(2, nn)) * 5.0
    cell, _ = orc.locate(xs, ys, x, y)
    state = [x, y, *v]
    for step in range(4):
        X, Y, VX, VY, VZ, alive, cnt, err = orc.move_all(bits, xs, ys, ei, ni, np.zeros((0, 6)), ex, ey,
                                                         np.zeros(5 * nn), -1.0, 0.05, 419396.657327663, 9.1095e-31,
                                                         2 * 11605.0, 5, step, orc.tag(0, 2), *state, cell)
        assert err == 0
        live = alive.astype(bool)
        assert np.all((X[live] > 0) & (X[live] <= n) & (Y[live] >= 0) & (Y[live] <= n))
        state = [a[live].astype(np.float64) for a in (X, Y, VX, VY, VZ)]
        cell, bad = orc.locate(xs, ys, state[0], state[1])
        assert bad == 0


# ------------------------------------------------------------------ wall injection
def test_injection_flux_constants_closed_form(orc):
    """b_amb of InjectBeams_2D.f90:42-50 is the mean inward flux velocity of a drifting
    Maxwellian, int_0^inf v N(v; vd, vt) dv (within the single-precision SQRT(2.) the
    reference uses); Bird's FS1/FS2 and Loadv's VMP follow their definitions."""
    from scipy import integrate
    for vd, vt in ((0.1, 1.0), (0.3, 2.0), (1.5, 0.7)):
        c = orc.inject_consts(vd, vt, vt * vt, 1.0, 3.0, 0.05)
        ref, _ = integrate.quad(lambda v: v * np.exp(-(v - vd) ** 2 / (2 * vt * vt)) / np.sqrt(2 * np.pi) / vt,
                                0, np.inf)
        assert c[0] == pytest.approx(ref, rel=1e-6)
        vr = vd / (np.sqrt(2) * vt)
        assert c[1] == pytest.approx(vr, rel=1e-7) and c[2] == pytest.approx(vr + np.sqrt(vr * vr + 2), rel=1e-7)
        assert c[6] == pytest.approx(np.sqrt(2) * vt, rel=1e-12)            # VMP with tmpj = vt^2, xm = 1
        assert c[7] == 3.0 and c[8] == pytest.approx(3.0 + 0.5 * ref * 0.05, rel=1e-6)
    # no drift: the reference takes the negative branch (slab behind R_start, |b| in the count)
    c = orc.inject_consts(0.0, 1.0, 1.0, 1.0, 3.0, 0.05)
    assert c[0] == pytest.approx(-1.0 / np.sqrt(2 * np.pi), rel=1e-12) and c[8] == 3.0


def test_injection_sampler_statistics(orc):
    """Bird's sampler draws the flux-weighted inward component v exp(-(v-vd)^2/2vt^2), v > 0;
    Loadv draws N(0, vt) transverse components (plus the ion drift); positions fill the slab
    and InjectPositioning matches a brute-force element search."""
    from scipy import integrate
    n, vd, vt = 60000, 0.2, 1.0
    c = orc.inject_consts(vd, vt, vt * vt, 1.0, 0.0, 0.05)
    x, y, vx, vy, vz, cell = orc.inject(64, n, 0, c, [vd, 0.25, -0.5], 0.0, 16.0, 9, 4, orc.tag(1, orc.RNG_INJECT),
                                        0.0, 0.0, 1.0, 1.0, 16, 16)
    w = lambda v: v * np.exp(-(v - vd) ** 2 / (2 * vt * vt))
    m1 = integrate.quad(lambda v: v * w(v), 0, np.inf)[0] / integrate.quad(w, 0, np.inf)[0]
    m2 = integrate.quad(lambda v: v * v * w(v), 0, np.inf)[0] / integrate.quad(w, 0, np.inf)[0]
    assert np.all(vx > 0)
    assert abs(vx.mean() - m1) < 5 * np.sqrt((m2 - m1 * m1) / n)
    for v, drift in ((vy, 0.25), (vz, -0.5)):
        assert abs(v.mean() - drift) < 5 * vt / np.sqrt(n)
        assert abs(v.std() - vt) < 0.02
    assert np.all((x >= c[7]) & (x <= c[8]) & (y >= 0) & (y <= 16))
    brute = np.floor(y).astype(int) + np.floor(x).astype(int) * 16
    assert np.array_equal(cell, brute)
    # electrons get no transverse drift
    _, _, _, vy_e, _, _ = orc.inject(64, 4000, 1, c, [vd, 0.25, -0.5], 0.0, 16.0, 9, 4, orc.tag(0, orc.RNG_INJECT),
                                     0.0, 0.0, 1.0, 1.0, 16, 16)
    assert abs(vy_e.mean()) < 5 * vt / np.sqrt(4000)


def test_injection_upper_edge_stays_inside(orc):
    """A slab 1e-3 high under the mesh's upper edge y = 16: about one FP32 Y in 2000 rounds onto
    16.0; those are stored as the float below 16, so every particle keeps an element (the top
    row) and none sits on the edge. FP64 Y never reaches the edge here."""
    c = orc.inject_consts(0.2, 1.0, 1.0, 1.0, 0.0, 0.05)
    below = float(np.nextafter(np.float32(16.0), np.float32(0.0)))
    for bits in (32, 64):
        x, y, vx, vy, vz, cell = orc.inject(bits, 40000, 0, c, [0.2, 0.0, 0.0], 16.0 - 1e-3, 16.0, 9, 4,
                                            orc.tag(1, orc.RNG_INJECT), 0.0, 0.0, 1.0, 1.0, 16, 16)
        assert np.all(y < 16.0) and np.all(cell >= 0) and np.all(cell % 16 == 15)
        hits = int(np.count_nonzero(y == below))
        if bits == 32:
            assert hits > 5                   # the clamp was exercised
            assert np.all(y == y.astype(np.float32))


def test_injection_count_rounding(orc):
    """add_N = INT(expect + R): floor or floor + 1, with mean expect over many steps."""
    counts = np.array([orc.inject_count(12.37, 3, s, orc.tag(0, orc.RNG_INJECT)) for s in range(4000)])
    assert set(np.unique(counts)) <= {12, 13}
    assert abs(counts.mean() - 12.37) < 5 * 0.5 / np.sqrt(4000)


# ------------------------------------------------------------------ threaded oracle
def _csr(A):
    rows, cols = np.nonzero(np.abs(A) > 0)
    rowptr = np.zeros(A.shape[0] + 1, dtype=np.int32)
    np.add.at(rowptr, rows + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), cols.astype(np.int32), A[rows, cols], np.diag(A).copy()


def test_threaded_oracle_matches_serial(orc):
    """The OpenMP split of the oracle's loops (used only for the multi-core CPU timing)
    reproduces the serial reference order: push, lookup, fixed-point deposition and PPR are
    bit-identical; PCG and the RHS, whose sums are split over threads, agree to rounding."""
    n = 16
    xs, ys, ni, ei, pp, _ = _mesh(n, circles=[(8.0, 8.0, 2.5)])
    nn = (n + 1) ** 2
    rng = np.random.default_rng(23)
    m = 40000
    x = 0.05 + (n - 0.1) * rng.random(m)
    y = (n - 1e-3) * rng.random(m)
    v = rng.normal(size=(3, m)) * rng.choice([0.1, 2.0, 20.0], size=m)
    ex, ey = rng.normal(size=(2, nn))
    coef5 = rng.normal(size=5 * nn)
    obj = np.array([[1, 8.0, 8.0, 0, 0, 6.25]])
    cell, _ = orc.locate(xs, ys, x, y)
    phi = rng.normal(size=nn)
    _, _, nt2, A = _poisson_system(orc, n)
    rp, col, val, diag = _csr(A)
    b = rng.normal(size=A.shape[0])
    rho = rng.normal(size=nn)
    dphi = rng.normal(size=nn)

    def run():
        mv = orc.move_all(64, xs, ys, ei, ni, obj, ex, ey, coef5, -1.0, 0.05, 419396.657327663, 9.1095e-31,
                          2 * 11605.0, 7, 3, orc.tag(0, 2), x, y, *v, cell)
        keep = mv[5].astype(bool)
        loc, bad = orc.locate(xs, ys, mv[0][keep], mv[1][keep])
        dep = orc.deposit(64, xs, ys, 0, *(a[keep] for a in mv[:5]), loc, 9.1095e-31, 4.19e5, True)
        pr = orc.ppr(xs, ys, pp, phi)
        pc = orc.pcg(rp, col, val, diag, b, np.zeros_like(b), 1e-10, 5000)
        rh = orc.rhs(xs, ys, 0, nt2, dphi, rho, A.shape[0])
        sel = orc.mcc_select(np.where(rng_cells >= 0, rng_cells, -1), 900000000, 7, 3, orc.tag(1, 3))
        return mv, loc, bad, dep, pr, pc, rh, sel

    rng_cells = np.random.default_rng(4).integers(-1, 9, size=m).astype(np.int32)

    ref = run()
    try:
        orc.set_threads(4)
        assert orc.get_threads() == 4
        got = run()
    finally:
        orc.set_threads(1)
    for a, b_ in zip(ref[0][:6], got[0][:6]):
        assert np.array_equal(a, b_)
    assert np.array_equal(ref[0][6], got[0][6]) and ref[0][7] == got[0][7]
    assert np.array_equal(ref[1], got[1]) and ref[2] == got[2]
    for a, b_ in zip(ref[3], got[3]):
        assert np.array_equal(a, b_)
    for a, b_ in zip(ref[4][:3], got[4][:3]):
        assert np.array_equal(a, b_)
    np.testing.assert_allclose(got[5][0], ref[5][0], rtol=1e-9, atol=1e-12)
    assert abs(got[5][2] - ref[5][2]) <= 1
    np.testing.assert_allclose(got[6], ref[6], rtol=1e-13, atol=1e-14)
    assert len(ref[7]) > 1000 and np.array_equal(ref[7], got[7])     # MCC selection, slot order kept


# ------------------------------------------------------------------ axisymmetric mover
def test_trig_against_libm(orc):
    """The mover's own sin / cos / atan (one IEEE sequence shared by the device and the
    oracle) stay within 1 ulp (sin, cos) and 2 ulp (atan) of the C library, with the exact
    special values and symmetries."""
    import math
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-8.0, 8.0, 4000), rng.uniform(-1e-6, 1e-6, 200), np.arange(-12, 13) * np.pi / 4])
    for x in xs:
        s, c = orc.sincos(x)
        for got, ref in ((s, math.sin(x)), (c, math.cos(x))):
            assert abs(got - ref) <= math.ulp(ref), (x, got, ref)
    ts = np.concatenate([rng.uniform(-3.0, 3.0, 4000), 10.0 ** rng.uniform(-12, 12, 4000), [0.41421356237309503]])
    for t in ts:
        for tt in (t, -t):
            assert abs(orc.atan(tt) - math.atan(tt)) <= 2 * math.ulp(math.atan(tt)), tt
        assert orc.atan(-t) == -orc.atan(t)
    assert orc.sincos(0.0) == (0.0, 1.0) and orc.atan(0.0) == 0.0
    assert orc.atan(1.0) == math.pi / 4 and orc.atan(math.inf) == math.pi / 2
    assert math.isnan(orc.atan(math.nan))


def _rz_case(orc, n_part=400, seed=5):
    n = 16
    xs = np.arange(n + 1, dtype=np.float64)            # axial 0..16
    ys = np.arange(n + 1, dtype=np.float64)            # radial 0..16
    from patch_ref import element_and_node_classes
    ni, ei = element_and_node_classes(n + 1, n + 1, xs, ys, ())
    rng = np.random.default_rng(seed)
    x = 4.0 + 8.0 * rng.random(n_part)
    y = 3.0 + 8.0 * rng.random(n_part)
    th = rng.uniform(-1.5, 4.6, n_part)
    v = rng.normal(size=(3, n_part)) * 0.6
    cell, bad = orc.locate(xs, ys, x, y)
    assert bad == 0
    return xs, ys, ni, ei, x, y, th, v, cell


def test_axisymmetric_push_field_free_is_a_straight_line(orc):
    """With no field the r-z mover moves the particle on a straight line of 3-D space:
    z += T vx, (r cos th, r sin th) += T (vy, vz), then back to (r, th) with th = atan(Y/X)
    (+ PI_LY when X <= 0), MoveOne.f90:246-263."""
    xs, ys, ni, ei, x, y, th, v, cell = _rz_case(orc)
    nn = len(xs) * len(ys)
    z = np.zeros(nn)
    dt = 0.25
    X, Y, VX, VY, VZ, alive, cnt, err, TH = orc.move_all(64, xs, ys, ei, ni, np.zeros((0, 6)), z, z, np.zeros(5 * nn),
                                                         -1.0, dt, 4e5, 9.1e-31, 2 * 11605.0, 1, 0, 0, x, y, *v,
                                                         cell, delta=1, th=th)
    assert err == 0 and alive.all()
    np.testing.assert_array_equal(VX, v[0]); np.testing.assert_array_equal(VY, v[1]); np.testing.assert_array_equal(VZ, v[2])
    cx = y * np.cos(th) + dt * v[1]
    cy = y * np.sin(th) + dt * v[2]
    np.testing.assert_allclose(X, x + dt * v[0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(Y, np.hypot(cx, cy), rtol=1e-14)
    expect_th = np.arctan(cy / cx) + np.where(cx <= 0, 3.14159265358979, 0.0)
    np.testing.assert_allclose(TH, expect_th, rtol=0, atol=1e-13)


def test_axisymmetric_push_uniform_fields_and_angular_momentum(orc):
    """Uniform E_z (efx) and E_r (efy): vx += qm E_z T, (vy, vz) += qm E_r T (cos th, sin th)
    (MoveOne.f90:221-236); a radial kick leaves the angular momentum about the axis
    x vz - y vy (Cartesian frame, before the move) unchanged."""
    xs, ys, ni, ei, x, y, th, v, cell = _rz_case(orc, seed=8)
    nn = len(xs) * len(ys)
    Ez, Er, qm, dt = 0.3, -0.7, -1.0, 0.05
    out = orc.move_all(64, xs, ys, ei, ni, np.zeros((0, 6)), np.full(nn, Ez), np.full(nn, Er), np.zeros(5 * nn), qm,
                       dt, 4e5, 9.1e-31, 2 * 11605.0, 1, 0, 0, x, y, *v, cell, delta=1, th=th)
    X, Y, VX, VY, VZ, alive, cnt, err, TH = out
    assert err == 0 and alive.all()
    np.testing.assert_allclose(VX, v[0] + qm * Ez * dt, rtol=0, atol=1e-15)
    np.testing.assert_allclose(VY, v[1] + qm * Er * dt * np.cos(th), rtol=0, atol=1e-15)
    np.testing.assert_allclose(VZ, v[2] + qm * Er * dt * np.sin(th), rtol=0, atol=1e-15)
    px, py = y * np.cos(th), y * np.sin(th)
    np.testing.assert_allclose(px * VZ - py * VY, px * v[2] - py * v[1], rtol=1e-13, atol=1e-13)
    # (r, th) after the move is the kicked straight line in the Cartesian frame
    cx, cy = px + dt * VY, py + dt * VZ
    np.testing.assert_allclose(Y * np.cos(TH), cx, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Y * np.sin(TH), cy, rtol=0, atol=1e-12)


def test_axisymmetric_outer_radius_absorbs(orc):
    """delta = 1: a particle carried past r = ymax is deleted (the build's outer-radius wall;
    the reference has no branch there), not reflected, and counts only as a deletion."""
    xs, ys, ni, ei, x, y, th, v, cell = _rz_case(orc, n_part=3)
    nn = len(xs) * len(ys)
    z = np.zeros(nn)
    x = np.array([8.0, 8.0, 8.0]); y = np.array([15.9, 15.9, 5.0]); th = np.array([0.0, 1.0, 2.0])
    vx = np.zeros(3); vy = np.array([4.0, 4.0 * np.cos(1.0), 0.1]); vz = np.array([0.0, 4.0 * np.sin(1.0), 0.0])
    cell, _ = orc.locate(xs, ys, x, y)
    out = orc.move_all(64, xs, ys, ei, ni, np.zeros((0, 6)), z, z, np.zeros(5 * nn), -1.0, 0.05, 4e5, 9.1e-31,
                       2 * 11605.0, 1, 0, 0, x, y, vx, vy, vz, cell, delta=1, th=th)
    X, Y, VX, VY, VZ, alive, cnt, err, TH = out
    assert err == 0 and alive.tolist() == [0, 0, 1]
    assert cnt[7] == 2 and cnt[1] == 0 and cnt[3] == 0


def test_sort_group_needs_a_domain(pkg):
    """pic_sort_group before pic_init / pic_setup_domain is a call-order error (host-side
    check only; run in a fresh process so no earlier test's library state is visible)."""
    import subprocess
    import sys
    code = ("import ctypes; lib = ctypes.CDLL(%r); n = ctypes.c_int32(-1); "
            "rc = lib.pic_sort_group(ctypes.byref(n)); print(rc, n.value)" % str(pkg.LIB_PATH))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rc, n = map(int, out.stdout.split())
    assert rc == 2 and n == -1        # PIC_ERR_STATE, output untouched
