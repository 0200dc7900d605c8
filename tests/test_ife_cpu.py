"""The oracle's immersed-finite-element (IFE) interface element (Linear_IFE_Basis_Coeff_2D,
Gauss_integration_local_stiffness_IFE, IFE_RHS_2D), pinned by what the mathematics fixes:
the defining conditions of the bilinear IFE basis, its reduction to the Q1 basis when both
sides have the same permittivity, exact reproduction of piecewise-linear functions that satisfy
the interface conditions, orientation independence, and second-order convergence on the
standard circular-interface problem with a closed-form solution."""
import numpy as np
import pytest

VX = np.array([0.0, 1.0, 1.0, 0.0])
VY = np.array([0.0, 0.0, 1.0, 1.0])


def _cases():
    """(inside flags, circle centre, radius) in element units [0,1]^2: all four lone corners
    (type 1, inside and outside corner) and the four opposite-edge cuts (type 2)."""
    out = []
    for k in range(4):
        c = np.array([VX[k], VY[k]]) + 0.35 * (np.array([VX[k], VY[k]]) - 0.5)
        out.append(c.tolist() + [0.62])                     # lone vertex inside
        c2 = np.array([VX[k], VY[k]]) + 1.4 * (np.array([VX[k], VY[k]]) - 0.5)
        out.append(c2.tolist() + [1.85])                    # lone vertex outside
    for cx, cy in ((0.5, -1.6), (0.5, 2.7), (-1.9, 0.5), (2.4, 0.4)):
        out.append([cx, cy, 2.0])                           # a nearly straight cut through two opposite edges
    return out


def _inside(cx, cy, r):
    return [int((VX[k] - cx) ** 2 + (VY[k] - cy) ** 2 < r * r) for k in range(4)]


def _phi(coef, i, pc, X, Y):
    c = coef[i, pc]
    return c[0] + c[1] * X + c[2] * Y + c[3] * X * Y


def _grad(coef, i, pc, X, Y):
    c = coef[i, pc]
    return np.array([c[1] + c[3] * Y, c[2] + c[3] * X])


def _piece_of_vertex(t, order):
    p = np.ones(4, dtype=int)
    p[order[0]] = 0
    if t == 2:
        p[order[3]] = 0
    return p


@pytest.mark.parametrize("beta_in,beta_out", [(4.0, 1.0), (1.0, 10.0), (1e4, 1.0)])
def test_ife_basis_defining_conditions(orc, beta_in, beta_out):
    types = set()
    for cx, cy, r in _cases():
        ins = _inside(cx, cy, r)
        t, K, M, coef, pts, order = orc.ife_element(0, 0.0, 0.0, 1.0, ins, cx, cy, r, beta_in, beta_out)
        assert t in (1, 2), (cx, cy, r, ins)
        types.add(t)
        D, E = pts
        # D and E lie on the circle and on the element boundary
        for P in (D, E):
            assert abs(np.hypot(P[0] - cx, P[1] - cy) - r) < 1e-13
            assert min(P[0], 1 - P[0], P[1], 1 - P[1]) < 1e-15
        piece = _piece_of_vertex(t, order)
        beta = [beta_in if ins[order[0]] else beta_out, beta_out if ins[order[0]] else beta_in]
        # each piece is on one side of the circle
        for k in range(4):
            assert (beta[piece[k]] == beta_in) == bool(ins[k])
        n = np.array([E[1] - D[1], -(E[0] - D[0])])
        gl = np.polynomial.legendre.leggauss(4)
        for i in range(4):
            # nodal: phi_i(A_k) = delta_ik
            for k in range(4):
                assert abs(_phi(coef, i, piece[k], VX[k], VY[k]) - (i == k)) < 1e-12
            # continuous along the whole segment DE
            for s in np.linspace(0, 1, 7):
                P = D + s * (E - D)
                assert abs(_phi(coef, i, 0, *P) - _phi(coef, i, 1, *P)) < 1e-12
            # beta * normal flux has no jump across DE (integrated along it)
            jump = 0.0
            for x, w in zip(*gl):
                P = D + 0.5 * (x + 1) * (E - D)
                jump += 0.5 * w * (beta[0] * _grad(coef, i, 0, *P) - beta[1] * _grad(coef, i, 1, *P)) @ n
            assert abs(jump) < 1e-11 * max(beta)
        # partition of unity on both pieces
        for pc in (0, 1):
            s = coef[:, pc].sum(axis=0)
            np.testing.assert_allclose(s, [1, 0, 0, 0], atol=1e-11)
        # K symmetric, constants in its kernel; M sums to the element area
        np.testing.assert_allclose(K, K.T, atol=1e-12 * abs(K).max())
        np.testing.assert_allclose(K.sum(axis=1), 0, atol=1e-11 * abs(K).max())
        assert M.sum() == pytest.approx(1.0, abs=1e-15 * max(beta) / min(beta) + 1e-14)
    assert types == {1, 2}


def test_ife_reduces_to_q1_without_a_jump(orc):
    """Equal permittivities: the basis is the bilinear Q1 basis, so K is the Q1 stiffness (and,
    axisymmetric, the r-weighted one); M is the edge-midpoint rule of the Q1 mass on the
    reference's sub-triangles, exact for the constant and linear moments."""
    for delta in (0, 1):
        for cx, cy, r in _cases():
            ins = _inside(cx, cy, r)
            y0 = 3.0 if delta else 0.0
            t, K, M, coef, pts, order = orc.ife_element(delta, 2.0, y0, 1.0, ins, cx + 2.0, cy + y0, r, 2.5, 2.5)
            Kq, Mq = orc.element_matrices(delta, 2.0, y0, 1.0, 1.0)
            np.testing.assert_allclose(K, 2.5 * Kq, rtol=0, atol=1e-12 * abs(Kq).max())
            q1 = np.array([[1, -1, -1, 1], [0, 1, 0, -1], [0, 0, 0, 1], [0, 0, 1, -1]], dtype=float)
            for pc in (0, 1):
                np.testing.assert_allclose(coef[:, pc], q1, atol=1e-12)
            # the midpoint rule is exact to degree 2: int N_i (Cartesian) and int r (r-z)
            if delta == 0:
                np.testing.assert_allclose(M.sum(axis=1), Mq.sum(axis=1), rtol=1e-13)
            else:
                assert M.sum() == pytest.approx(Mq.sum(), rel=1e-14)


def test_ife_reproduces_piecewise_linear_solutions(orc):
    """u = (n . (x - D)) / beta on each side is continuous with continuous beta * flux across the
    line DE; the IFE interpolant sum_k u(A_k) phi_k reproduces it on both pieces."""
    for beta_in, beta_out in ((4.0, 1.0), (1.0, 7.0)):
        for cx, cy, r in _cases():
            ins = _inside(cx, cy, r)
            t, K, M, coef, pts, order = orc.ife_element(0, 0.0, 0.0, 1.0, ins, cx, cy, r, beta_in, beta_out)
            D, E = pts
            n = np.array([E[1] - D[1], -(E[0] - D[0])])
            piece = _piece_of_vertex(t, order)
            beta = [beta_in if ins[order[0]] else beta_out, beta_out if ins[order[0]] else beta_in]
            u = lambda X, Y, pc: (n @ (np.array([X, Y]) - D)) / beta[pc]
            nodal = [u(VX[k], VY[k], piece[k]) for k in range(4)]
            rng = np.random.default_rng(3)
            for _ in range(20):
                X, Y = rng.random(2)
                for pc in (0, 1):
                    interp = sum(nodal[k] * _phi(coef, k, pc, X, Y) for k in range(4))
                    assert abs(interp - u(X, Y, pc)) < 1e-12


def test_ife_orientation_independent(orc):
    """Rotating the configuration by a quarter turn about the element centre permutes K and M
    the same way (the classification picks another A1, the space is the same)."""
    for cx, cy, r in _cases():
        ins = _inside(cx, cy, r)
        t, K, M, *_ = orc.ife_element(0, 0.0, 0.0, 1.0, ins, cx, cy, r, 4.0, 1.0)
        # rotate by +90 degrees: (x, y) -> (1 - y, x); vertex k moves to vertex k + 1
        rx, ry = 1 - cy, cx
        ins2 = _inside(rx, ry, r)
        assert ins2 == [ins[3], ins[0], ins[1], ins[2]]
        t2, K2, M2, *_ = orc.ife_element(0, 0.0, 0.0, 1.0, ins2, rx, ry, r, 4.0, 1.0)
        p = [3, 0, 1, 2]     # new vertex k is old vertex p[k]
        np.testing.assert_allclose(K2, K[np.ix_(p, p)], atol=1e-12)
        if t == 1:           # the type-1 sub-triangles are symmetric about A1; type 2 picks E by edge
            np.testing.assert_allclose(M2, M[np.ix_(p, p)], atol=1e-14)


def _circle_problem(orc, n, beta_in, beta_out=1.0, r0=0.5031, mistake=None):
    """-div(beta grad u) = f on [-1,1]^2 with a circular interface of radius r0 (Li, Lin and Wu
    2003): u = r^3/beta_in inside, r^3/beta_out + (1/beta_in - 1/beta_out) r0^3 outside,
    f = -9 r, Dirichlet data on the box from u. Returns the max and RMS nodal errors."""
    import scipy.sparse.linalg as spl
    xs = np.linspace(-1.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    R = np.hypot(X, Y).ravel()
    exact = np.where(R < r0, R ** 3 / beta_in, R ** 3 / beta_out + (1 / beta_in - 1 / beta_out) * r0 ** 3)
    bnd = ((X == -1) | (X == 1) | (Y == -1) | (Y == 1)).ravel()
    nt2 = np.zeros((n + 1) ** 2, dtype=np.int32)
    nt2[~bnd] = np.arange(1, (~bnd).sum() + 1)
    nt2[bnd] = -1
    A, c, B, types = orc.assemble_ife(xs, xs, 0, [(0.0, 0.0, r0, beta_in, False)], nt2, np.where(bnd, exact, 0.0),
                                      mistake=mistake)
    err = spl.spsolve(A.tocsc(), B @ (-9.0 * R) - c) - exact[~bnd]
    return np.abs(err).max(), np.sqrt(np.mean(err ** 2)), types


@pytest.mark.parametrize("beta_in", [4.0, 0.25, 1e4])
def test_ife_converges_on_circular_interface(orc, beta_in):
    """Galerkin bilinear IFE without penalties (the reference's stiffness: its interior-penalty
    edge terms act only between refinement levels): the nodal RMS error falls at close to second
    order on every halving of h, the maximum error (at nodes next to the interface) about as h."""
    mx, rms = [], []
    for n in (16, 32, 64, 128):
        e, r, types = _circle_problem(orc, n, beta_in)
        assert set(types) == {1, 2}
        mx.append(e)
        rms.append(r)
    assert all(rms[i + 1] < rms[i] / 1.8 for i in range(3)), rms
    assert rms[3] < rms[0] / 20 and mx[3] < mx[0] / 4, (rms, mx)


def test_ife_planted_errors_are_caught(orc):
    """The circular-interface problem separates the IFE element from plausible mistakes: plain
    Q1 on the cut elements, or the two permittivities swapped there."""
    _, r_ife, _ = _circle_problem(orc, 64, 4.0)
    for mistake in ("q1_cut", "swap"):
        _, r_bad, _ = _circle_problem(orc, 64, 4.0, mistake=mistake)
        assert r_bad > 4 * r_ife, (mistake, r_bad, r_ife)
