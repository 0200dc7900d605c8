"""CPU checks of the learned-multigrid oracle (oracle/learned.py): adjointness of the
transposed convolution, the Galerkin hierarchy, symmetry and definiteness of the V-cycle,
the untrained network equal to the classical Jacobi V-cycle, preconditioning quality, and
the analytic gradient against central differences of the FP64 loss."""
import numpy as np
import pytest
import scipy.sparse as sp

import learned as lrn


def q1_operator(n, holes=True):
    """Jacobi-scaled Q1 stiffness (unit centre, -1/8 neighbours) on an n x n node grid,
    Dirichlet walls and (optionally) three Dirichlet discs; returns A' and the unknown mask."""
    N = n * n
    xs = np.arange(n) / (n - 1)
    circ = [(0.25, 0.5, 0.08), (0.5, 0.5, 0.08), (0.75, 0.5, 0.08)] if holes else []
    unk = np.zeros((n, n), bool)
    for i in range(1, n - 1):
        for j in range(1, n - 1):
            unk[i, j] = not any((xs[i] - cx) ** 2 + (xs[j] - cy) ** 2 <= r * r for cx, cy, r in circ)
    rows, cols, vals = [], [], []
    for i in range(1, n - 1):
        for j in range(1, n - 1):
            if not unk[i, j]:
                continue
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    if unk[i + di, j + dj]:
                        rows.append(i * n + j); cols.append((i + di) * n + j + dj)
                        vals.append(1.0 if (di == 0 and dj == 0) else -1.0 / 8.0)
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N)), unk.reshape(-1).astype(float)


def pcg_iterations(A, b, Mf, tol=1e-8, maxit=2000):
    x = np.zeros_like(b); r = b.copy(); z = Mf(r); p = z.copy(); rz = r @ z; nb = np.linalg.norm(b)
    for k in range(maxit):
        q = A @ p; a = rz / (p @ q); x += a * p; r -= a * q
        if np.linalg.norm(r) / nb < tol:
            return k + 1
        z = Mf(r); rz2 = r @ z; p = z + rz2 / rz * p; rz = rz2
    return maxit


def test_transposed_convolution_is_adjoint():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((9, 7, 16))
    z = rng.standard_normal((9, 7, 32))
    w = rng.standard_normal((9, 32, 16))
    lhs = np.sum(lrn.conv_same(x, w) * z)
    rhs = np.sum(x * lrn.conv_same(z, np.transpose(w, (0, 2, 1)), sign=-1))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_level_sizes_and_galerkin_hierarchy():
    # coarsening stops at the first level of <= 33 x 33 nodes ...
    assert lrn.level_sizes(513, 513)[0] == [513, 257, 129, 65, 33]
    assert lrn.level_sizes(1025, 1025)[0] == [1025, 513, 257, 129, 65, 33]
    assert lrn.level_sizes(513, 129) == ([513, 257, 129, 65, 33], [129, 65, 33, 17, 9])
    # ... once every learned level has a coarser one, and never below about 3 x 3
    assert lrn.level_sizes(17, 17, learned_levels=2)[0] == [17, 9, 5]
    assert lrn.level_sizes(65, 65, learned_levels=4)[0] == [65, 33, 17, 9, 5]
    assert lrn.level_sizes(9, 9, learned_levels=4)[0] == [9, 5, 3]
    assert lrn.level_sizes(513, 513, dense_nodes=0)[0] == [513, 257, 129, 65, 33, 17, 9, 5, 3]
    A, mask = q1_operator(17)
    lev = lrn.build_hierarchy(A, 17, 17, dense_nodes=0)
    assert [l["H"] for l in lev] == [17, 9, 5, 3]
    for l in range(len(lev) - 1):
        Ac = lev[l + 1]["A"].toarray()
        assert np.allclose(Ac, Ac.T)                      # Galerkin operators stay symmetric
        P = lev[l]["P"]
        # bilinear interpolation reproduces constants away from Dirichlet nodes
        ones_c = np.ones(P.shape[1])
        f = P @ ones_c
        inner = (lev[l]["mask"] > 0)
        assert np.all(f[inner] <= 1.0 + 1e-12)
    # exact coarsest solve
    last = lev[-1]
    idx = np.nonzero(last["mask"])[0]
    Ad = last["A"].toarray()[np.ix_(idx, idx)]
    assert np.allclose(last["inv"][np.ix_(idx, idx)] @ Ad, np.eye(len(idx)))


def test_untrained_network_is_the_jacobi_vcycle_and_preconditions():
    n = 65
    A, mask = q1_operator(n)
    lev = lrn.build_hierarchy(A, n, n)
    lay = lrn.layout(n, n, [16], learned_levels=2)
    p = lrn.initial_params(lay, small=0.0)
    lay0 = lrn.layout(n, n, [16], learned_levels=0)
    rng = np.random.default_rng(4)
    r = rng.standard_normal(n * n) * mask
    y1, _ = lrn.forward(p, lay, lev, r)
    y0, _ = lrn.forward(np.zeros(0), lay0, lev, r)
    assert np.allclose(y1, y0, rtol=1e-12, atol=1e-14)
    it_mg = pcg_iterations(A, r, lambda v: lrn.forward(p, lay, lev, v)[0])
    it_j = pcg_iterations(A, r, lambda v: v)
    assert it_mg <= 12 and it_j >= 8 * it_mg, (it_mg, it_j)


def test_vcycle_is_symmetric_positive_definite():
    rng = np.random.default_rng(1)
    n = 33
    A, mask = q1_operator(n)
    lev = lrn.build_hierarchy(A, n, n)
    lay = lrn.layout(n, n, [16, 32], learned_levels=2)
    p = lrn.initial_params(lay, small=0.05)
    u = rng.standard_normal(n * n) * mask
    v = rng.standard_normal(n * n) * mask
    Mu = lrn.forward(p, lay, lev, u)[0]
    Mv = lrn.forward(p, lay, lev, v)[0]
    assert abs(u @ Mv - v @ Mu) <= 1e-10 * abs(u @ Mv)
    assert v @ Mv > 0


def _fd(p, lay, lev, pairs, d, eps=1e-6):
    return (lrn.loss_only(p + eps * d, lay, lev, pairs) - lrn.loss_only(p - eps * d, lay, lev, pairs)) / (2 * eps)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(2)
    n = 17
    A, mask = q1_operator(n)
    lev = lrn.build_hierarchy(A, n, n)
    lay = lrn.layout(n, n, [16, 32], learned_levels=2)
    p = lrn.initial_params(lay, small=0.2)
    pairs = []
    for _ in range(2):
        e = rng.standard_normal(n * n) * mask
        pairs.append((A @ e, e))
    loss, grad = lrn.loss_and_grad(p, lay, lev, pairs)
    assert np.isclose(loss, lrn.loss_only(p, lay, lev, pairs), rtol=1e-12)
    coords = []
    for l in range(lay["learned"]):
        C = lay["C"][l]
        coords += [lay["off_in"][l] + 4 * C + 0, lay["off_in"][l] + 2 * C + 3,
                   lay["off_q"][l], lay["off_q"][l] + C + 2]
    for c in coords:
        d = np.zeros_like(p); d[c] = 1.0
        fd = _fd(p, lay, lev, pairs, d)
        assert abs(fd - grad[c]) <= 1e-5 * max(1e-3, abs(fd)), (c, fd, grad[c])
    for _ in range(3):
        d = rng.standard_normal(p.size)
        fd = _fd(p, lay, lev, pairs, d)
        assert abs(fd - grad @ d) <= 1e-6 * max(1.0, abs(fd))


def test_planted_mistake_in_gradient_is_caught():
    """Dropping one of the two site pairings (the post-smoother's) breaks the
    central-difference agreement: the check can see that mistake."""
    rng = np.random.default_rng(3)
    n = 17
    A, mask = q1_operator(n)
    lev = lrn.build_hierarchy(A, n, n)
    lay = lrn.layout(n, n, [16], learned_levels=1)
    p = lrn.initial_params(lay, small=0.2)
    e = rng.standard_normal(n * n) * mask
    pairs = [(A @ e, e)]
    _, grad = lrn.loss_and_grad(p, lay, lev, pairs)
    net = lrn._Net(p, lay, lev)
    sr, sg = {}, {}
    y = net.vcycle(0, pairs[0][0], sr)
    ea = e @ (A @ e)
    g = 2 * mask * (A @ (y - e)) / ea
    net.vcycle(0, g, sg)
    u, v = sg[0][1], sr[0][0]                   # only the first pairing
    uu = lrn._as_chan(lev[0]["dh"] * u, n, n); vv = lrn._as_chan(lev[0]["dh"] * v, n, n)
    U = lrn.conv_same(uu, net.w_in[0]); V = lrn.conv_same(vv, net.w_in[0])
    dS = np.einsum("pqo,pqi->oi", U, V)
    wrong = ((dS + dS.T) @ net.q[0]).reshape(-1)
    c = lay["off_q"][0]
    d = np.zeros_like(p); d[c] = 1.0
    fd = _fd(p, lay, lev, pairs, d)
    assert abs(fd - grad[c]) <= 1e-5 * max(1e-3, abs(fd))
    assert abs(fd - wrong[0]) > 1e-2 * abs(fd)


def _tap_gram(u, v, H, W):
    """G[t][t2] = sum_p u(p + d_t) v(p + d_t2) over the grid, zero outside."""
    up = np.zeros((H + 2, W + 2)); up[1:-1, 1:-1] = u.reshape(H, W)
    vp = np.zeros((H + 2, W + 2)); vp[1:-1, 1:-1] = v.reshape(H, W)
    G = np.zeros((9, 9))
    for t, (di, dj) in enumerate(lrn.TAPS):
        for t2, (ei, ej) in enumerate(lrn.TAPS):
            G[t, t2] = np.sum(up[1 + di:1 + di + H, 1 + dj:1 + dj + W] * vp[1 + ei:1 + ei + H, 1 + ej:1 + ej + W])
    return G


def test_gradient_contracts_to_tap_pair_gram_matrix():
    """The library's gradient of a learned level, from the tap-pair Gram matrix G summed over
    both pairings (Gs = G + G^T): dIn = Gs W S and dQ = (W^T Gs W) Q, equals the oracle's
    factored gradient (lifts and correlations over pixels) on a rectangular grid; with G in
    place of Gs (a dropped transpose) it does not."""
    rng = np.random.default_rng(7)
    n = 17
    A, mask = q1_operator(n)
    lev = lrn.build_hierarchy(A, n, n)
    lay = lrn.layout(n, n, [16], learned_levels=1)
    p = lrn.initial_params(lay, small=0.3)
    e = rng.standard_normal(n * n) * mask
    pairs = [(A @ e, e)]
    _, grad = lrn.loss_and_grad(p, lay, lev, pairs)
    net = lrn._Net(p, lay, lev)
    sr, sg = {}, {}
    y = net.vcycle(0, pairs[0][0], sr)
    g = 2 * mask * (A @ (y - e)) / (e @ (A @ e))
    net.vcycle(0, g, sg)
    dh = lev[0]["dh"]
    G = _tap_gram(dh * sg[0][1], dh * sr[0][0], n, n) + _tap_gram(dh * sg[0][0], dh * sr[0][1], n, n)
    w_in, q = lrn.unpack(p, lay)
    Wm, Q = w_in[0][:, :, 0], q[0]
    S = Q @ Q.T
    C = lay["C"][0]
    g_in = grad[lay["off_in"][0]: lay["off_in"][0] + 9 * C].reshape(9, C)
    g_q = grad[lay["off_q"][0]: lay["off_q"][0] + C * C].reshape(C, C)
    Gs = G + G.T
    assert np.allclose(Gs @ Wm @ S, g_in, rtol=1e-10, atol=1e-12 * np.abs(g_in).max())
    assert np.allclose(Wm.T @ Gs @ Wm @ Q, g_q, rtol=1e-10, atol=1e-12 * np.abs(g_q).max())
    assert np.abs(2 * G @ Wm @ S - g_in).max() > 1e-3 * np.abs(g_in).max()


def test_collapsed_smoother_equals_factored_smoother():
    """In^T S In applied through the channels equals the 81 tap-pair weights the library
    applies, up to the grid edge (and the pair matrix is symmetric)."""
    rng = np.random.default_rng(5)
    n = 11
    lay = lrn.layout(n, n, [16], learned_levels=1)
    p = lrn.initial_params(lay, small=0.3)
    w_in, q = lrn.unpack(p, lay)
    k81 = lrn.collapsed_pairs(w_in[0], q[0])
    assert np.allclose(k81, k81.T)
    v = rng.standard_normal((n, n))
    E = lrn.conv_same(lrn._as_chan(v.reshape(-1), n, n), np.einsum("oc,tci->toi", q[0] @ q[0].T, w_in[0]))
    y_fact = lrn.conv_same(E, np.transpose(w_in[0], (0, 2, 1)), sign=-1)[:, :, 0]
    y_col = lrn.apply_collapsed(k81, v)
    assert np.allclose(y_fact, y_col, rtol=1e-12, atol=1e-12)


def test_stencil_equals_tap_pairs_on_interior_nodes():
    """The 5 x 5 stencil the library applies away from the grid edge equals the 81 tap-pair
    form there (not on the outer ring, where the pair form drops off-grid intermediates)."""
    rng = np.random.default_rng(6)
    n = 12
    lay = lrn.layout(n, n, [16], learned_levels=1)
    p = lrn.initial_params(lay, small=0.3)
    w_in, q = lrn.unpack(p, lay)
    k81 = lrn.collapsed_pairs(w_in[0], q[0])
    k25 = lrn.collapsed_stencil(k81)
    assert np.isclose(k25.sum(), k81.sum())
    v = rng.standard_normal((n, n))
    y81 = lrn.apply_collapsed(k81, v)
    y25 = lrn.apply_stencil(k25, v)
    assert np.allclose(y81[1:-1, 1:-1], y25[1:-1, 1:-1], rtol=1e-12, atol=1e-12)
    assert not np.allclose(y81[0, :], y25[0, :])
