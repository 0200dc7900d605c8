"""CPU checks of the learned-preconditioner oracle (oracle/learned.py): adjointness of the
transposed convolutions, symmetry and semi-definiteness of M, and the analytic gradient
against central differences of the FP64 loss."""
import numpy as np
import pytest
import scipy.sparse as sp

import learned as lrn


def q1_scaled_operator(n):
    """Jacobi-scaled Q1 stiffness (8/3 centre, -1/3 neighbours) on an n x n node grid with
    Dirichlet boundary nodes: A' on node vectors, mask of the unknown nodes."""
    N = n * n
    mask = np.zeros(N)
    rows, cols, vals = [], [], []
    idx = lambda i, j: i * n + j
    for i in range(1, n - 1):
        for j in range(1, n - 1):
            mask[idx(i, j)] = 1
            for di in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    ii, jj = i + di, j + dj
                    if 1 <= ii < n - 1 and 1 <= jj < n - 1:
                        rows.append(idx(i, j)); cols.append(idx(ii, jj))
                        vals.append(1.0 if (di == 0 and dj == 0) else -1.0 / 8.0)
    return sp.csr_matrix((vals, (rows, cols)), shape=(N, N)), mask


def random_params(lay, rng, scale=0.3):
    p = rng.standard_normal(lay["nparams"]) * scale
    # only input channel 0 of In exists
    w = p[lay["off_in"]: lay["off_in"] + 9 * lay["C"][0] * 16].reshape(9, lay["C"][0], 16)
    w[:, :, 1:] = 0.0
    return p


def test_transposed_convolutions_are_adjoints():
    rng = np.random.default_rng(0)
    for H, W in ((9, 9), (10, 7)):
        Ho, Wo = (H + 1) // 2, (W + 1) // 2
        x = rng.standard_normal((H, W, 16))
        y = rng.standard_normal((Ho, Wo, 32))
        w = rng.standard_normal((9, 32, 16))
        lhs = np.sum(lrn.conv_down(x, w, Ho, Wo) * y)
        rhs = np.sum(x * lrn.conv_up(y, w, H, W))
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
        z = rng.standard_normal((H, W, 32))
        lhs = np.sum(lrn.conv_same(x, w) * z)
        rhs = np.sum(x * lrn.conv_same(z, np.transpose(w, (0, 2, 1)), sign=-1))
        assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_network_is_symmetric_positive_semidefinite():
    rng = np.random.default_rng(1)
    n = 17
    lay = lrn.layout(n, n, [16, 16, 32], max_levels=4)
    p = random_params(lay, rng)
    _, mask = q1_scaled_operator(n)
    u = rng.standard_normal(n * n) * mask
    v = rng.standard_normal(n * n) * mask
    Mu = lrn.forward(p, lay, u, mask)[0]
    Mv = lrn.forward(p, lay, v, mask)[0]
    assert abs(u @ Mv - v @ Mu) <= 1e-10 * (abs(u @ Mv) + 1e-30)
    assert v @ Mv >= -1e-12 * np.linalg.norm(v) ** 2


def test_level_sizes_follow_vertex_coarsening():
    H, W = lrn.level_sizes(513, 513)
    assert H == [513, 257, 129, 65, 33, 17, 9, 5, 3]
    H, W = lrn.level_sizes(65, 65, 10)
    assert H == [65, 33, 17, 9, 5, 3]


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(2)
    n = 13
    lay = lrn.layout(n, n, [16, 16, 32], max_levels=3)
    A, mask = q1_scaled_operator(n)
    p = random_params(lay, rng)
    pairs = []
    for _ in range(2):
        e = rng.standard_normal(n * n) * mask
        pairs.append((A @ e, e))
    loss, grad = lrn.loss_and_grad(p, lay, pairs, A, mask)
    assert np.isclose(loss, lrn.loss_only(p, lay, pairs, A, mask), rtol=1e-12)
    eps = 1e-6
    # one coordinate in every tensor, plus random directions
    coords = [lay["off_in"] + (4 * lay["C"][0] + 3) * 16]
    coords += [lay["off_down"][l] + 5 for l in range(1, lay["levels"])]
    coords += [lay["off_q"][l] + 7 for l in range(lay["levels"])]
    for c in coords:
        d = np.zeros_like(p); d[c] = 1.0
        fd = (lrn.loss_only(p + eps * d, lay, pairs, A, mask) - lrn.loss_only(p - eps * d, lay, pairs, A, mask)) / (2 * eps)
        assert abs(fd - grad[c]) <= 1e-5 * max(1.0, abs(fd)), (c, fd, grad[c])
    for _ in range(3):
        d = random_params(lay, rng, 1.0)
        fd = (lrn.loss_only(p + eps * d, lay, pairs, A, mask) - lrn.loss_only(p - eps * d, lay, pairs, A, mask)) / (2 * eps)
        assert abs(fd - grad @ d) <= 1e-6 * max(1.0, abs(fd))


def test_planted_sign_error_in_gradient_is_caught():
    """The central-difference check detects a dropped transposed-use term (the kind of
    mistake the adjoint derivation invites)."""
    rng = np.random.default_rng(3)
    n = 9
    lay = lrn.layout(n, n, [16], max_levels=2)
    A, mask = q1_scaled_operator(n)
    p = random_params(lay, rng)
    e = rng.standard_normal(n * n) * mask
    pairs = [(A @ e, e)]
    _, grad = lrn.loss_and_grad(p, lay, pairs, A, mask)
    # recompute dIn without the corr(D_0, g) term
    w_in, w_dn, q = lrn.unpack(p, lay)
    y, E, D = lrn.forward(p, lay, pairs[0][0], mask)
    ea = e @ (A @ e)
    g = 2 * mask * (A @ (y - e)) / ea
    S = [qq @ qq.T for qq in q]
    G = lrn._encode(g, w_in, w_dn, lay, lambda a: a)
    Hh = lrn._decode(G, w_dn, S, lay, lambda a: a)
    rin = np.zeros((n, n, 16)); rin[:, :, 0] = pairs[0][0].reshape(n, n)
    wrong = lrn.corr_same(Hh[0], rin).reshape(-1)
    c = lay["off_in"] + (4 * lay["C"][0] + 0) * 16
    d = np.zeros_like(p); d[c] = 1.0
    eps = 1e-6
    fd = (lrn.loss_only(p + eps * d, lay, pairs, A, mask) - lrn.loss_only(p - eps * d, lay, pairs, A, mask)) / (2 * eps)
    assert abs(fd - grad[c]) <= 1e-5 * max(1.0, abs(fd))
    assert abs(fd - wrong[c - lay["off_in"]]) > 1e-3 * max(1.0, abs(fd))
