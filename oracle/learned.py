"""CPU restatement of the learned multigrid preconditioner (numpy/scipy, FP64).

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the GPU kernels in
pic-_amd/csrc/k_learn.hip, never by the library. The network has no counterpart in the
reference; this file restates the build's own definition.

Hierarchy: level 0 is the Jacobi-scaled node-space operator A' (unit diagonal on unknown
nodes). Level l+1 has the nodes (I, J) <-> fine (2I, 2J); P_l is bilinear interpolation
with the rows of fine Dirichlet nodes (zero diagonal) removed, R_l = P_l^T, and
A_{l+1} = P_l^T A_l P_l (Galerkin). Coarsening stops at the first level with at most
DENSE_NODES nodes once every learned level has a coarser one (or at about 3 x 3, or at
max_levels); that coarsest level is solved exactly with a dense inverse.

Symmetric V(1,1)-cycle y = V_l(r):
    x  = K_l r
    rc = R_l (r - A_l x);  ec = V_{l+1}(rc)
    x2 = x + P_l ec
    y  = x2 + K_l (r - A_l x2)
Smoothers: learned on levels l < Ls,  K_l = Dh In_l^T S_l In_l Dh  (Dh = diag(A_l)^-1/2,
In_l a 3x3 conv 1 -> C_l channels with weights W_in[t][c], S_l = Q_l Q_l^T), damped Jacobi
omega/diag elsewhere. Parameters (flat, the library's layout): per learned level W_in [9][C],
then per learned level Q [C][C].
Loss over pairs (r, e = A'^-1 r):  sum (y - e)^T A' (y - e) / (e^T A' e).
For symmetric K and M, with g = dL/dy, the gradient of a learned smoother is
    dL/dK = s2(g) (x) r_l(r) + r_l(g) (x) s2(r)
where r_l(.) and s2(.) are the inputs of its pre- and post-smoothing in the V-cycle run on
r or on g (see tests/test_learned_cpu.py for the central-difference check). This file keeps
the factored form (lifts U = In Dh u, V = In Dh v, correlations over pixels); the library
contracts the same gradient to a tap-pair Gram matrix, so agreement checks that algebra.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]
OMEGA = 0.8


DENSE_NODES = 1089            # 33 x 33: the largest coarsest level solved by its dense inverse


def level_sizes(H0: int, W0: int, max_levels: int = 12, learned_levels: int = 2, dense_nodes: int = DENSE_NODES):
    """Node-grid sizes of the hierarchy: halve (vertex-centred) while the coarsest level has
    more than dense_nodes nodes or fewer than max(2, learned_levels + 1) levels exist, down to
    about 3 x 3 and at most max_levels levels (dense_nodes = 0: always down to 3 x 3)."""
    H, W = [H0], [W0]
    min_levels = max(2, learned_levels + 1)
    while (len(H) < max_levels and min(H[-1], W[-1]) > 3 and
           (H[-1] * W[-1] > dense_nodes or len(H) < min_levels)):
        H.append((H[-1] + 1) // 2)
        W.append((W[-1] + 1) // 2)
    return H, W


def layout(H0: int, W0: int, channels, learned_levels: int = 2, max_levels: int = 12,
           dense_nodes: int = DENSE_NODES) -> dict:
    H, W = level_sizes(H0, W0, max_levels, learned_levels, dense_nodes)
    L = len(H)
    Ls = min(learned_levels, L - 1)
    C = [channels[min(l, len(channels) - 1)] for l in range(Ls)]
    off, off_in, off_q = 0, [], []
    for l in range(Ls):
        off_in.append(off)
        off += 9 * C[l]
    for l in range(Ls):
        off_q.append(off)
        off += C[l] * C[l]
    return {"levels": L, "learned": Ls, "H": H, "W": W, "C": C, "off_in": off_in, "off_q": off_q, "nparams": off}


def unpack(params, lay):
    p = np.asarray(params, dtype=np.float64)
    w_in = [p[o:o + 9 * c].reshape(9, c, 1) for o, c in zip(lay["off_in"], lay["C"])]
    q = [p[o:o + c * c].reshape(c, c) for o, c in zip(lay["off_q"], lay["C"])]
    return w_in, q


# ------------------------------------------------------------------ hierarchy
def prolongation(Hf, Wf, Hc, Wc, maskf):
    rows, cols, vals = [], [], []
    for i in range(Hf):
        ci = [(i // 2, 1.0)] if i % 2 == 0 else [((i - 1) // 2, 0.5), ((i + 1) // 2, 0.5)]
        for j in range(Wf):
            f = i * Wf + j
            if maskf[f] == 0:
                continue
            cj = [(j // 2, 1.0)] if j % 2 == 0 else [((j - 1) // 2, 0.5), ((j + 1) // 2, 0.5)]
            for I, wi in ci:
                for J, wj in cj:
                    if I < Hc and J < Wc:
                        rows.append(f); cols.append(I * Wc + J); vals.append(wi * wj)
    return sp.csr_matrix((vals, (rows, cols)), shape=(Hf * Wf, Hc * Wc))


def build_hierarchy(A0, H0, W0, max_levels: int = 12, learned_levels: int = 2, dense_nodes: int = DENSE_NODES):
    H, W = level_sizes(H0, W0, max_levels, learned_levels, dense_nodes)
    levels = []
    A = sp.csr_matrix(A0)
    for l in range(len(H)):
        d = A.diagonal()
        mask = (d != 0).astype(np.float64)
        lev = {"H": H[l], "W": W[l], "A": A, "d": d, "mask": mask,
               "dinv": np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 0.0),
               "dh": np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 0.0)}
        if l + 1 < len(H):
            P = prolongation(H[l], W[l], H[l + 1], W[l + 1], mask)
            lev["P"] = P
            A = (P.T @ A @ P).tocsr()
        levels.append(lev)
    last = levels[-1]
    idx = np.nonzero(last["mask"])[0]
    inv = np.zeros((last["H"] * last["W"],) * 2)
    if len(idx):
        inv[np.ix_(idx, idx)] = np.linalg.inv(last["A"].toarray()[np.ix_(idx, idx)])
    last["inv"] = inv
    return levels


# ------------------------------------------------------------------ convolutions
def conv_same(x, w, sign=1):
    """out[p] = sum_t w[t] @ x[p + sign*d_t] (zero outside)."""
    H, W, _ = x.shape
    out = np.zeros((H, W, w.shape[1]))
    xp = np.zeros((H + 2, W + 2, x.shape[2]))
    xp[1:-1, 1:-1] = x
    for t, (di, dj) in enumerate(TAPS):
        si, sj = sign * di, sign * dj
        out += xp[1 + si:1 + si + H, 1 + sj:1 + sj + W] @ w[t].T
    return out


def corr_same(o, x):
    """dW[t][co][ci] = sum_p o[p, co] x[p + d_t, ci]."""
    H, W, _ = o.shape
    xp = np.zeros((H + 2, W + 2, x.shape[2]))
    xp[1:-1, 1:-1] = x
    return np.stack([np.einsum("pqo,pqi->oi", o, xp[1 + di:1 + di + H, 1 + dj:1 + dj + W]) for di, dj in TAPS])


def _as_chan(v, H, W):
    x = np.zeros((H, W, 1))
    x[:, :, 0] = v.reshape(H, W)
    return x


# ------------------------------------------------------------------ network
class _Net:
    """The V-cycle in FP64 (the library applies the smoother as the collapsed FP32 stencil,
    which equals the factored form up to rounding)."""
    def __init__(self, params, lay, levels):
        self.lay, self.levels = lay, levels
        w_in, q = unpack(params, lay)
        self.w_in, self.q = w_in, q
        self.S = [qq @ qq.T for qq in q]
        self.SW = [np.einsum("oc,tci->toi", S, w) for S, w in zip(self.S, w_in)]   # S In, [9][C][1]
        self.WT = [np.transpose(w, (0, 2, 1)) for w in w_in]                          # In^T, [9][1][C]

    def smooth(self, l, v):
        lev = self.levels[l]
        if l < self.lay["learned"]:
            H, W = lev["H"], lev["W"]
            u = _as_chan(lev["dh"] * v, H, W)
            E = conv_same(u, self.SW[l])
            return conv_same(E, self.WT[l], sign=-1)[:, :, 0].reshape(-1) * lev["dh"]
        return OMEGA * lev["dinv"] * v

    def vcycle(self, l, r, sites):
        lev = self.levels[l]
        if l == len(self.levels) - 1:
            return lev["inv"] @ r
        A, P = lev["A"], lev["P"]
        x = self.smooth(l, r)
        ec = self.vcycle(l + 1, P.T @ (r - A @ x), sites)
        x2 = x + P @ ec
        s2 = r - A @ x2
        if l < self.lay["learned"]:
            sites[l] = (r.copy(), s2.copy())
        return x2 + self.smooth(l, s2)


def collapsed_pairs(w_in, q):
    """The learned smoother In^T S In as 81 tap-pair weights k81[t][t2] (the library's form)."""
    S = q @ q.T
    W = w_in[:, :, 0]
    return W @ S @ W.T


def apply_collapsed(k81, u):
    """y(p) = sum_t [p - d_t on the grid] sum_t2 k81[t][t2] u(p - d_t + d_t2) on a 2-D field."""
    H, W = u.shape
    up = np.zeros((H + 4, W + 4)); up[2:-2, 2:-2] = u
    y = np.zeros((H, W))
    for t, (ti, tj) in enumerate(TAPS):
        valid = np.zeros((H, W), bool)
        valid[max(0, ti):H + min(0, ti), max(0, tj):W + min(0, tj)] = True   # p - d_t on the grid
        acc = np.zeros((H, W))
        for t2, (si, sj) in enumerate(TAPS):
            oi, oj = 2 - ti + si, 2 - tj + sj
            acc += k81[t, t2] * up[oi:oi + H, oj:oj + W]
        y += np.where(valid, acc, 0.0)
    return y


def collapsed_stencil(k81):
    """The 5 x 5 stencil k25[di+2][dj+2] = sum over tap pairs with d_t2 - d_t = (di, dj) of
    k81[t][t2]: equal to the tap-pair form wherever every p - d_t is on the grid."""
    k25 = np.zeros((5, 5))
    for t, (ti, tj) in enumerate(TAPS):
        for t2, (si, sj) in enumerate(TAPS):
            k25[si - ti + 2, sj - tj + 2] += k81[t, t2]
    return k25


def apply_stencil(k25, u):
    """y(p) = sum_delta k25[delta] u(p + delta), zero outside the grid."""
    H, W = u.shape
    up = np.zeros((H + 4, W + 4)); up[2:-2, 2:-2] = u
    y = np.zeros((H, W))
    for di in range(5):
        for dj in range(5):
            y += k25[di, dj] * up[di:di + H, dj:dj + W]
    return y


def forward(params, lay, levels, r):
    """y = M r on a flat node vector; returns y and the smoother inputs per learned level."""
    net = _Net(params, lay, levels)
    sites = {}
    y = net.vcycle(0, np.asarray(r, np.float64), sites)
    return y, sites


def loss_and_grad(params, lay, levels, pairs):
    """Summed loss over pairs [(r, e)] and its gradient (flat, library layout)."""
    net = _Net(params, lay, levels)
    A = levels[0]["A"]
    m = levels[0]["mask"]
    grad = np.zeros(lay["nparams"])
    total = 0.0
    for r, e in pairs:
        r = np.asarray(r, np.float64)
        e = np.asarray(e, np.float64)
        sr = {}
        y = net.vcycle(0, r, sr)
        ea = float(e @ (A @ e))
        res = y - e
        Ares = A @ res
        total += float(res @ Ares) / ea
        g = 2.0 * m * Ares / ea
        sg = {}
        net.vcycle(0, g, sg)
        for l in range(lay["learned"]):
            lev = levels[l]
            H, W, dh = lev["H"], lev["W"], lev["dh"]
            dS = np.zeros_like(net.S[l])
            dW = np.zeros_like(net.w_in[l])
            for u, v in ((sg[l][1], sr[l][0]), (sg[l][0], sr[l][1])):
                uu = _as_chan(dh * u, H, W)
                vv = _as_chan(dh * v, H, W)
                U = conv_same(uu, net.w_in[l])
                V = conv_same(vv, net.w_in[l])
                SU = conv_same(uu, net.SW[l])
                SV = conv_same(vv, net.SW[l])
                dS += np.einsum("pqo,pqi->oi", U, V)
                dW += corr_same(SV, uu) + corr_same(SU, vv)
            dQ = (dS + dS.T) @ net.q[l]
            grad[lay["off_in"][l]: lay["off_in"][l] + dW.size] += dW.reshape(-1)
            grad[lay["off_q"][l]: lay["off_q"][l] + dQ.size] += dQ.reshape(-1)
    return total, grad


def loss_only(params, lay, levels, pairs):
    A = levels[0]["A"]
    tot = 0.0
    for r, e in pairs:
        y, _ = forward(params, lay, levels, r)
        res = y - np.asarray(e, np.float64)
        tot += float(res @ (A @ res)) / float(np.asarray(e) @ (A @ np.asarray(e)))
    return tot


def initial_params(lay, seed: int = 1, small: float = 0.01):
    """The library's initialisation: channel 0 of In is the identity tap, S = omega e0 e0^T
    (Q_00 = sqrt(omega)), so the untrained smoother is damped Jacobi."""
    rng = np.random.default_rng(seed)
    p = np.zeros(lay["nparams"])
    for l in range(lay["learned"]):
        C = lay["C"][l]
        w = p[lay["off_in"][l]: lay["off_in"][l] + 9 * C].reshape(9, C)
        w[:, :] = small * rng.uniform(-1, 1, (9, C))
        w[:, 0] = 0.0
        w[4, 0] = 1.0
        q = p[lay["off_q"][l]: lay["off_q"][l] + C * C].reshape(C, C)
        q[:] = small * rng.uniform(-1, 1, (C, C))
        q[0, 0] = np.sqrt(OMEGA)
    return p


def scaled_operator(rowptr, col, val, diag, unknown_node, nnodes):
    """A' = D^-1/2 A D^-1/2 in node numbering (scipy CSR), with the D^-1/2 vector."""
    un = np.asarray(unknown_node) - 1
    n = len(diag)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    s = 1.0 / np.sqrt(np.asarray(diag))
    v = np.asarray(val) * s[rows] * s[np.asarray(col)]
    A = sp.csr_matrix((v, (un[rows], un[np.asarray(col)])), shape=(nnodes, nnodes))
    sinv = np.zeros(nnodes)
    sinv[un] = s
    return A, sinv
