"""CPU restatement of the learned preconditioner (numpy, FP64 with optional BF16 emulation).

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker of the GPU network kernels
(pic-_amd/csrc/k_learn.hip), never by the library. The network has no counterpart in the
reference; this file restates the build's own definition:

  E_0 = In r (3x3), E_l = Down_l E_{l-1} (3x3, stride 2: coarse node I <-> fine node 2I),
  D_{L-1} = S_{L-1} E_{L-1}, D_l = S_l E_l + Down_{l+1}^T D_{l+1}, y = m * In^T D_0,
  S_l = Q_l Q_l^T, loss = sum_pairs (y - e)^T A' (y - e) / (e^T A' e).

Weights are stored per tap t = 3*(di+1) + (dj+1) as W[t][co][ci]; activations as
(H, W, C) arrays (node (i, j) -> [i, j]). The gradient is computed by the adjoint
recursions (G = encoder of g, H = decoder of G) and checked against central differences
of the loss in tests/test_learned_cpu.py.
"""
from __future__ import annotations

import numpy as np

TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def level_sizes(H0: int, W0: int, max_levels: int = 10):
    H, W = [H0], [W0]
    while len(H) < max_levels and min(H[-1], W[-1]) > 3:
        H.append((H[-1] + 1) // 2)
        W.append((W[-1] + 1) // 2)
    return H, W


def unpack(params, lay):
    """Views of the flat parameter vector in the library's layout."""
    L, C = lay["levels"], lay["C"]
    p = np.asarray(params, dtype=np.float64)
    w_in = p[lay["off_in"]: lay["off_in"] + 9 * C[0] * 16].reshape(9, C[0], 16)
    w_dn = [None] + [p[lay["off_down"][l]: lay["off_down"][l] + 9 * C[l] * C[l - 1]].reshape(9, C[l], C[l - 1])
                     for l in range(1, L)]
    q = [p[lay["off_q"][l]: lay["off_q"][l] + C[l] * C[l]].reshape(C[l], C[l]) for l in range(L)]
    return w_in, w_dn, q


def layout(H0: int, W0: int, channels, max_levels: int = 10) -> dict:
    H, W = level_sizes(H0, W0, max_levels)
    L = len(H)
    C = [channels[min(l, len(channels) - 1)] for l in range(L)]
    off = 0
    off_in = off
    off += 9 * C[0] * 16
    off_down = [None]
    for l in range(1, L):
        off_down.append(off)
        off += 9 * C[l] * C[l - 1]
    off_q = []
    for l in range(L):
        off_q.append(off)
        off += C[l] * C[l]
    return {"levels": L, "H": H, "W": W, "C": C, "off_in": off_in, "off_down": off_down, "off_q": off_q,
            "nparams": off}


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


def conv_down(x, w, Ho, Wo):
    """out[I, J] = sum_t w[t] @ x[2I + di, 2J + dj] (zero outside)."""
    H, W, _ = x.shape
    xp = np.zeros((2 * Ho + 2, 2 * Wo + 2, x.shape[2]))
    xp[1:H + 1, 1:W + 1] = x
    out = np.zeros((Ho, Wo, w.shape[1]))
    for t, (di, dj) in enumerate(TAPS):
        out += xp[1 + di:1 + di + 2 * Ho:2, 1 + dj:1 + dj + 2 * Wo:2] @ w[t].T
    return out


def conv_up(y, w, H, W):
    """Adjoint of conv_down with the same w: out[2I + di, 2J + dj] += w[t]^T @ y[I, J]."""
    Ho, Wo, _ = y.shape
    outp = np.zeros((2 * Ho + 2, 2 * Wo + 2, w.shape[2]))
    for t, (di, dj) in enumerate(TAPS):
        outp[1 + di:1 + di + 2 * Ho:2, 1 + dj:1 + dj + 2 * Wo:2] += y @ w[t]
    return outp[1:H + 1, 1:W + 1]


def corr_same(o, x):
    """dW[t][co][ci] = sum_p o[p, co] x[p + d_t, ci]."""
    H, W, _ = o.shape
    xp = np.zeros((H + 2, W + 2, x.shape[2]))
    xp[1:-1, 1:-1] = x
    return np.stack([np.einsum("pqo,pqi->oi", o, xp[1 + di:1 + di + H, 1 + dj:1 + dj + W]) for di, dj in TAPS])


def corr_down(o, x):
    """dW[t][co][ci] = sum_P o[P, co] x[2P + d_t, ci]."""
    Ho, Wo, _ = o.shape
    H, W, _ = x.shape
    xp = np.zeros((2 * Ho + 2, 2 * Wo + 2, x.shape[2]))
    xp[1:H + 1, 1:W + 1] = x
    return np.stack([np.einsum("pqo,pqi->oi", o, xp[1 + di:1 + di + 2 * Ho:2, 1 + dj:1 + dj + 2 * Wo:2])
                     for di, dj in TAPS])


# ------------------------------------------------------------------ network
def _encode(x0, w_in, w_dn, lay, rnd):
    H, W, C = lay["H"], lay["W"], lay["C"]
    xin = np.zeros((H[0], W[0], 16))
    xin[:, :, 0] = rnd(x0.reshape(H[0], W[0]))
    E = [rnd(conv_same(xin, rnd(w_in)))]
    for l in range(1, lay["levels"]):
        E.append(rnd(conv_down(E[l - 1], rnd(w_dn[l]), H[l], W[l])))
    return E


def _decode(E, w_dn, S, lay, rnd):
    L = lay["levels"]
    H, W = lay["H"], lay["W"]
    D = [None] * L
    D[L - 1] = rnd(E[L - 1] @ rnd(S[L - 1]).T)
    for l in range(L - 2, -1, -1):
        D[l] = rnd(conv_up(D[l + 1], rnd(w_dn[l + 1]), H[l], W[l]) + E[l] @ rnd(S[l]).T)
    return D


def forward(params, lay, r, mask, emulate_bf16: bool = False):
    """y = M r on a flat node vector r (length H0*W0); returns y and the tensors E, D."""
    rnd = bf16 if emulate_bf16 else (lambda a: a)
    w_in, w_dn, q = unpack(params, lay)
    S = [qq @ qq.T for qq in q]
    E = _encode(np.asarray(r, np.float64), w_in, w_dn, lay, rnd)
    D = _decode(E, w_dn, S, lay, rnd)
    y = conv_same(D[0], rnd(np.transpose(w_in, (0, 2, 1))), sign=-1)[:, :, 0].reshape(-1)
    return y * (np.asarray(mask) > 0), E, D


def loss_and_grad(params, lay, pairs, A, mask, emulate_bf16: bool = False):
    """Summed loss over pairs [(r, e)] and its gradient (flat, library layout).
    A: the scaled operator (scipy sparse or dense) on node vectors."""
    rnd = bf16 if emulate_bf16 else (lambda a: a)
    L, H, W, C = lay["levels"], lay["H"], lay["W"], lay["C"]
    w_in, w_dn, q = unpack(params, lay)
    S = [qq @ qq.T for qq in q]
    m = (np.asarray(mask) > 0).astype(np.float64)
    grad = np.zeros(lay["nparams"])
    total = 0.0
    for r, e in pairs:
        r = np.asarray(r, np.float64)
        e = np.asarray(e, np.float64)
        y, E, D = forward(params, lay, r, mask, emulate_bf16)
        ea = float(e @ (A @ e))
        res = y - e
        Ares = A @ res
        total += float(res @ Ares) / ea
        g = 2.0 * m * Ares / ea
        G = _encode(g, w_in, w_dn, lay, rnd)
        Hh = _decode(G, w_dn, S, lay, rnd)
        rin = np.zeros((H[0], W[0], 16)); rin[:, :, 0] = rnd(r.reshape(H[0], W[0]))
        gin = np.zeros((H[0], W[0], 16)); gin[:, :, 0] = rnd(g.reshape(H[0], W[0]))
        d_in = corr_same(Hh[0], rin) + corr_same(D[0], gin)
        grad[lay["off_in"]: lay["off_in"] + d_in.size] += d_in.reshape(-1)
        for l in range(1, L):
            d = corr_down(Hh[l], E[l - 1]) + corr_down(D[l], G[l - 1])
            grad[lay["off_down"][l]: lay["off_down"][l] + d.size] += d.reshape(-1)
        for l in range(L):
            dS = np.einsum("pqo,pqi->oi", G[l], E[l])
            dQ = (dS + dS.T) @ q[l]
            grad[lay["off_q"][l]: lay["off_q"][l] + dQ.size] += dQ.reshape(-1)
    return total, grad


def loss_only(params, lay, pairs, A, mask):
    tot = 0.0
    for r, e in pairs:
        y, _, _ = forward(params, lay, r, mask)
        res = y - np.asarray(e, np.float64)
        tot += float(res @ (A @ res)) / float(np.asarray(e) @ (A @ np.asarray(e)))
    return tot


def scaled_operator(rowptr, col, val, diag, unknown_node, nnodes):
    """A' = D^-1/2 A D^-1/2 in node numbering (scipy CSR), with the D^-1/2 vector."""
    import scipy.sparse as sp
    un = np.asarray(unknown_node) - 1
    n = len(diag)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    s = 1.0 / np.sqrt(np.asarray(diag))
    v = np.asarray(val) * s[rows] * s[np.asarray(col)]
    A = sp.csr_matrix((v, (un[rows], un[np.asarray(col)])), shape=(nnodes, nnodes))
    sinv = np.zeros(nnodes)
    sinv[un] = s
    return A, sinv
