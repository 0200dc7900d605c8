"""Oracle (test infrastructure, never on the product path) of the learned warm start of the
field solve: a linear convolutional encoder-decoder on the node grid and the gradient of its
energy-norm loss, in FP64 numpy.

The network has no counterpart in the reference (the reference's IFE_Solve_2D.f90:73-98 warm
starts from the previous potential only). What it must reproduce is the library's definition:

  h_0 = m_0 . lift(r)                         3x3, 1 -> C
  h_l = m_l . conv(R h_{l-1}; E_l)            l = 1..L-1
  g_{L-1} = h_{L-1}
  g_l = m_l . conv(P g_{l+1} + h_l; D_l)      l = L-2..0
  y   = m_0 . out(g_0)                        3x3, C -> 1

with taps t <-> node offset (t // 3 - 1, t % 3 - 1), zero outside the grid, level l+1 holding
the nodes (I, J) <-> (2I, 2J) of level l, P bilinear interpolation from the coarse nodes
inside the grid, R = P^T / 4, m_l the unknown-node masks (m_{l+1}(I, J) = m_l(2I, 2J)).
Loss per pair (r = A' e): (y - e)^T A' (y - e) / (e^T A' e).

`bf16=True` rounds to bfloat16 (round to nearest even, through FP32) exactly where the
library stores or stages BF16 values: the C -> C weights, every C-channel activation and
backward quantity, and the staged convolution inputs (R h, P g + h). Then the library's
results differ from this oracle only by its FP32 summation order. The backward pass is
derived here from the forward definition (adjoint of each step); the CPU tests pin it by
central differences of the FP64 loss and by adjoint identities.
"""
from __future__ import annotations

import numpy as np


def level_sizes(H0: int, W0: int, levels: int = 6):
    H, W = [H0], [W0]
    while len(H) < levels and min(H[-1], W[-1]) > 5:
        H.append((H[-1] + 1) // 2)
        W.append((W[-1] + 1) // 2)
    return H, W


def level_masks(mask0, H, W):
    """m_0 from the unknown mask (node numbering i * W + j), m_{l+1}(I, J) = m_l(2I, 2J)."""
    ms = [np.asarray(mask0, np.float64).reshape(H[0], W[0])]
    for l in range(1, len(H)):
        ms.append(ms[-1][0::2, 0::2][: H[l], : W[l]].copy())
    return ms


def layout(H, W, C: int):
    L = len(H)
    off, o = {}, 0
    off["lift"] = o; o += 9 * C
    off["out"] = o; o += 9 * C
    for l in range(1, L):
        off[f"enc{l}"] = o; o += 9 * C * C
    for l in range(L - 1):
        off[f"dec{l}"] = o; o += 9 * C * C
    return {"L": L, "C": C, "H": list(H), "W": list(W), "off": off, "nparams": o}


def unpack(p, lay):
    C, off = lay["C"], lay["off"]
    w = {"lift": p[off["lift"]: off["lift"] + 9 * C].reshape(9, C),
         "out": p[off["out"]: off["out"] + 9 * C].reshape(9, C)}
    for k, o in off.items():
        if k.startswith("enc") or k.startswith("dec"):
            w[k] = p[o: o + 9 * C * C].reshape(9, C, C)
    return w


def bf16(x):
    f = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(x, di, dj):
    """x(i + di, j + dj) with zeros outside the grid (x: [H, W, ...])."""
    H, W = x.shape[:2]
    out = np.zeros_like(x)
    i0, i1 = max(0, -di), min(H, H - di)
    j0, j1 = max(0, -dj), min(W, W - dj)
    out[i0:i1, j0:j1] = x[i0 + di:i1 + di, j0 + dj:j1 + dj]
    return out


def conv(x, w):
    """y(p)[o] = sum_t sum_i w[t][i][o] x(p + d_t)[i]; x [H, W, Ci], w [9, Ci, Co]."""
    y = 0.0
    for t in range(9):
        y = y + _shift(x, t // 3 - 1, t % 3 - 1) @ w[t]
    return y


def conv_adjoint(z, w):
    """Adjoint of conv in x: x_bar(q)[i] = sum_t sum_o w[t][i][o] z(q - d_t)[o]."""
    wt = np.transpose(w[::-1], (0, 2, 1))          # W'[t][o][i] = W[8 - t][i][o]
    return conv(z, wt)


def lift(s, w):
    """1 -> C: h(p)[c] = sum_t w[t][c] s(p + d_t); s [H, W], w [9, C]."""
    return conv(s[:, :, None], w[:, None, :])


def out_conv(x, w):
    """C -> 1: y(p) = sum_t sum_c w[t][c] x(p + d_t)[c]."""
    return conv(x, w[:, :, None])[:, :, 0]


def restrict(x, Hc, Wc):
    """(R x)(I, J) = sum_{u, v in -1..1} w_u w_v x(2I + u, 2J + v), w = 1/2 (centre), 1/4; fine
    nodes inside the grid only."""
    Hf, Wf = x.shape[:2]
    xp = np.zeros((Hf + 3, Wf + 3) + x.shape[2:])
    xp[1:Hf + 1, 1:Wf + 1] = x
    y = 0.0
    for du in (-1, 0, 1):
        for dv in (-1, 0, 1):
            wgt = (0.5 if du == 0 else 0.25) * (0.5 if dv == 0 else 0.25)
            y = y + wgt * xp[1 + du: 1 + du + 2 * Hc: 2, 1 + dv: 1 + dv + 2 * Wc: 2]
    return y


def interp(x, Hf, Wf):
    """Bilinear interpolation from coarse nodes (I, J) <-> fine (2I, 2J), coarse nodes outside
    the grid read as zero."""
    Hc, Wc = x.shape[:2]
    xp = np.zeros((Hc + 1, Wc + 1) + x.shape[2:])
    xp[:Hc, :Wc] = x
    y = np.zeros((Hf, Wf) + x.shape[2:])
    ie, io, je, jo = (Hf + 1) // 2, Hf // 2, (Wf + 1) // 2, Wf // 2
    y[0::2, 0::2] = xp[:ie, :je]
    y[1::2, 0::2] = 0.5 * (xp[:io, :je] + xp[1:io + 1, :je])
    y[0::2, 1::2] = 0.5 * (xp[:ie, :jo] + xp[:ie, 1:jo + 1])
    y[1::2, 1::2] = 0.25 * (xp[:io, :jo] + xp[1:io + 1, :jo] + xp[:io, 1:jo + 1] + xp[1:io + 1, 1:jo + 1])
    return y


def forward(p, lay, masks, r, bf16_round: bool = False, mistake: str | None = None):
    """y = N(r) for one node vector r; returns (y flat, cache for the backward pass)."""
    L, H, W = lay["L"], lay["H"], lay["W"]
    q = bf16 if bf16_round else (lambda v: v)
    w = unpack(np.asarray(p, np.float64), lay)
    wc = {k: (q(v) if k.startswith(("enc", "dec")) else v) for k, v in w.items()}
    m = [mk[:, :, None] for mk in masks]
    r2 = np.asarray(r, np.float64).reshape(H[0], W[0])
    h, t, u, g = [None] * L, [None] * L, [None] * L, [None] * L
    h[0] = q(m[0] * lift(r2, wc["lift"]))
    for l in range(1, L):
        t[l] = q(restrict(h[l - 1], H[l], W[l]))
        h[l] = q(m[l] * conv(t[l], wc[f"enc{l}"]))
    g[L - 1] = h[L - 1]
    for l in range(L - 2, -1, -1):
        up = interp(g[l + 1], H[l], W[l])
        if mistake == "no_skip":
            u[l] = q(up)
        else:
            u[l] = q(up + h[l])
        g[l] = q(m[l] * conv(u[l], wc[f"dec{l}"]))
    y = masks[0] * out_conv(g[0], wc["out"])
    return y.reshape(-1), {"r": r2, "h": h, "t": t, "u": u, "g": g, "w": wc}


def backward(lay, masks, cache, gy, bf16_round: bool = False, mistake: str | None = None):
    """Gradient of <gy, N(r)> in the parameters (flat, the library's layout)."""
    L, H, W, C = lay["L"], lay["H"], lay["W"], lay["C"]
    q = bf16 if bf16_round else (lambda v: v)
    m = [mk[:, :, None] for mk in masks]
    w, h, t, u, g = cache["w"], cache["h"], cache["t"], cache["u"], cache["g"]
    gy2 = np.asarray(gy, np.float64).reshape(H[0], W[0]) * masks[0]
    grad = {}
    # y = m0 . out(g0)
    grad["out"] = np.stack([np.sum(_shift(g[0], tt // 3 - 1, tt % 3 - 1) * gy2[:, :, None], axis=(0, 1))
                            for tt in range(9)])
    wl = w["out"][::-1] if mistake != "no_flip" else w["out"]
    gb = [None] * L
    gb[0] = q(m[0] * lift(gy2, wl))               # m0 . out^T(m0 gy)
    ub, hb, tb = [None] * L, [None] * L, [None] * L

    def wgrad(x, z):
        return np.stack([np.einsum("ijc,ijd->cd", _shift(x, tt // 3 - 1, tt % 3 - 1), z) for tt in range(9)])

    for l in range(L - 1):
        grad[f"dec{l}"] = wgrad(u[l], gb[l])
        ub[l] = q(conv_adjoint(gb[l], w[f"dec{l}"]))
        gb[l + 1] = q(m[l + 1] * 4.0 * restrict(ub[l], H[l + 1], W[l + 1]))
    for l in range(L - 1, 0, -1):
        z = gb[L - 1] if l == L - 1 else hb[l]
        grad[f"enc{l}"] = wgrad(t[l], z)
        tb[l] = q(conv_adjoint(z, w[f"enc{l}"]))
        hb[l - 1] = q(m[l - 1] * (ub[l - 1] + 0.25 * interp(tb[l], H[l - 1], W[l - 1])))
    r2 = cache["r"]
    grad["lift"] = np.stack([np.sum(_shift(r2, tt // 3 - 1, tt % 3 - 1)[:, :, None] * hb[0], axis=(0, 1))
                             for tt in range(9)])
    flat = np.zeros(lay["nparams"])
    for k, o in lay["off"].items():
        v = grad[k].reshape(-1)
        flat[o: o + v.size] = v
    return flat


def loss_and_grad(p, lay, masks, A, pairs, bf16_round: bool = False, mistake: str | None = None):
    """Summed loss over pairs [(r, e)] (r = A e on the unknowns, A the Jacobi-scaled node
    operator as a scipy matrix) and its gradient."""
    total, grad = 0.0, np.zeros(lay["nparams"])
    mflat = masks[0].reshape(-1)
    for r, e in pairs:
        y, cache = forward(p, lay, masks, r, bf16_round, mistake)
        res = y - e
        ar = A @ res
        ea = float(e @ (A @ e))
        total += float(res @ ar) / ea
        gy = 2.0 * ar / ea * mflat
        grad += backward(lay, masks, cache, gy, bf16_round, mistake)
    return total, grad


def initial_params(lay, seed: int = 1):
    """The library's initial weights (xorshift64 stream, uniform with variance 1 / fan-in)."""
    C = lay["C"]
    st = [(seed * 0x9E3779B97F4A7C15 + 3) & 0xFFFFFFFFFFFFFFFF]

    def rnd():
        s = st[0]
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        st[0] = s
        return float(np.float32((s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0))

    a1 = float(np.sqrt(np.float32(3.0 / 9.0)))
    aC = float(np.sqrt(np.float32(3.0 / (9.0 * C))))
    p = np.zeros(lay["nparams"], np.float32)
    off = lay["off"]
    for k in range(9 * C):
        p[off["lift"] + k] = np.float32(np.float32(a1) * np.float32(rnd()))
    for k in range(9 * C):
        p[off["out"] + k] = np.float32(np.float32(aC) * np.float32(rnd()))
    for l in range(1, lay["L"]):
        for k in range(9 * C * C):
            p[off[f"enc{l}"] + k] = np.float32(np.float32(aC) * np.float32(rnd()))
    for l in range(lay["L"] - 1):
        for k in range(9 * C * C):
            p[off[f"dec{l}"] + k] = np.float32(np.float32(aC) * np.float32(rnd()))
    return p
