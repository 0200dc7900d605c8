"""CPU checks of the warm-start oracle (oracle/warm.py): the transfer operators and the
convolution adjoint, linearity of the network, its gradient against central differences of
the FP64 loss (with planted mistakes that the check catches), the bfloat16 rounding rule, and
the level sizes and masks."""
import numpy as np
import pytest

import warm as wm
from test_learned_cpu import q1_operator


def test_bf16_rounds_to_nearest_even():
    one = 1.0
    assert wm.bf16(one + 2.0 ** -8) == one                      # tie, even mantissa below
    assert wm.bf16(one + 3 * 2.0 ** -8) == one + 2.0 ** -6      # tie, odd below -> up
    assert wm.bf16(one + 2.0 ** -8 + 2.0 ** -20) == one + 2.0 ** -7
    assert wm.bf16(-3.0) == -3.0
    x = np.random.default_rng(0).standard_normal(1000)
    r = wm.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * 1.0000001)


def test_level_sizes_and_masks():
    assert wm.level_sizes(513, 513)[0] == [513, 257, 129, 65, 33, 17]
    assert wm.level_sizes(1025, 1025)[0] == [1025, 513, 257, 129, 65, 33]
    assert wm.level_sizes(257, 129, 8) == ([257, 129, 65, 33, 17, 9], [129, 65, 33, 17, 9, 5])
    assert wm.level_sizes(17, 17, 8)[0] == [17, 9, 5]
    m0 = np.zeros((9, 9)); m0[1:-1, 1:-1] = 1.0; m0[4, 4] = 0.0
    ms = wm.level_masks(m0.reshape(-1), [9, 5, 3], [9, 5, 3])
    assert ms[1][2, 2] == 0.0 and ms[1][1, 1] == 1.0 and ms[1][0, 0] == 0.0
    assert ms[2].shape == (3, 3)


@pytest.mark.parametrize("Hf,Wf", [(17, 17), (16, 11)])
def test_interpolation_is_four_times_restriction_adjoint(Hf, Wf):
    rng = np.random.default_rng(1)
    Hc, Wc = (Hf + 1) // 2, (Wf + 1) // 2
    xc = rng.standard_normal((Hc, Wc, 3))
    yf = rng.standard_normal((Hf, Wf, 3))
    lhs = np.sum(wm.interp(xc, Hf, Wf) * yf)
    rhs = 4.0 * np.sum(xc * wm.restrict(yf, Hc, Wc))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # interpolation reproduces bilinear functions away from the far edge, restriction of a
    # constant is the constant in the interior
    I, J = np.meshgrid(np.arange(Hc), np.arange(Wc), indexing="ij")
    lin = (2.0 * I + 3.0 * J + 1.0)[:, :, None]
    i, j = np.meshgrid(np.arange(Hf), np.arange(Wf), indexing="ij")
    f = wm.interp(lin, Hf, Wf)[:, :, 0]
    ref = 1.0 * i + 1.5 * j + 1.0
    inner = (i < 2 * (Hc - 1)) & (j < 2 * (Wc - 1))
    assert np.allclose(f[inner], ref[inner])
    rc = wm.restrict(np.ones((Hf, Wf, 1)), Hc, Wc)[1:-1, 1:-1]
    assert np.allclose(rc, 1.0)


def test_convolution_adjoint_and_taps():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((9, 7, 16))
    z = rng.standard_normal((9, 7, 8))
    w = rng.standard_normal((9, 16, 8))
    assert abs(np.sum(wm.conv(x, w) * z) - np.sum(x * wm.conv_adjoint(z, w))) <= 1e-11 * np.sum(np.abs(x))
    # tap t reads the node at offset (t // 3 - 1, t % 3 - 1)
    d = np.zeros((5, 5, 1)); d[2, 2, 0] = 1.0
    for t in range(9):
        wt = np.zeros((9, 1, 1)); wt[t, 0, 0] = 1.0
        y = wm.conv(d, wt)[:, :, 0]
        assert y[2 - (t // 3 - 1), 2 - (t % 3 - 1)] == 1.0 and y.sum() == 1.0


def _setup(n=17, C=16, levels=3, seed=3):
    A, mask = q1_operator(n)
    H, W = wm.level_sizes(n, n, levels)
    lay = wm.layout(H, W, C)
    ms = wm.level_masks(mask, H, W)
    rng = np.random.default_rng(seed)
    p = wm.initial_params(lay, seed=5).astype(np.float64)
    pairs = []
    for _ in range(2):
        e = rng.standard_normal(n * n) * mask
        pairs.append((A @ e, e))
    return A, mask, lay, ms, p, pairs


def test_network_is_linear_and_masked():
    A, mask, lay, ms, p, pairs = _setup()
    r1, r2 = pairs[0][0], pairs[1][0]
    y1 = wm.forward(p, lay, ms, r1)[0]
    y2 = wm.forward(p, lay, ms, r2)[0]
    y12 = wm.forward(p, lay, ms, 2.0 * r1 - 0.5 * r2)[0]
    assert np.allclose(y12, 2.0 * y1 - 0.5 * y2, rtol=1e-12, atol=1e-12 * np.abs(y1).max())
    assert np.all(y1[mask == 0] == 0.0)


def _fd_check(mistake=None):
    A, mask, lay, ms, p, pairs = _setup()
    loss, grad = wm.loss_and_grad(p, lay, ms, A, pairs, mistake=mistake)
    rng = np.random.default_rng(4)
    worst = 0.0
    for name, o in lay["off"].items():
        n = 9 * lay["C"] if name in ("lift", "out") else 9 * lay["C"] ** 2
        for k in rng.choice(n, 3, replace=False):
            d = np.zeros_like(p); d[o + k] = 1.0
            eps = 1e-5
            lp = wm.loss_and_grad(p + eps * d, lay, ms, A, pairs)[0]
            lm = wm.loss_and_grad(p - eps * d, lay, ms, A, pairs)[0]
            fd = (lp - lm) / (2 * eps)
            worst = max(worst, abs(fd - grad[o + k]) / (abs(fd) + 1e-3 * np.abs(grad).max()))
    return worst


def test_gradient_matches_central_differences():
    assert _fd_check() < 1e-6


@pytest.mark.parametrize("mistake", ["no_flip"])
def test_gradient_check_catches_planted_mistake(mistake):
    assert _fd_check(mistake) > 1e-2


def test_planted_forward_mistake_changes_output():
    A, mask, lay, ms, p, pairs = _setup()
    y = wm.forward(p, lay, ms, pairs[0][0])[0]
    y_bad = wm.forward(p, lay, ms, pairs[0][0], mistake="no_skip")[0]
    assert np.linalg.norm(y - y_bad) > 1e-2 * np.linalg.norm(y)


def test_bf16_emulation_stays_close_to_fp64():
    A, mask, lay, ms, p, pairs = _setup()
    y = wm.forward(p, lay, ms, pairs[0][0])[0]
    yb = wm.forward(p, lay, ms, pairs[0][0], bf16_round=True)[0]
    err = np.linalg.norm(yb - y) / np.linalg.norm(y)
    assert 0.0 < err < 3e-2
