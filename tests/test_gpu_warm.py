"""GPU checks of the learned warm start (the encoder-decoder on BF16 matrix cores): with
sparse +-1 weights and integer inputs every sum is exact, and the output equals the oracle's
(which rounds to bfloat16 where the library does) bit for bit; with the random initial
weights the output is within 1e-2 of the oracle in BF16 and in FP64 (the random network
amplifies a single BF16 rounding flip, which the FP32 summation order decides, to ~5e-5 of
its output); the training gradient per weight tensor within 1e-2 on levels 0-1, 3e-2 on
level 2 and 1e-1 below (the coarse tensors' gradients are 1e-3..1e-5 of the fine ones and
sums with cancellation: in the oracle itself nine BF16 rounding flips move the level-3 and
level-4 gradients by 2-3%, a wrong tap, channel or transfer moves them by O(1));
bit-identical gradients on repeated calls,
and the gain on cold solves of held-out charge fields: trained on pairs harvested from plain
cold solves of one run, the warm start makes the learned FCG need fewer iterations on the
charge of an independent run, at the same 1e-8 residual."""
import numpy as np
import pytest

import learned as lrn
import warm as wm

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def wnet(gpu_lib, pkg):
    say("case: warm-start setup small64")
    dom = pkg.workload("small64")["domain"]
    pl = gpu_lib
    pl.setup(dom, bits=32, solver=1, tol=1e-8)
    pl.learn_setup(12, [16], seed=5)
    A, sinv = lrn.scaled_operator(pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL"),
                                  pl.setup_array("A_DIAG"), pl.setup_array("UNKNOWN_NODE"), dom.nnx * dom.nny)
    return {"dom": dom, "A": A, "mask": (sinv > 0).astype(np.float64)}


def _oracle_net(wnet, levels, C):
    dom = wnet["dom"]
    H, W = wm.level_sizes(dom.nnx, dom.nny, levels)
    return wm.layout(H, W, C), wm.level_masks(wnet["mask"], H, W)


@pytest.mark.parametrize("C", [16, 32])
def test_forward_matches_oracle(gpu_lib, wnet, C):
    pl, mask = gpu_lib, wnet["mask"]
    say(f"case: warm-start forward small64 C={C}")
    pl.warm_setup(5, C, seed=3)
    lay, ms = _oracle_net(wnet, 5, C)
    lg = pl.warm_layout()
    assert lg["levels"] == lay["L"] and lg["H"] == lay["H"] and lg["W"] == lay["W"] and lg["C"] == C
    p32 = pl.warm_params()
    assert p32.size == lay["nparams"]
    assert np.allclose(p32, wm.initial_params(lay, 3), rtol=1e-6, atol=0)
    p = p32.astype(np.float64)
    rng = np.random.default_rng(21)
    for trial in range(2):
        r = (rng.standard_normal(mask.size) * mask).astype(np.float32)
        y = pl.warm_apply(r).astype(np.float64)
        yb = wm.forward(p, lay, ms, r.astype(np.float64), bf16_round=True)[0]
        y64 = wm.forward(p, lay, ms, r.astype(np.float64))[0]
        eb, e64 = rel(y, yb), rel(y, y64)
        say(f"  trial {trial}: rel err vs bf16 oracle {eb:.2e}, vs FP64 oracle {e64:.2e}")
        assert eb < 1e-2
        assert e64 < 1e-2
        assert np.all(y[mask == 0] == 0.0)


def sparse_unit_params(lay, rng):
    """One +-1 weight per output channel of every convolution (three in the output one): every
    activation is a short sum of dyadic numbers, exact in FP32 on both sides."""
    C = lay["C"]
    p = np.zeros(lay["nparams"], np.float32)
    for name, o in lay["off"].items():
        if name == "lift":
            for c in range(C):
                p[o + rng.integers(9) * C + c] = rng.choice([-1.0, 1.0])
        elif name == "out":
            for _ in range(3):
                p[o + rng.integers(9) * C + rng.integers(C)] = rng.choice([-1.0, 1.0])
        else:
            for co in range(C):
                p[o + (rng.integers(9) * C + rng.integers(C)) * C + co] = rng.choice([-1.0, 1.0])
    return p


@pytest.mark.parametrize("C", [16, 32])
def test_forward_exact_case_is_bit_identical(gpu_lib, wnet, C):
    pl, mask = gpu_lib, wnet["mask"]
    say(f"case: warm-start forward, exact arithmetic, small64 C={C}")
    pl.warm_setup(5, C, seed=3)
    lay, ms = _oracle_net(wnet, 5, C)
    rng = np.random.default_rng(31)
    for trial in range(3):
        p = sparse_unit_params(lay, rng)
        pl.warm_set_params(p)
        r = (rng.integers(-2, 3, mask.size) * mask).astype(np.float32)
        y = pl.warm_apply(r).astype(np.float64)
        yb = wm.forward(p.astype(np.float64), lay, ms, r.astype(np.float64), bf16_round=True)[0]
        nd = int(np.sum(y != yb))
        say(f"  trial {trial}: |y| {np.linalg.norm(yb):.3e}, nodes differing {nd}")
        assert np.linalg.norm(yb) > 0 and nd == 0


@pytest.mark.parametrize("C", [16, 32])
def test_gradient_matches_oracle(gpu_lib, wnet, C):
    pl, mask, A = gpu_lib, wnet["mask"], wnet["A"]
    say(f"case: warm-start gradient small64 C={C}")
    pl.warm_setup(5, C, seed=4)
    lay, ms = _oracle_net(wnet, 5, C)
    p = pl.warm_params().astype(np.float64)
    rng = np.random.default_rng(22)
    es = [(rng.standard_normal(mask.size) * mask).astype(np.float32) for _ in range(2)]
    rs = [(A @ e.astype(np.float64)).astype(np.float32) for e in es]
    pl.learn_set_pairs(np.stack(rs), np.stack(es))
    loss_g, g_g = pl.warm_gradient([0, 1], lay["nparams"])
    loss_g2, g_g2 = pl.warm_gradient([0, 1], lay["nparams"])
    assert loss_g == loss_g2 and np.array_equal(g_g, g_g2), "gradient not reproducible"
    pairs = [(r.astype(np.float64), e.astype(np.float64)) for r, e in zip(rs, es)]
    loss_o, g_o = wm.loss_and_grad(p, lay, ms, A, pairs, bf16_round=True)
    say(f"  loss gpu {loss_g:.6e} oracle {loss_o:.6e}")
    assert abs(loss_g - loss_o) <= 1e-3 * abs(loss_o)
    C2 = C * C
    for name, o in lay["off"].items():
        n = 9 * C if name in ("lift", "out") else 9 * C2
        a, b = g_g[o:o + n].astype(np.float64), g_o[o:o + n]
        err = rel(a, b)
        level = 0 if name in ("lift", "out") else int(name[3:])
        tol = 1e-2 if level <= 1 else (3e-2 if level == 2 else 1e-1)
        say(f"  grad {name}: |g| {np.linalg.norm(b):.3e} rel err {err:.2e} (bound {tol:.0e})")
        assert err < tol, name


def test_warm_start_cuts_cold_solve_iterations(gpu_lib, pkg):
    """Three-conductor geometry at 128^2: plain cold solves of the charge of 48 steps harvest
    the initial-residual pairs; held-out: the charge of 12 steps of an independent run."""
    pl = gpu_lib
    say("case: warm start, cold solves, cfg3s128")
    pkg.build_workload(pl, "cfg3s128", seed=77)
    pl.initial_field()
    held = []
    for _ in range(12):
        pl.step(1)
        held.append(pl.field("RHO"))
    pkg.build_workload(pl, "cfg3s128", seed=2024)
    pl.initial_field()
    pl.learn_setup(12, [16, 16], seed=7)
    phi0 = pl.setup_array("DIRICHLET_PHI").copy()
    cap = 512
    for _ in range(48):
        pl.learn_harvest(False, cap, 1000)
        pl.step(1)
        rho = pl.field("RHO")
        pl.learn_harvest(True, cap, 1000)
        phi = phi0.copy()
        pl.IFE_Solve_2D(Rho=rho, Phi=phi)
    pl.learn_harvest(False, cap, 1000)
    pl.warm_setup(6, 16, seed=3)
    ninit = pl.warm_info()["initial_pairs"]
    say(f"  initial-residual pairs {ninit} of {pl.learn_info()['dataset']}")
    assert ninit >= 40
    l0, l0s = pl.warm_eval(ninit)
    lf, ll = pl.warm_train(600, batch=8, lr=2e-3, seed=5)
    l1, l1s = pl.warm_eval(ninit)
    say(f"  loss raw {l0:.4e} -> {l1:.4e}, after line search {l0s:.4e} -> {l1s:.4e} (batches {lf:.4e} .. {ll:.4e})")
    assert l1 < l0 and l1s < 0.5

    def solve_all(warm):
        pl.warm_enable(warm)
        pl.set_solver(2)
        its, phis = [], []
        for rho in held:
            phi = phi0.copy()
            pl.IFE_Solve_2D(Rho=rho, Phi=phi)
            info = pl.last_solve()
            assert info["relres"] <= 1e-8
            its.append(info["iterations"])
            phis.append(phi)
        pl.set_solver(1)
        pl.warm_enable(False)
        return np.array(its), phis

    it0, ph0 = solve_all(False)
    it1, ph1 = solve_all(1)
    info = pl.warm_info()
    say(f"  FCG iterations without {it0.tolist()}\n                 with    {it1.tolist()}; last alpha {info['last_alpha']:.4f}")
    assert info["applied"] >= len(held)
    assert it1.sum() < it0.sum()
    for a, b in zip(ph0, ph1):
        assert rel(b, a) < 1e-6
    # the captured graph (replayed once per solve, the solve's own launches and host waits in
    # between) gives the same bits as the kernels launched one by one
    it2, ph2 = solve_all(2)
    assert np.array_equal(it1, it2)
    for a, b in zip(ph1, ph2):
        assert np.array_equal(a, b)
