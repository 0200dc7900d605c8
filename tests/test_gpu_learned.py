"""GPU checks of the learned preconditioner: network forward and gradient against the
numpy oracle (BF16 emulated; stated bounds 1e-2 relative for the output and 5e-2 for each
gradient tensor), training on pairs harvested from the plain solver, the learned FCG
against the plain CG at the same 1e-8 residual, and the fallback on a useless network."""
import numpy as np
import pytest

import learned as lrn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def smooth_rho(dom, rng, nblobs=4):
    x = np.arange(dom.nnx)[:, None] * np.ones((1, dom.nny))
    y = np.ones((dom.nnx, 1)) * np.arange(dom.nny)[None, :]
    rho = np.zeros((dom.nnx, dom.nny))
    for _ in range(nblobs):
        cx, cy = rng.uniform(0, dom.nnx - 1), rng.uniform(0, dom.nny - 1)
        w = rng.uniform(2.0, 0.25 * dom.nnx)
        rho += rng.normal() * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * w * w))
    rho += 0.05 * rng.normal(size=rho.shape)
    return rho.reshape(-1)


@pytest.fixture(scope="module")
def net(gpu_lib, pkg):
    say("case: learned network setup small64")
    w = pkg.workload("small64")
    dom = w["domain"]
    pl = gpu_lib
    pl.setup(dom, bits=32, solver=1, tol=1e-8)
    pl.learn_setup(10, (16, 16, 32), seed=5)
    lay = pl.learn_layout()
    A, sinv = lrn.scaled_operator(pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL"),
                                  pl.setup_array("A_DIAG"), pl.setup_array("UNKNOWN_NODE"), dom.nnx * dom.nny)
    return {"dom": dom, "lay": lay, "A": A, "mask": (sinv > 0).astype(np.float64)}


def perturbed_params(pl, lay, rng, scale=0.2):
    p = pl.learn_params().astype(np.float64)
    p += scale * rng.standard_normal(p.size)
    w = p[lay["off_in"]: lay["off_in"] + 9 * lay["C"][0] * 16].reshape(9, lay["C"][0], 16)
    w[:, :, 1:] = 0.0
    return p.astype(np.float32)


def test_layout_matches_oracle(gpu_lib, net):
    dom, lay = net["dom"], net["lay"]
    ref = lrn.layout(dom.nnx, dom.nny, [16, 16, 32])
    for k in ("levels", "H", "W", "C", "off_in", "off_q", "nparams"):
        assert lay[k] == ref[k], k
    assert lay["off_down"][1:] == ref["off_down"][1:]


def test_forward_matches_oracle(gpu_lib, net):
    pl, lay, mask = gpu_lib, net["lay"], net["mask"]
    say("case: learned forward small64")
    rng = np.random.default_rng(7)
    p = perturbed_params(pl, lay, rng)
    pl.learn_set_params(p)
    for trial in range(2):
        r = (rng.standard_normal(mask.size) * mask).astype(np.float32)
        y_gpu = pl.learn_apply(r).astype(np.float64)
        y_or, _, _ = lrn.forward(p.astype(np.float64), lay, r.astype(np.float64), mask, emulate_bf16=True)
        err = np.linalg.norm(y_gpu - y_or) / np.linalg.norm(y_or)
        say(f"  forward trial {trial}: rel err {err:.2e}")
        assert err < 1e-2
        assert np.all(y_gpu[mask == 0] == 0)


def test_gradient_matches_oracle(gpu_lib, net):
    pl, lay, mask, A = gpu_lib, net["lay"], net["mask"], net["A"]
    say("case: learned gradient small64")
    rng = np.random.default_rng(8)
    p = perturbed_params(pl, lay, rng, 0.1)
    pl.learn_set_params(p)
    es = [(rng.standard_normal(mask.size) * mask) for _ in range(2)]
    rs = [(A @ e).astype(np.float32) for e in es]
    es32 = [e.astype(np.float32) for e in es]
    pl.learn_set_pairs(np.stack(rs), np.stack(es32))
    loss_gpu, g_gpu = pl.learn_gradient([0, 1], lay["nparams"])
    pairs = [(r.astype(np.float64), e.astype(np.float64)) for r, e in zip(rs, es32)]
    loss_or, g_or = lrn.loss_and_grad(p.astype(np.float64), lay, pairs, A, mask, emulate_bf16=True)
    say(f"  loss gpu {loss_gpu:.6e} oracle {loss_or:.6e}")
    assert abs(loss_gpu - loss_or) <= 1e-2 * abs(loss_or)
    segs = [("In", lay["off_in"], 9 * lay["C"][0] * 16)]
    segs += [(f"Down{l}", lay["off_down"][l], 9 * lay["C"][l] * lay["C"][l - 1]) for l in range(1, lay["levels"])]
    segs += [(f"Q{l}", lay["off_q"][l], lay["C"][l] ** 2) for l in range(lay["levels"])]
    for name, o, n in segs:
        a, b = g_gpu[o:o + n].astype(np.float64), g_or[o:o + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        say(f"  grad {name}: |g| {np.linalg.norm(b):.3e} rel err {err:.2e}")
        assert err < 5e-2, name


def test_training_and_learned_solve(gpu_lib, net):
    pl, dom = gpu_lib, net["dom"]
    say("case: harvest + train + learned FCG small64")
    rng = np.random.default_rng(9)
    pl.learn_setup(10, (16, 16, 32), seed=5)
    pl.learn_harvest(True, 256, 1)
    phi0 = pl.setup_array("DIRICHLET_PHI").copy()
    for k in range(8):
        phi = phi0.copy()
        pl.IFE_Solve_2D(Rho=smooth_rho(dom, rng), Phi=phi)
    pl.learn_harvest(False, 256, 1)
    info = pl.learn_info()
    say(f"  harvested {info['dataset']} pairs")
    assert info["dataset"] >= 16
    l0 = pl.learn_eval(info["dataset"])
    lf, ll = pl.learn_train(300, batch=4, lr=2e-3, seed=3)
    l1 = pl.learn_eval(info["dataset"])
    say(f"  loss before {l0:.4f} after {l1:.4f} (first batch {lf:.4f}, last {ll:.4f})")
    assert l1 < 0.8 * l0
    # held-out RHS: plain vs learned at the same 1e-8 residual, same warm start
    rho = smooth_rho(dom, np.random.default_rng(1234))
    phi_p = phi0.copy()
    pl.set_solver(1)
    pl.IFE_Solve_2D(Rho=rho, Phi=phi_p)
    plain = pl.last_solve()
    phi_l = phi0.copy()
    pl.set_solver(2)
    pl.IFE_Solve_2D(Rho=rho, Phi=phi_l)
    learned = pl.last_solve()
    li = pl.learn_info()
    pl.set_solver(1)
    say(f"  plain {plain}  learned {learned}  fcg {li['last_fcg_iterations']} fallbacks {li['fallbacks']}")
    assert plain["relres"] <= 1e-8 and learned["relres"] <= 1e-8
    assert li["fallbacks"] == 0
    assert 3 * learned["iterations"] <= plain["iterations"]
    assert np.linalg.norm(phi_l - phi_p) / np.linalg.norm(phi_p) < 1e-6


def test_fallback_to_plain_cg(gpu_lib, net):
    pl, dom = gpu_lib, net["dom"]
    say("case: learned FCG fallback small64")
    pl.learn_setup(10, (16, 16, 32), seed=5)
    pl.learn_set_params(np.zeros(net["lay"]["nparams"], np.float32))     # M = 0: FCG breaks down at once
    fb0 = pl.learn_info()["fallbacks"]
    rho = smooth_rho(dom, np.random.default_rng(99))
    phi = pl.setup_array("DIRICHLET_PHI").copy()
    pl.set_solver(2)
    pl.IFE_Solve_2D(Rho=rho, Phi=phi)
    info = pl.last_solve()
    pl.set_solver(1)
    assert info["relres"] <= 1e-8
    assert pl.learn_info()["fallbacks"] == fb0 + 1
