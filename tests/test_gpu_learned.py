"""GPU checks of the learned multigrid preconditioner: Galerkin hierarchy and network output
(FP32 V-cycle with the collapsed 5x5 smoother) against the FP64 numpy oracle (stated bound
1e-4 relative), the training gradient (FP32 V-cycles, tap-pair Gram matrices accumulated in
FP32 per thread and FP64 across blocks) against the oracle's factored FP64 gradient (2e-3
relative per tensor), training on pairs harvested from the plain
solver, the learned FCG against the plain CG at the same 1e-8 residual, and the fallback to
the plain CG when the network is unusable."""
import numpy as np
import pytest

import learned as lrn

pytestmark = pytest.mark.gpu
CHANNELS = [16, 32]


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
    pl.learn_setup(12, CHANNELS, seed=5)
    lay = pl.learn_layout()
    A, sinv = lrn.scaled_operator(pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL"),
                                  pl.setup_array("A_DIAG"), pl.setup_array("UNKNOWN_NODE"), dom.nnx * dom.nny)
    levels = lrn.build_hierarchy(A, dom.nnx, dom.nny)
    return {"dom": dom, "lay": lay, "A": A, "mask": (sinv > 0).astype(np.float64), "levels": levels}


def perturbed_params(pl, lay, rng, scale):
    p = pl.learn_params().astype(np.float64)
    p += scale * rng.standard_normal(p.size)
    return p.astype(np.float32)


def test_layout_matches_oracle(gpu_lib, net):
    dom, lay = net["dom"], net["lay"]
    ref = lrn.layout(dom.nnx, dom.nny, CHANNELS, learned_levels=len(CHANNELS))
    for k in ("levels", "learned", "H", "W", "C", "off_in", "off_q", "nparams"):
        assert lay[k] == ref[k], k


def test_untrained_network_is_the_jacobi_vcycle(gpu_lib, net):
    pl, lay, mask = gpu_lib, net["lay"], net["mask"]
    say("case: learned untrained = Jacobi V-cycle small64")
    p0 = lrn.initial_params(lay, small=0.0).astype(np.float32)
    pl.learn_set_params(p0)
    rng = np.random.default_rng(6)
    r = (rng.standard_normal(mask.size) * mask).astype(np.float32)
    y_gpu = pl.learn_apply(r).astype(np.float64)
    dom = net["dom"]
    lay0 = lrn.layout(dom.nnx, dom.nny, CHANNELS, learned_levels=0)
    y_or, _ = lrn.forward(np.zeros(0), lay0, net["levels"], r.astype(np.float64))
    err = np.linalg.norm(y_gpu - y_or) / np.linalg.norm(y_or)
    say(f"  rel err vs classical V-cycle {err:.2e}")
    assert err < 1e-4
    assert np.all(y_gpu[mask == 0] == 0)


def test_forward_matches_oracle(gpu_lib, net):
    pl, lay, mask = gpu_lib, net["lay"], net["mask"]
    say("case: learned forward small64")
    rng = np.random.default_rng(7)
    p = perturbed_params(pl, lay, rng, 0.05)
    pl.learn_set_params(p)
    for trial in range(2):
        r = (rng.standard_normal(mask.size) * mask).astype(np.float32)
        y_gpu = pl.learn_apply(r).astype(np.float64)
        y_or, _ = lrn.forward(p.astype(np.float64), lay, net["levels"], r.astype(np.float64))
        err = np.linalg.norm(y_gpu - y_or) / np.linalg.norm(y_or)
        say(f"  forward trial {trial}: rel err {err:.2e}")
        assert err < 1e-4


def test_gradient_matches_oracle(gpu_lib, net):
    pl, lay, mask, A = gpu_lib, net["lay"], net["mask"], net["A"]
    say("case: learned gradient small64")
    rng = np.random.default_rng(8)
    p = perturbed_params(pl, lay, rng, 0.05)
    pl.learn_set_params(p)
    es = [(rng.standard_normal(mask.size) * mask) for _ in range(2)]
    rs = [(A @ e).astype(np.float32) for e in es]
    es32 = [e.astype(np.float32) for e in es]
    pl.learn_set_pairs(np.stack(rs), np.stack(es32))
    loss_gpu, g_gpu = pl.learn_gradient([0, 1], lay["nparams"])
    pairs = [(r.astype(np.float64), e.astype(np.float64)) for r, e in zip(rs, es32)]
    loss_or, g_or = lrn.loss_and_grad(p.astype(np.float64), lay, net["levels"], pairs)
    say(f"  loss gpu {loss_gpu:.6e} oracle {loss_or:.6e}")
    assert abs(loss_gpu - loss_or) <= 1e-3 * abs(loss_or)
    segs = [(f"In{l}", lay["off_in"][l], 9 * lay["C"][l]) for l in range(lay["learned"])]
    segs += [(f"Q{l}", lay["off_q"][l], lay["C"][l] ** 2) for l in range(lay["learned"])]
    for name, o, n in segs:
        a, b = g_gpu[o:o + n].astype(np.float64), g_or[o:o + n]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        say(f"  grad {name}: |g| {np.linalg.norm(b):.3e} rel err {err:.2e}")
        assert err < 2e-3, name


def test_training_and_learned_solve(gpu_lib, net):
    pl, dom = gpu_lib, net["dom"]
    say("case: harvest + train + learned FCG small64")
    rng = np.random.default_rng(9)
    pl.learn_setup(12, CHANNELS, seed=5)
    pl.learn_harvest(True, 256, 1)
    phi0 = pl.setup_array("DIRICHLET_PHI").copy()
    for k in range(8):
        phi = phi0.copy()
        pl.IFE_Solve_2D(Rho=smooth_rho(dom, rng), Phi=phi)
    pl.learn_harvest(False, 256, 1)
    info = pl.learn_info()
    say(f"  harvested {info['dataset']} pairs")
    assert info["dataset"] >= 16
    rho = smooth_rho(dom, np.random.default_rng(1234))
    # held-out RHS: plain, untrained and trained network at the same 1e-8 residual, same warm start
    res = {}
    for name, mode in (("plain", 1), ("untrained", 2)):
        pl.set_solver(mode)
        phi = phi0.copy()
        pl.IFE_Solve_2D(Rho=rho, Phi=phi)
        res[name] = (pl.last_solve(), phi)
    l0 = pl.learn_eval(info["dataset"])
    lf, ll = pl.learn_train(200, batch=4, lr=2e-3, seed=3)
    l1 = pl.learn_eval(info["dataset"])
    say(f"  loss before {l0:.4e} after {l1:.4e} (first batch {lf:.4e}, last {ll:.4e})")
    assert l1 < l0
    pl.set_solver(2)
    phi = phi0.copy()
    pl.IFE_Solve_2D(Rho=rho, Phi=phi)
    res["trained"] = (pl.last_solve(), phi)
    li = pl.learn_info()
    pl.set_solver(1)
    for k, (info_k, _) in res.items():
        say(f"  {k}: {info_k}")
    assert all(v[0]["relres"] <= 1e-8 for v in res.values())
    assert li["fallbacks"] == 0
    assert 3 * res["trained"][0]["iterations"] <= res["plain"][0]["iterations"]
    assert res["trained"][0]["iterations"] <= res["untrained"][0]["iterations"]
    phi_p = res["plain"][1]
    assert np.linalg.norm(res["trained"][1] - phi_p) / np.linalg.norm(phi_p) < 1e-6


def test_fallback_to_plain_cg(gpu_lib, net):
    pl, dom = gpu_lib, net["dom"]
    say("case: learned FCG fallback small64")
    pl.learn_setup(12, CHANNELS, seed=5)
    # a network that produces NaN: the FCG detects the breakdown and the plain CG finishes
    pl.learn_set_params(np.full(net["lay"]["nparams"], np.nan, np.float32))
    fb0 = pl.learn_info()["fallbacks"]
    rho = smooth_rho(dom, np.random.default_rng(99))
    phi = pl.setup_array("DIRICHLET_PHI").copy()
    pl.set_solver(2)
    pl.IFE_Solve_2D(Rho=rho, Phi=phi)
    info = pl.last_solve()
    pl.set_solver(1)
    assert info["relres"] <= 1e-8
    assert pl.learn_info()["fallbacks"] == fb0 + 1
    assert np.all(np.isfinite(phi))


def test_forward_one_learned_level_matches_oracle(gpu_lib, pkg):
    """129 x 129 nodes, one learned level: the Jacobi level 65 x 65 by separate launches, 33 x 33
    the dense coarsest level, against the FP64 oracle at the same 1e-4 bound."""
    say("case: learned forward small128, one learned level")
    dom = pkg.workload("small128")["domain"]
    pl = gpu_lib
    pl.setup(dom, bits=32, solver=1, tol=1e-8)
    pl.learn_setup(12, [16], seed=5)
    lay = pl.learn_layout()
    assert lay["H"] == [129, 65, 33]
    A, sinv = lrn.scaled_operator(pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL"),
                                  pl.setup_array("A_DIAG"), pl.setup_array("UNKNOWN_NODE"), dom.nnx * dom.nny)
    levels = lrn.build_hierarchy(A, dom.nnx, dom.nny, learned_levels=1)
    mask = (sinv > 0).astype(np.float64)
    rng = np.random.default_rng(17)
    p = perturbed_params(pl, lay, rng, 0.05)
    pl.learn_set_params(p)
    r = (rng.standard_normal(mask.size) * mask).astype(np.float32)
    y_gpu = pl.learn_apply(r).astype(np.float64)
    y_or, _ = lrn.forward(p.astype(np.float64), lay, levels, r.astype(np.float64))
    err = np.linalg.norm(y_gpu - y_or) / np.linalg.norm(y_or)
    say(f"  rel err {err:.2e}")
    assert err < 1e-4


def test_trained_gain_on_heldout_charge_fields(gpu_lib, pkg):
    """The stand-alone Poisson protocol at 256^2 (the three-conductor geometry): charge fields
    of 24 steps of an independent run, solved cold to 1e-8 with the untrained (damped-Jacobi)
    and the trained V-cycle (2 learned levels, minibatches of 8 with a quarter harvested
    pairs). The trained network needs strictly fewer FCG iterations in total and more on no
    field."""
    say("case: held-out charge fields, cfg3s256, trained vs untrained")
    pl = gpu_lib
    pkg.build_workload(pl, "cfg3s256", seed=77)
    pl.initial_field()
    held = []
    for _ in range(24):
        pl.step(1)
        held.append(pl.field("RHO"))
    pkg.build_workload(pl, "cfg3s256", seed=2024)
    pl.initial_field()
    pl.learn_setup(12, CHANNELS[:1] * 2, seed=7)
    p_untrained = pl.learn_params().copy()
    pl.learn_harvest(True, 1024, 1)
    for _ in range(60):
        pl.step(1)
        if pl.learn_info()["dataset"] >= 1024:
            break
    pl.learn_harvest(False, 1024, 1)
    phi0 = pl.setup_array("DIRICHLET_PHI").copy()

    def iterations():
        pl.set_solver(2)
        its = []
        for rho in held:
            phi = phi0.copy()
            pl.IFE_Solve_2D(Rho=rho, Phi=phi)
            info = pl.last_solve()
            assert info["relres"] <= 1e-8
            its.append(info["iterations"])
        pl.set_solver(1)
        return np.array(its)

    it_un = iterations()
    pl.learn_data_mix(0.25)
    lf, ll = pl.learn_train(3000, batch=8, lr=5e-3, seed=11)
    it_tr = iterations()
    pl.learn_set_params(p_untrained)
    say(f"  pairs {pl.learn_info()['dataset']}, loss first {lf:.4e} last {ll:.4e}")
    say(f"  untrained {it_un.tolist()}\n  trained   {it_tr.tolist()}")
    assert it_tr.sum() < it_un.sum()
    assert np.all(it_tr <= it_un)
