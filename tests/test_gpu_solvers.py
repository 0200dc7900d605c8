"""GPU checks of the comparison solver (solver_mode 3: FP64 PCG with a multicolour IC(0)
preconditioner) against the reference-semantics FP64 Jacobi-PCG (solver_mode 0) on the same
RHS and warm start: both reach the relative residual, the potentials agree within what that
residual allows, IC(0) needs fewer iterations, and the solve time is reported."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", ["small64", "cfg3s128", "dielrz64"])
def test_ic0_pcg_matches_jacobi_pcg(gpu_lib, pkg, name):
    say(f"case: IC(0)-PCG vs Jacobi-PCG, {name}")
    pl = gpu_lib
    w = pkg.workload(name)
    dom = w["domain"]
    pl.setup(dom, bits=64, solver=0, tol=1e-10)
    rng = np.random.default_rng(3)
    nn = dom.nnx * dom.nny
    x = np.arange(dom.nnx)[:, None] / (dom.nnx - 1)
    y = np.arange(dom.nny)[None, :] / (dom.nny - 1)
    rho = (np.sin(3 * np.pi * x) * np.cos(2 * np.pi * y)).reshape(-1) + 0.1 * rng.standard_normal(nn)
    phi0 = pl.setup_array("DIRICHLET_PHI").copy()
    out = {}
    for mode in (0, 3, 3):                   # the second IC solve reuses the cached factor
        pl.set_solver(mode)
        phi = phi0.copy()
        pl.IFE_Solve_2D(Rho=rho, Phi=phi)
        out.setdefault(mode, []).append((pl.last_solve(), pl.last_solve_ms(), phi))
    pl.set_solver(0)
    (i0, ms0, p0), = out[0]
    (i3, ms3, p3), (i3b, _, p3b) = out[3]
    say(f"  jacobi {i0['iterations']} it {ms0:.2f} ms, ic0 {i3['iterations']} it {ms3:.2f} ms")
    assert i0["relres"] <= 1e-10 and i3["relres"] <= 1e-10
    assert np.linalg.norm(p3 - p0) / np.linalg.norm(p0) < 1e-8
    assert np.array_equal(p3, p3b) and i3["iterations"] == i3b["iterations"]
    assert i3["iterations"] < 0.9 * i0["iterations"]
    assert ms0 > 0 and ms3 > 0
