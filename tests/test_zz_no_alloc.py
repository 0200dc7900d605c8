"""pic_MY_JPCG_Solver_noalloc (no counterpart in the reference; for callers that cannot
allocate device memory) against pic_MY_JPCG_Solver, the entry point it shares its kernels
with: the same solution, residual, ITER and status, bit for bit, on two calls in a row (the
second on a smaller system, so stale scratch contents of the first would show), and the
size limits refused before any device work."""
import ctypes as C

import numpy as np
import pytest


def say(msg):
    print(msg, flush=True)


def _system(pl, pkg, case, seed):
    w = pkg.workload(case)
    pl.setup(w["domain"], bits=64, solver=0, tol=1e-10)
    rp = pl.setup_array("A_ROWPTR")
    col = pl.setup_array("A_COL")
    val = pl.setup_array("A_VAL")
    diag = pl.setup_array("A_DIAG")
    mat, nz = pkg.sparse_from_csr(rp, col, val)
    n = len(diag)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(n)
    x0 = 0.1 * rng.standard_normal(n)
    return mat, nz, n, diag, b, x0


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["small64", "small24"])
def test_noalloc_matches_jpcg(gpu_lib, pkg, case):
    """Run in this order (64^2, then 24^2): the second call reuses the scratch the first
    filled with a larger system."""
    pl = gpu_lib
    say(f"case: MY_JPCG_Solver_noalloc vs MY_JPCG_Solver, {case}")
    mat, nz, n, diag, b, x0 = _system(pl, pkg, case, 5)
    pkg.check(pkg.lib().pic_set_pcg_params(1e-10, 50000))
    R1, X1 = b.copy(), x0.copy()
    it1 = pl.MY_JPCG_Solver(mat, R1, X1, diag, nz, n)
    R2, X2 = b.copy(), x0.copy()
    it2 = pl.MY_JPCG_Solver_noalloc(mat, R2, X2, diag, nz, n)
    say(f"  N={n} NZ={nz} ITER {it1} / {it2}")
    assert it1 == it2 and it1 > 10
    assert np.array_equal(X1, X2)
    assert np.array_equal(R1, R2)
    # a second call on the same scratch gives the same bits again
    R3, X3 = b.copy(), x0.copy()
    assert pl.MY_JPCG_Solver_noalloc(mat, R3, X3, diag, nz, n) == it1
    assert np.array_equal(X3, X1) and np.array_equal(R3, R1)


@pytest.mark.gpu
def test_noalloc_limits_and_errors(gpu_lib, pkg):
    pl = gpu_lib
    say("case: MY_JPCG_Solver_noalloc limits")
    mat, nz, n, diag, b, x0 = _system(pl, pkg, "small16", 3)
    it = C.c_int()
    big = 262144 + 1
    bb = np.zeros(4)
    # refused on the host before the matrix is read: N beyond the scratch
    rc = pkg.lib().pic_MY_JPCG_Solver_noalloc(mat, bb.ctypes.data_as(C.POINTER(C.c_double)),
                                              bb.ctypes.data_as(C.POINTER(C.c_double)),
                                              bb.ctypes.data_as(C.POINTER(C.c_double)), big + 1, big, C.byref(it))
    assert rc == 4 and "exceed the scratch limits" in pkg.lib().pic_last_error().decode()
    # maxit = 1: IERR = 1 from both entry points
    pkg.check(pkg.lib().pic_set_pcg_params(1e-12, 1))
    for fn in (pl.MY_JPCG_Solver, pl.MY_JPCG_Solver_noalloc):
        with pytest.raises(pkg.PicError) as ei:
            fn(mat, b.copy(), x0.copy(), diag, nz, n)
        assert ei.value.code == 5
    pkg.check(pkg.lib().pic_set_pcg_params(1e-10, 50000))
    # a decreasing row pointer is refused by both
    bad, nzb = pkg.sparse_from_csr(pl.setup_array("A_ROWPTR"), pl.setup_array("A_COL"), pl.setup_array("A_VAL"))
    bad[3].SROW, bad[4].SROW = bad[4].SROW, bad[3].SROW
    for fn in (pl.MY_JPCG_Solver, pl.MY_JPCG_Solver_noalloc):
        with pytest.raises(pkg.PicError) as ei:
            fn(bad, b.copy(), x0.copy(), diag, nzb, n)
        assert ei.value.code == 1
