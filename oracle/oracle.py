"""ctypes front of the CPU oracle (liboracle.so built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py; the GPU library never uses it. See oracle.c for the
reference routine each function restates.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> None:
    subprocess.run(["make", "-C", str(HERE)], check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _lib.or_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib.or_sincos.restype = None
        _lib.or_atan.argtypes = [C.c_double]
        _lib.or_atan.restype = C.c_double
        _lib.or_set_threads.argtypes = [C.c_int]
        _lib.or_get_threads.restype = C.c_int
        _lib.or_philox_block.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib.or_move_all.restype = C.c_int
        _lib.or_move_all.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                                     C.c_uint32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.or_locate.restype = C.c_int
        _lib.or_locate.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
        _lib.or_deposit.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int64,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
        _lib.or_normalize.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.or_brinv.restype = C.c_int
        _lib.or_brinv.argtypes = [C.c_void_p, C.c_int]
        _lib.or_ppr.restype = C.c_int
        _lib.or_ppr.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p]
        _lib.or_ppr_gather.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.or_pcg.restype = C.c_int
        _lib.or_pcg.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_double, C.c_int, C.POINTER(C.c_int)]
        _lib.or_element_matrices.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p,
                                             C.c_void_p]
        _lib.or_ife_element.restype = C.c_int
        _lib.or_ife_element.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.or_rhs.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int, C.c_void_p]
        _lib.or_stiffness_apply.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                            C.c_void_p, C.c_void_p]
        _lib.or_mcc_select.restype = C.c_int64
        _lib.or_mcc_select.argtypes = [C.c_int64, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                       C.c_void_p]
        _lib.or_mcc_collide.restype = C.c_int
        _lib.or_mcc_collide.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.or_inject_consts.restype = None
        _lib.or_inject_consts.argtypes = [C.c_double] * 6 + [C.c_void_p]
        _lib.or_inject_count.restype = C.c_int64
        _lib.or_inject_count.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]
        _lib.or_inject.restype = None
        _lib.or_inject.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                   C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, C.c_int] + [C.c_void_p] * 7
        _lib.or_set_plant.argtypes = [C.c_int]
        _lib.or_set_plant.restype = None
        _lib.or_select_prob.restype = None
        _lib.or_select_prob.argtypes = [C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.or_post_test.restype = None
        _lib.or_post_test.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_uint64,
                                      C.c_uint32, C.c_void_p]
        _lib.or_vel_maxwell_test.restype = None
        _lib.or_vel_maxwell_test.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_void_p]
        _lib.or_load_velocity.restype = None
        _lib.or_load_velocity.argtypes = [C.c_int, C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_void_p]
        _lib.or_cell_volume.restype = None
        _lib.or_cell_volume.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        _lib.or_mcc_table.restype = C.c_double
        _lib.or_mcc_table.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                      C.c_void_p, C.c_void_p]
    return _lib


def sincos(x: float) -> tuple[float, float]:
    s, c = C.c_double(), C.c_double()
    lib().or_sincos(float(x), C.byref(s), C.byref(c))
    return s.value, c.value


def atan(t: float) -> float:
    return lib().or_atan(float(t))


def set_threads(n: int) -> None:
    """OpenMP threads of the particle, lookup, deposition, RHS, PCG and PPR loops (1 = the
    serial reference order, the default)."""
    lib().or_set_threads(int(n))


def get_threads() -> int:
    return int(lib().or_get_threads())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox_block(c, k, o)
    return np.array(list(o), dtype=np.uint32)


def tag(species: int, purpose: int) -> int:
    return (purpose << 16) | (species & 0xFFFF)


RNG_EMIT, RNG_SELECT, RNG_COLLIDE = 2, 3, 4


def move_all(bits, xs, ys, element_index, node_index, objects, efx, efy, coef5, qm, dt, v_ref, mass, init_temp_K,
             seed, step, tag_emit, x, y, vx, vy, vz, cell, delta=0, th=None):
    """MoveAll on copies; returns (x, y, vx, vy, vz, alive, counters, err), plus the azimuth th
    when delta = 1 (axisymmetric: x axial, y radial, th = PO%Z)."""
    ft = np.float64 if bits == 64 else np.float32
    st = [np.array(a, dtype=ft) for a in (x, y, vx, vy, vz)]
    tha = np.array(th if th is not None else np.zeros(len(cell)), dtype=ft)
    f = [np.ascontiguousarray(a, dtype=ft) for a in (efx, efy, coef5)]
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    alive = np.zeros(len(cell), dtype=np.int32)
    cnt = np.zeros(8, dtype=np.int64)
    objs = np.ascontiguousarray(objects, dtype=np.float64)
    err = lib().or_move_all(bits, len(xs), len(ys), _p(xs), _p(ys), _p(element_index), _p(node_index),
                            len(objects), _p(objs), int(delta), _p(f[0]), _p(f[1]), _p(f[2]), qm, dt, v_ref, mass,
                            init_temp_K, seed, step, tag_emit, len(cell), *[_p(a) for a in st], _p(tha), _p(cell),
                            _p(alive), _p(cnt))
    if delta == 1:
        return (*st, alive, cnt, err, tha)
    return (*st, alive, cnt, err)


def locate(xs, ys, x, y) -> tuple[np.ndarray, int]:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(len(x), dtype=np.int32)
    bad = lib().or_locate(len(xs), len(ys), _p(xs), _p(ys), len(x), _p(x), _p(y), _p(out))
    return out, bad


def deposit(bits, xs, ys, delta, x, y, vx, vy, vz, cell, mass, v_ref, fix):
    ft = np.float64 if bits == 64 else np.float32
    arrs = [np.ascontiguousarray(a, dtype=ft) for a in (x, y, vx, vy, vz)]
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    nn = len(xs) * len(ys)
    w, e = np.zeros(nn), np.zeros(nn)
    lib().or_deposit(bits, len(xs), len(ys), _p(xs), _p(ys), delta, len(cell), *[_p(a) for a in arrs], _p(cell),
                     mass, v_ref, 1 if fix else 0, _p(w), _p(e))
    return w, e


def normalize(delta, weight, qs_norm, L_ref, n_ref, vol, w, e, rho):
    rs, ek = np.zeros_like(w), np.zeros_like(w)
    lib().or_normalize(len(w), delta, weight, qs_norm, L_ref, n_ref, _p(np.ascontiguousarray(vol)), _p(w), _p(e),
                       _p(rs), _p(ek), _p(rho))
    return rs, ek


def brinv(a: np.ndarray) -> tuple[np.ndarray, int]:
    b = np.ascontiguousarray(a, dtype=np.float64).copy()
    ok = lib().or_brinv(_p(b), b.shape[0])
    return b, ok


def ppr(xs, ys, point_patch, phi):
    nn = len(xs) * len(ys)
    coef, efx, efy = np.zeros(6 * nn), np.zeros(nn), np.zeros(nn)
    pp = np.ascontiguousarray(point_patch, dtype=np.int32)
    ns = lib().or_ppr(len(xs), len(ys), _p(xs), _p(ys), _p(pp), _p(np.ascontiguousarray(phi)), _p(coef), _p(efx),
                      _p(efy))
    return coef, efx, efy, ns


def ppr_gather(xs, ys, element_index, node_index, coef, x, y, ele0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    e = np.ascontiguousarray(ele0, dtype=np.int32)
    ex, ey = np.zeros_like(x), np.zeros_like(x)
    lib().or_ppr_gather(len(xs), len(ys), _p(xs), _p(ys), _p(element_index), _p(node_index), _p(coef), len(x), _p(x),
                        _p(y), _p(e), _p(ex), _p(ey))
    return ex, ey


def pcg(rowptr, col, val, diag, b, x0, tol, maxit):
    """MY_PCG restated: returns (x, r, iter, ierr)."""
    b = np.array(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    it = C.c_int(-1)
    ierr = lib().or_pcg(len(diag), _p(np.ascontiguousarray(rowptr, dtype=np.int32)),
                        _p(np.ascontiguousarray(col, dtype=np.int32)), _p(np.ascontiguousarray(val)),
                        _p(np.ascontiguousarray(diag)), _p(b), _p(x), tol, maxit, C.byref(it))
    return x, b, it.value, ierr


def element_matrices(delta, x0, y0, hx, hy):
    K, M = np.zeros(16), np.zeros(16)
    lib().or_element_matrices(delta, x0, y0, hx, hy, _p(K), _p(M))
    return K.reshape(4, 4), M.reshape(4, 4)


def ife_element(delta, x0, y0, h, inside, cx, cy, rad, beta_in, beta_out):
    """One circle-cut element: (type, K 4x4, M 4x4, coef[4 basis][2 pieces][4] in X=(x-x0)/h,
    Y=(y-y0)/h, cut points D, E in X, Y, local indices of A1..A4)."""
    K, M, coef, pts = np.zeros(16), np.zeros(16), np.zeros(32), np.zeros(4)
    order = np.zeros(4, dtype=np.int32)
    ins = np.ascontiguousarray(inside, dtype=np.int32)
    t = lib().or_ife_element(delta, x0, y0, h, _p(ins), cx, cy, rad, beta_in, beta_out, _p(K), _p(M), _p(coef),
                             _p(pts), _p(order))
    return t, K.reshape(4, 4), M.reshape(4, 4), coef.reshape(4, 2, 4), pts.reshape(2, 2), order


def assemble_ife(xs, ys, delta, circles, node_type2, dphi, mistake=None):
    """Element-by-element restatement of the stiffness on unknowns A, the Dirichlet lift c (A_xt
    times the Dirichlet values) and the RHS operator B (unknowns x nodes) for a square mesh with
    circular objects (cx, cy, r, eps, conductor): vacuum elements Q1 with beta 1, elements inside
    an object Q1 with beta eps, elements cut by an object of eps != 1 through ife_element (a
    conductor's inside nodes stay Dirichlet), cut elements of eps = 1 Q1
    (Generate_stiffness_matrix_local_IFE.f90, Global_RHS_PDE_2D.f90).
    mistake plants an error for the tests that show the checks catch it: "q1_cut" treats the
    dielectric's cut elements as plain Q1, "swap" swaps the two permittivities on them."""
    import scipy.sparse as sp
    nnx, nny = len(xs), len(ys)
    h = xs[1] - xs[0]
    nu = int(node_type2.max())
    inside = np.full(nnx * nny, -1)
    for k, (cx, cy, r, eps, cond) in enumerate(circles):
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        m = ((X - cx) ** 2 + (Y - cy) ** 2 < r * r).ravel() & (inside < 0)
        inside[m] = k
    Ar, Ac, Av, Br, Bc, Bv = [], [], [], [], [], []
    c = np.zeros(nu)
    types = []
    for ci in range(nnx - 1):
        for cj in range(nny - 1):
            nodes = [ci * nny + cj, (ci + 1) * nny + cj, (ci + 1) * nny + cj + 1, ci * nny + cj + 1]
            objs = [inside[n] for n in nodes]
            kin = sum(o >= 0 for o in objs)
            K, M = element_matrices(delta, xs[ci], ys[cj], xs[ci + 1] - xs[ci], ys[cj + 1] - ys[cj])
            if kin == 4:
                cx, cy, r, eps, cond = circles[objs[0]]
                K = K * eps
            elif kin > 0:
                k = next(o for o in objs if o >= 0)
                cx, cy, r, eps, cond = circles[k]
                if eps != 1.0 and mistake != "q1_cut":
                    b_in, b_out = (1.0, eps) if mistake == "swap" else (eps, 1.0)
                    t, K2, M2, _, _, _ = ife_element(delta, xs[ci], ys[cj], h, [o >= 0 for o in objs], cx, cy, r,
                                                     b_in, b_out)
                    assert t in (1, 2), t
                    K, M = K2, M2
                    types.append(t)
            for a in range(4):
                u = node_type2[nodes[a]]
                if u <= 0:
                    continue
                for b in range(4):
                    ub = node_type2[nodes[b]]
                    if ub > 0:
                        Ar.append(u - 1); Ac.append(ub - 1); Av.append(K[a, b])
                    else:
                        c[u - 1] += K[a, b] * dphi[nodes[b]]
                    Br.append(u - 1); Bc.append(nodes[b]); Bv.append(M[a, b])
    A = sp.csr_matrix((Av, (Ar, Ac)), shape=(nu, nu))
    B = sp.csr_matrix((Bv, (Br, Bc)), shape=(nu, nnx * nny))
    return A, c, B, types


def rhs(xs, ys, delta, node_type2, dphi, rho, nunk):
    out = np.zeros(nunk)
    lib().or_rhs(len(xs), len(ys), _p(xs), _p(ys), delta, _p(np.ascontiguousarray(node_type2, dtype=np.int32)),
                 _p(np.ascontiguousarray(dphi)), _p(np.ascontiguousarray(rho)), nunk, _p(out))
    return out


def stiffness_apply(xs, ys, delta, node_type2, xu):
    xu = np.ascontiguousarray(xu, dtype=np.float64)
    out = np.zeros_like(xu)
    lib().or_stiffness_apply(len(xs), len(ys), _p(xs), _p(ys), delta,
                             _p(np.ascontiguousarray(node_type2, dtype=np.int32)), len(xu), _p(xu), _p(out))
    return out


def mcc_select(cell, threshold, seed, step, tag_):
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    out = np.zeros(len(cell), dtype=np.int32)
    k = lib().or_mcc_select(len(cell), _p(cell), threshold, seed, step, tag_, _p(out))
    return out[:k]


def mcc_collide(electron, slot, seed, step, tag_, v_ref, mass, gas_mass, gas_temp, nreact, nsigma, emin, eint, emax,
                prob, rtype, thr, v):
    v = np.array(v, dtype=np.float64)
    ne, ni = np.zeros(3), np.zeros(3)
    idx = lib().or_mcc_collide(electron, slot, seed, step, tag_, v_ref, mass, gas_mass, gas_temp, nreact, nsigma, emin,
                               eint, emax, _p(np.ascontiguousarray(prob)),
                               _p(np.ascontiguousarray(rtype, dtype=np.int32)), _p(np.ascontiguousarray(thr)), _p(v),
                               _p(ne), _p(ni))
    return idx, v, ne, ni


def mcc_table(nreact, nsigma, emin, eint, mass, gas_mass, density, sigma):
    prob = np.zeros(nreact * nsigma)
    smax = lib().or_mcc_table(nreact, nsigma, emin, eint, mass, gas_mass, density,
                              _p(np.ascontiguousarray(sigma, dtype=np.float64)), _p(prob))
    return prob, smax


RNG_INJECT = 5


def inject_consts(vd1, vt, tmpj, xm, r_start, dt):
    """b_amb, v_r, FS1, FS2, QA, sqrt(2) vt, VMP, XL, XU of InjectBeams_2D / Loadv."""
    c = np.zeros(9)
    lib().or_inject_consts(vd1, vt, tmpj, xm, r_start, dt, _p(c))
    return c


def inject_count(expect_no_r, seed, step, tag_):
    return int(lib().or_inject_count(expect_no_r, seed, step, tag_))


def inject(bits, n, electron, consts, vd, y_lo, y_hi, seed, step, tag_, xmin, ymin, mapx, mapy, nex, ney, delta=0):
    """n injected particles of one source: x, y, vx, vy, vz, 0-based cell (-1 outside), and on
    an axisymmetric domain (delta = 1) the azimuth as a seventh array."""
    out = [np.zeros(n) for _ in range(5)] + [np.zeros(n, dtype=np.int32)]
    th = np.zeros(n) if delta == 1 else None
    c = np.ascontiguousarray(consts, dtype=np.float64)
    v = np.ascontiguousarray(vd, dtype=np.float64)
    lib().or_inject(bits, n, electron, _p(c), _p(v), y_lo, y_hi, seed, step, tag_, xmin, ymin, mapx, mapy, nex, ney,
                    *[_p(a) for a in out], _p(th))
    return out + [th] if delta == 1 else out


# ------------------------------------------------------------------ pins of the MCC and set-up restatements
class planted:
    """Context manager: a planted mistake in the oracle (codes in oracle.c, or_set_plant),
    for the tests that show a check catches it."""

    def __init__(self, code: int):
        self.code = code

    def __enter__(self):
        lib().or_set_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib().or_set_plant(0)
        return False


def select_prob(energy, prob, thr, nreact, nsigma, emin, eint, emax):
    """SelectProbility's reaction probabilities at `energy` (eV), before the draw."""
    P = np.zeros(4)
    lib().or_select_prob(float(energy), nreact, nsigma, emin, eint, emax, _p(np.ascontiguousarray(prob, dtype=float)),
                         _p(np.ascontiguousarray(thr, dtype=float)), _p(P))
    return P[:nreact]


def post_collision(g, v, energy, mass_ratio, iso, seed, slot):
    """PostCollisionVelocity of a projectile (m/s) with relative velocity g (m/s)."""
    out = np.zeros(3)
    lib().or_post_test(_p(np.ascontiguousarray(g, dtype=float)), _p(np.ascontiguousarray(v, dtype=float)),
                       float(energy), float(mass_ratio), int(iso), seed, slot, _p(out))
    return out


def vel_maxwell(seed, n, mass, temp):
    """n target-gas velocities (m/s) of VelMaxInit, one Philox stream each."""
    v = np.zeros((n, 3))
    lib().or_vel_maxwell_test(seed, n, mass, temp, _p(v))
    return v


def load_velocity(kind, seed, n, temp_K, mass):
    """n initial velocities of VelKappaInit (kind 0) or VelPolyInit (kind 2)."""
    v = np.zeros((n, 3))
    lib().or_load_velocity(int(kind), seed, n, float(temp_K), float(mass), _p(v))
    return v


def cell_volume(xs, ys, delta):
    """Cell_Volume_zwz(2,:) in node order (i * nny + j)."""
    xs = np.ascontiguousarray(xs, dtype=float)
    ys = np.ascontiguousarray(ys, dtype=float)
    vol = np.zeros(len(xs) * len(ys))
    lib().or_cell_volume(len(xs), len(ys), _p(xs), _p(ys), int(delta), _p(vol))
    return vol
