"""Python binding of the MI355X IFE-PIC hot-path library (libpichip.so, include/pic_hip.h).

The functions below keep the reference subroutine names and argument meaning
(MoveAll, GetParChg_2D, IFE_Solve_2D, GetEField_SIDG_PPR, MCC, MY_JPCG_Solver,
Get_ParticleEField_PPR, InjectPositioning, ParticlePositioning, generate_Elementmap);
every C status other than PIC_OK raises PicError carrying the library's message, the
way the reference STOPs or reports IERR. The compute path is the HIP library only:
importing this module raises if libpichip.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpichip.so"

# ----------------------------------------------------------------- constants
PIC_OK = 0
ERRORS = {
    1: "PIC_ERR_ARG", 2: "PIC_ERR_STATE", 3: "PIC_ERR_HIP", 4: "PIC_ERR_CAPACITY", 5: "PIC_ERR_PCG_MAXITER",
    6: "PIC_ERR_PCG_BREAKDOWN", 7: "PIC_ERR_NONFINITE", 8: "PIC_ERR_POSITION", 9: "PIC_ERR_UNSUPPORTED",
    10: "PIC_ERR_BOUNDARY", 11: "PIC_ERR_NO_DEVICE", 12: "PIC_ERR_OVERFLOW",
}
SETUP = dict(HP=1, HT=2, ELEMENT_INDEX=3, NODE_INDEX=4, NODE_TYPE=5, CELL_VOLUME=6, A_ROWPTR=7, A_COL=8, A_VAL=9,
             A_DIAG=10, B_ROWPTR=11, B_COL=12, B_VAL=13, RHS_CONST=14, DIRICHLET_PHI=15, POINT_PATCH=16,
             UNKNOWN_NODE=17)
FIELD = dict(RHO=1, PHI=2, EFX=3, EFY=4, COEF_PPR=5, RHO_S=6, RHO_P=7, EK_TOT=8, RHS=9, EK_S=10)

# physical constants of the reference (MCC_jw/code/base/Constants.f90)
ELECTRON_MASS = 9.1095e-31
ELECTRON_CHARGE = 1.6022e-19
KB = 1.3807e-23
EVTOK = 11605.0


class PicError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


# ----------------------------------------------------------------- C structs
class Sparse(C.Structure):
    """TYPE SPARSE (IFE_MAIN_PARAM.f90:8-11)."""
    _fields_ = [("K", C.c_double), ("JCOL", C.c_int32), ("SROW", C.c_int32)]


class ParticleOne(C.Structure):
    """TYPE ParticleOne (MCC_jw/code/base/ParticleOne.f90:4-6)."""
    _fields_ = [(n, C.c_double) for n in ("X", "Y", "Z", "Vx", "Vy", "Vz", "Ax", "Ay", "Az", "Wq")] + \
               [("Location", C.c_int32), ("pad_", C.c_int32)]


class Object(C.Structure):
    _fields_ = [("shape", C.c_int32), ("conductor", C.c_int32), ("dimensions", C.c_double * 2),
                ("locations", (C.c_double * 2) * 4), ("phi", C.c_double), ("eps", C.c_double)]


class DomainConfig(C.Structure):
    _fields_ = [("nnx", C.c_int32), ("nny", C.c_int32), ("delta", C.c_int32), ("n_objects", C.c_int32),
                ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("objects", Object * 8), ("wall_dirichlet", C.c_int32 * 4), ("wall_phi", C.c_double * 4),
                ("n_ref", C.c_double), ("T_ref_eV", C.c_double), ("dt", C.c_double)]


class SpeciesParams(C.Structure):
    _fields_ = [("charge_C", C.c_double), ("mass_kg", C.c_double), ("weight", C.c_double),
                ("init_temperature_K", C.c_double)]


class GasParams(C.Structure):
    _fields_ = [("mass_kg", C.c_double), ("temperature_K", C.c_double), ("density_m3", C.c_double)]


class Injection(C.Structure):
    _fields_ = [("density_m3", C.c_double), ("temp_eV", C.c_double), ("drift_ms", C.c_double * 3),
                ("r_start", C.c_double), ("y_lo", C.c_double), ("y_hi", C.c_double)]


_P = C.POINTER
_dp = _P(C.c_double)
_ip = _P(C.c_int32)
_i64p = _P(C.c_int64)
_fp = _P(C.c_float)

_PROTOS = {
    "pic_init": (C.c_int, [C.c_int]),
    "pic_finalize": (C.c_int, []),
    "pic_last_error": (C.c_char_p, []),
    "pic_device_count": (C.c_int, [_ip]),
    "pic_set_precision": (C.c_int, [C.c_int, C.c_int]),
    "pic_setup_domain": (C.c_int, [_P(DomainConfig)]),
    "pic_get_sizes": (C.c_int, [_ip, _ip, _ip, _i64p, _i64p]),
    "pic_get_setup_array": (C.c_int, [C.c_int, C.c_void_p, C.c_int64]),
    "pic_set_species": (C.c_int, [C.c_int, _P(SpeciesParams), C.c_int64]),
    "pic_load_maxwellian": (C.c_int, [C.c_int, C.c_int32, C.c_double, C.c_uint64]),
    "pic_load_distribution": (C.c_int, [C.c_int, C.c_int32, C.c_double, C.c_uint64, C.c_int]),
    "pic_load_bundle": (C.c_int, [C.c_int, _P(ParticleOne), C.c_int64]),
    "pic_store_bundle": (C.c_int, [C.c_int, _P(ParticleOne), C.c_int64, _i64p]),
    "pic_load_particles": (C.c_int, [C.c_int, C.c_int64, _dp, _dp, _dp, _dp, _dp, _ip]),
    "pic_get_particles": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, C.c_int64, _i64p]),
    "pic_set_azimuth": (C.c_int, [C.c_int, C.c_int64, _dp]),
    "pic_DumpParticleBundle": (C.c_int, [C.c_int, C.c_char_p]),
    "pic_LoadParticleBundle": (C.c_int, [C.c_int, C.c_char_p, _i64p]),
    "pic_Dump_2D": (C.c_int, [C.c_int, C.c_char_p]),
    "pic_Restart_2D": (C.c_int, [C.c_int, C.c_char_p]),
    "pic_Output_To_Tecplot_IJ": (C.c_int, [C.c_int, C.c_char_p]),
    "pic_format_fortran_e": (C.c_int, [C.c_double, C.c_int, C.c_int, C.c_char_p]),
    "pic_get_azimuth": (C.c_int, [C.c_int, _dp, C.c_int64, _i64p]),
    "pic_get_counts": (C.c_int, [C.c_int, _i64p, _i64p, _i64p]),
    "pic_get_mcc_events": (C.c_int, [C.c_int, _i64p, C.c_int]),
    "pic_get_counters": (C.c_int, [C.c_int, _i64p, C.c_int]),
    "pic_MoveAll": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "pic_GetParChg_2D": (C.c_int, [C.c_int]),
    "pic_IFE_Solve_2D": (C.c_int, [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double]),
    "pic_last_solve_info": (C.c_int, [_ip, _dp, _ip]),
    "pic_last_solve_time": (C.c_int, [_dp]),
    "pic_set_pcg_params": (C.c_int, [C.c_double, C.c_int32]),
    "pic_GetEField_SIDG_PPR": (C.c_int, []),
    "pic_MCC": (C.c_int, []),
    "pic_MCCBundleInit": (C.c_int, [_P(GasParams), C.c_int, C.c_int]),
    "pic_mcc_set_table": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _ip, _dp, C.c_double,
                                    C.c_double, C.c_int]),
    "pic_mcc_get_table": (C.c_int, [C.c_int, _dp, C.c_int64, _dp, _dp]),
    "pic_set_injection": (C.c_int, [C.c_int, _P(Injection)]),
    "pic_InjectBeams_2D": (C.c_int, [C.c_int, C.c_int, _i64p]),
    "pic_step": (C.c_int, [C.c_int]),
    "pic_initial_field": (C.c_int, []),
    "pic_set_run_params": (C.c_int, [C.c_int32, C.c_uint64, C.c_int32]),
    "pic_sort_particles": (C.c_int, []),
    "pic_sort_group": (C.c_int, [_ip]),
    "pic_get_step": (C.c_int, [_i64p]),
    "pic_set_step": (C.c_int, [C.c_int64]),
    "pic_get_field": (C.c_int, [C.c_int, _dp, C.c_int64]),
    "pic_set_field": (C.c_int, [C.c_int, _dp, C.c_int64]),
    "pic_get_phase_times": (C.c_int, [_dp, C.c_int]),
    "pic_MY_JPCG_Solver": (C.c_int, [_P(Sparse), _dp, _dp, _dp, C.c_int, C.c_int, _ip]),
    "pic_MY_JPCG_Solver_noalloc": (C.c_int, [_P(Sparse), _dp, _dp, _dp, C.c_int, C.c_int, _ip]),
    "pic_Get_ParticleEField_PPR": (C.c_int, [C.c_int64, _dp, _dp, _ip, _dp, _dp]),
    "pic_InjectPositioning": (C.c_int, [C.c_int64, _dp, _dp, _ip]),
    "pic_ParticlePositioning": (C.c_int, [C.c_int64, _dp, _dp, _ip, _ip]),
    "pic_locate": (C.c_int, [C.c_int64, _dp, _dp, _ip]),
    "pic_generate_Elementmap": (C.c_int, [_ip, C.c_int]),
    "pic_learn_setup": (C.c_int, [C.c_int32, _ip, C.c_int32, C.c_uint64]),
    "pic_learn_harvest": (C.c_int, [C.c_int32, C.c_int32, C.c_int32]),
    "pic_learn_train": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, _dp, _dp]),
    "pic_learn_eval": (C.c_int, [C.c_int32, _dp]),
    "pic_learn_info": (C.c_int, [_ip, _ip, _i64p, _dp]),
    "pic_learn_finetune": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_uint64]),
    "pic_learn_data_mix": (C.c_int, [C.c_double]),
    "pic_learn_finetune_info": (C.c_int, [_i64p, _dp, _dp]),
    "pic_learn_apply": (C.c_int, [_fp, _fp, C.c_int64]),
    "pic_learn_get_params": (C.c_int, [_fp, C.c_int64, _i64p]),
    "pic_learn_set_params": (C.c_int, [_fp, C.c_int64]),
    "pic_learn_layout": (C.c_int, [_ip, _ip, _ip, _ip, _i64p]),
    "pic_learn_set_pairs": (C.c_int, [C.c_int32, _fp, _fp]),
    "pic_learn_get_pair": (C.c_int, [C.c_int32, _fp, _fp, _dp]),
    "pic_learn_gradient": (C.c_int, [C.c_int32, _ip, _dp, _fp, C.c_int64]),
    "pic_warm_setup": (C.c_int, [C.c_int32, C.c_int32, C.c_uint64]),
    "pic_warm_train": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_uint64, _dp, _dp]),
    "pic_warm_enable": (C.c_int, [C.c_int32]),
    "pic_warm_eval": (C.c_int, [C.c_int32, _dp, _dp]),
    "pic_warm_apply": (C.c_int, [_fp, _fp, C.c_int64]),
    "pic_warm_get_params": (C.c_int, [_fp, C.c_int64, _i64p]),
    "pic_warm_set_params": (C.c_int, [_fp, C.c_int64]),
    "pic_warm_layout": (C.c_int, [_ip, _ip, _ip, _ip]),
    "pic_warm_gradient": (C.c_int, [C.c_int32, _ip, _dp, _fp, C.c_int64]),
    "pic_warm_info": (C.c_int, [_ip, _ip, _dp, _dp, _i64p]),
}


def header_symbols() -> list[str]:
    """Every function name declared in include/pic_hip.h."""
    import re
    text = (_HERE.parent / "include" / "pic_hip.h").read_text()
    return sorted(set(re.findall(r"\b(pic_[A-Za-z0-9_]+)\s*\(", text)))


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} not built: run __graft_entry__.build() (the HIP library is the only compute path)")
    lib = C.CDLL(str(p))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def check(rc: int) -> None:
    if rc != PIC_OK:
        raise PicError(rc, lib().pic_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f(a):
    return a.ctypes.data_as(_fp)


# ----------------------------------------------------------------- configuration
@dataclass
class Circle:
    x: float
    y: float
    r: float
    phi: float = 0.0
    conductor: bool = True
    eps: float = 1.0


@dataclass
class Domain:
    """A synthetic IFE-PIC workload (what mesh.inp / object.inp / pic.inp carry)."""
    nnx: int
    nny: int
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = 0.0
    ymax: float = 1.0
    delta: int = 0
    objects: list = field(default_factory=list)
    wall_dirichlet: tuple = (1, 1, 1, 1)
    wall_phi: tuple = (0.0, 0.0, 0.0, 0.0)
    n_ref: float = 1e21
    T_ref_eV: float = 1.0
    dt: float = 0.05

    def to_c(self) -> DomainConfig:
        c = DomainConfig()
        c.nnx, c.nny, c.delta, c.n_objects = self.nnx, self.nny, self.delta, len(self.objects)
        c.xmin, c.xmax, c.ymin, c.ymax = self.xmin, self.xmax, self.ymin, self.ymax
        for k, o in enumerate(self.objects):
            c.objects[k].shape = 1
            c.objects[k].conductor = 1 if o.conductor else 0
            c.objects[k].dimensions[0] = o.r
            c.objects[k].locations[0][0] = o.x
            c.objects[k].locations[0][1] = o.y
            c.objects[k].phi = o.phi
            c.objects[k].eps = o.eps
        for w in range(4):
            c.wall_dirichlet[w] = int(self.wall_dirichlet[w])
            c.wall_phi[w] = float(self.wall_phi[w])
        c.n_ref, c.T_ref_eV, c.dt = self.n_ref, self.T_ref_eV, self.dt
        return c

    def object_array(self) -> np.ndarray:
        """6 doubles per object (shape, x0, y0, x1, y1, r^2) for the CPU oracle."""
        a = np.zeros((max(len(self.objects), 1), 6))
        for k, o in enumerate(self.objects):
            a[k] = (1, o.x, o.y, 0.0, 0.0, o.r * o.r)
        return a

    def normalisation(self) -> dict:
        """Input_2D.f90:356-398 reference values."""
        eps0, e, me = 8.8542e-12, ELECTRON_CHARGE, ELECTRON_MASS
        T = self.T_ref_eV * EVTOK
        return dict(L_ref=np.sqrt(eps0 * KB * T / self.n_ref / e / e),
                    time_ref=np.sqrt(eps0 * me / self.n_ref / e / e),
                    v_ref=np.sqrt(KB * T / me), phi_ref=KB * T / e)


MASS_RATIO = 73440.0
ELECTRON = dict(charge_C=-ELECTRON_CHARGE, mass_kg=ELECTRON_MASS)
ARGON_ION = dict(charge_C=ELECTRON_CHARGE, mass_kg=MASS_RATIO * ELECTRON_MASS)


def workload(name: str) -> dict:
    """The synthetic workloads: geometry on the unit square scaled to the mesh extent at
    h = 1 L_ref (the reference's mesh.inp convention, hx = 1.0)."""
    if name == "cfg1":            # 64^2, one conductor at -1, FP64 reference-equivalent
        n = 64
        dom = Domain(n + 1, n + 1, 0.0, float(n), 0.0, float(n),
                     objects=[Circle(0.5 * n, 0.5 * n, 0.15 * n, -1.0)])
        return dict(domain=dom, ppc=100, bits=64, solver=0, tol=1e-10, Te=2.0, Ti=0.026)
    if name in ("cfg3", "cfg5") or name.startswith("cfg3s"):
        # three conductors at -1, 0.5, 0, mixed precision: cfg3 on 512^2, its layout scaled to
        # 1024^2 for cfg5 and to n^2 for the cfg3s<n> parity cases
        n = {"cfg3": 512, "cfg5": 1024}.get(name) or int(name[5:])
        dom = Domain(n + 1, n + 1, 0.0, float(n), 0.0, float(n),
                     objects=[Circle(0.25 * n, 0.30 * n, 0.10 * n, -1.0), Circle(0.65 * n, 0.65 * n, 0.12 * n, 0.5),
                              Circle(0.70 * n, 0.22 * n, 0.08 * n, 0.0)])
        # 71 per vacuum element: 16.8M particles per species at 512^2 (the conductors take 9.7% of
        # the area); cfg5: 35 per vacuum element, 33.1M per species at 1024^2. Both species are
        # injected through the left wall as a Maxwellian drifting inward at 0.1 v_th
        return dict(domain=dom, ppc=35 if name == "cfg5" else 71, bits=32, solver=1, tol=1e-8, Te=2.0, Ti=0.026,
                    inject_drift_vth=0.1)
    if name == "cfg2":            # 256^2 r-z, a conductor at -1 and a dielectric disk (eps 4), mixed precision
        n = 256
        dom = Domain(n + 1, n + 1, 0.0, float(n), 0.0, float(n), delta=1,
                     objects=[Circle(0.30 * n, 0.45 * n, 0.10 * n, -1.0),
                              Circle(0.68 * n, 0.50 * n, 0.12 * n, 0.0, conductor=False, eps=4.0)],
                     wall_dirichlet=(1, 1, 0, 1))
        return dict(domain=dom, ppc=64, bits=32, solver=1, tol=1e-8, Te=2.0, Ti=0.026)
    if name.startswith("diel"):   # dielectric parity cases diel<N> (Cartesian), dielrz<N> (r-z)
        rz = name.startswith("dielrz")
        n = int(name[6 if rz else 4:] or 24)
        dom = Domain(n + 1, n + 1, 0.0, float(n), 0.0, float(n), delta=1 if rz else 0,
                     objects=[Circle(0.27 * n, 0.48 * n, 0.12 * n, -1.0),
                              Circle(0.66 * n, 0.47 * n, 0.19 * n, 0.0, conductor=False, eps=4.0)],
                     wall_dirichlet=(1, 1, 0, 1) if rz else (1, 1, 1, 1))
        return dict(domain=dom, ppc=8, bits=64, solver=0, tol=1e-12, Te=2.0, Ti=0.026)
    if name.startswith("rz"):     # axisymmetric parity cases rz<N>: x axial, y radial, the axis (bottom) natural
        n = int(name[2:] or 16)
        dom = Domain(n + 1, n + 1, 0.0, float(n), 0.0, float(n), delta=1,
                     objects=[Circle(0.5 * n, 0.5 * n, 0.15 * n, -1.0)], wall_dirichlet=(1, 1, 0, 1))
        return dict(domain=dom, ppc=8, bits=64, solver=0, tol=1e-12, Te=2.0, Ti=0.026)
    if name.startswith("skew"):   # parity cases skew<N>: origin (0.3, -0.1), spacings 0.7 x 0.9 (neither a
        # power of two nor exact in FP32), so the lookup, the deposition weights and the node lines take
        # their general paths (FP64 location, divisions by the spacing, FP32 node lines rounded from FP64)
        n = int(name[4:] or 24)
        x0, y0, hx, hy = 0.3, -0.1, 0.7, 0.9
        dom = Domain(n + 1, n + 1, x0, x0 + hx * n, y0, y0 + hy * n,
                     objects=[Circle(x0 + 0.5 * hx * n, y0 + 0.5 * hy * n, 0.15 * hx * n, -1.0)])
        return dict(domain=dom, ppc=8, bits=64, solver=0, tol=1e-12, Te=2.0, Ti=0.026)
    if name.startswith("small"):  # small parity cases: small<N>
        n = int(name[5:] or 16)
        dom = Domain(n + 1, n + 1, 0.0, float(n), 0.0, float(n),
                     objects=[Circle(0.5 * n, 0.5 * n, 0.15 * n, -1.0)])
        return dict(domain=dom, ppc=8, bits=64, solver=0, tol=1e-12, Te=2.0, Ti=0.026)
    raise KeyError(name)


# ----------------------------------------------------------------- high level API
class PicLibrary:
    """Owns the device state (the reference's module globals)."""

    def __init__(self, device: int = 0):
        check(lib().pic_init(device))

    # set-up
    def setup(self, dom: Domain, bits: int = 64, solver: int = 0, tol: float = 1e-9, maxit: int = 50000):
        check(lib().pic_set_precision(bits, solver))
        c = dom.to_c()
        check(lib().pic_setup_domain(C.byref(c)))
        check(lib().pic_set_pcg_params(tol, maxit))
        self.dom = dom
        self.bits = bits

    def set_solver(self, solver: int) -> None:
        """0 FP64 Jacobi-PCG, 1 mixed precision, 2 mixed precision with the learned FCG,
        3 FP64 multicolour IC(0)-PCG (comparison)."""
        check(lib().pic_set_precision(self.bits, solver))

    def sizes(self):
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        d, e = C.c_int64(), C.c_int64()
        check(lib().pic_get_sizes(C.byref(a), C.byref(b), C.byref(c), C.byref(d), C.byref(e)))
        return dict(nnodes=a.value, nelem=b.value, nunknowns=c.value, nnz_A=d.value, nnz_B=e.value)

    def setup_array(self, name: str) -> np.ndarray:
        s = self.sizes()
        nn, ne, nu = s["nnodes"], s["nelem"], s["nunknowns"]
        spec = dict(HP=(2 * nn, np.float64), HT=(5 * ne, np.int32), ELEMENT_INDEX=(ne, np.int32),
                    NODE_INDEX=(nn, np.int32), NODE_TYPE=(2 * nn, np.int32), CELL_VOLUME=(2 * nn, np.float64),
                    A_ROWPTR=(nu + 1, np.int32), A_COL=(s["nnz_A"], np.int32), A_VAL=(s["nnz_A"], np.float64),
                    A_DIAG=(nu, np.float64), B_ROWPTR=(nu + 1, np.int32), B_COL=(s["nnz_B"], np.int32),
                    B_VAL=(s["nnz_B"], np.float64), RHS_CONST=(nu, np.float64), DIRICHLET_PHI=(nn, np.float64),
                    POINT_PATCH=(nn * 50, np.int32), UNKNOWN_NODE=(nu, np.int32))[name]
        a = np.zeros(spec[0], dtype=spec[1])
        check(lib().pic_get_setup_array(SETUP[name], a.ctypes.data_as(C.c_void_p), a.size))
        return a

    def set_species(self, isp: int, charge_C: float, mass_kg: float, weight: float, init_temperature_K: float,
                    capacity: int):
        p = SpeciesParams(charge_C, mass_kg, weight, init_temperature_K)
        check(lib().pic_set_species(isp, C.byref(p), int(capacity)))

    def load_maxwellian(self, isp: int, ppc: int, temp_eV: float, seed: int):
        check(lib().pic_load_maxwellian(isp, ppc, temp_eV, seed))

    def load_distribution(self, isp: int, ppc: int, temp_eV: float, seed: int, kind: int):
        """Initial load with the reference's MaxKappa velocity law: 0 kappa, 1 Maxwellian,
        2 polytropic (MCCInterface.f90:168-176)."""
        check(lib().pic_load_distribution(isp, ppc, temp_eV, seed, kind))

    def load_particles(self, isp, x, y, vx, vy, vz, loc=None, th=None):
        """th: the azimuth PO%Z of each particle (axisymmetric domains)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, vx, vy, vz)]
        lp = None if loc is None else _i(np.ascontiguousarray(loc, dtype=np.int32))
        check(lib().pic_load_particles(isp, len(arrs[0]), *[_d(a) for a in arrs], lp))
        if th is not None:
            t = np.ascontiguousarray(th, dtype=np.float64)
            check(lib().pic_set_azimuth(isp, len(t), _d(t)))

    # files in the reference's formats
    def dump_bundle(self, isp: int, path: str):
        check(lib().pic_DumpParticleBundle(isp, str(path).encode()))

    def load_bundle_file(self, isp: int, path: str) -> int:
        n = C.c_int64()
        check(lib().pic_LoadParticleBundle(isp, str(path).encode(), C.byref(n)))
        return n.value

    def dump_2d(self, lap: int, directory: str):
        check(lib().pic_Dump_2D(lap, str(directory).encode()))

    def restart_2d(self, lap: int, directory: str):
        check(lib().pic_Restart_2D(lap, str(directory).encode()))

    def tecplot_field(self, it: int, directory: str):
        check(lib().pic_Output_To_Tecplot_IJ(it, str(directory).encode()))

    def azimuth(self, isp: int) -> np.ndarray:
        """PO%Z of the live particles, in the order of particles() (axisymmetric domains)."""
        ns = C.c_int64()
        check(lib().pic_get_counts(isp, None, C.byref(ns), None))
        out = np.zeros(max(ns.value, 1))
        n = C.c_int64()
        check(lib().pic_get_azimuth(isp, _d(out), len(out), C.byref(n)))
        return out[:n.value]

    def particles(self, isp: int) -> dict:
        nl, ns = C.c_int64(), C.c_int64()
        check(lib().pic_get_counts(isp, None, C.byref(ns), None))
        cap = max(ns.value, 1)
        out = {k: np.zeros(cap) for k in ("x", "y", "vx", "vy", "vz")}
        loc = np.zeros(cap, dtype=np.int32)
        n = C.c_int64()
        check(lib().pic_get_particles(isp, *[_d(out[k]) for k in ("x", "y", "vx", "vy", "vz")], _i(loc), cap,
                                      C.byref(n)))
        res = {k: v[:n.value] for k, v in out.items()}
        res["loc"] = loc[:n.value]
        return res

    def counts(self, isp: int) -> dict:
        nl, ns = C.c_int64(), C.c_int64()
        c = (C.c_int64 * 8)()
        check(lib().pic_get_counts(isp, C.byref(nl), C.byref(ns), c))
        return dict(nlive=nl.value, nslots=ns.value, emit=c[0], right=c[1], object=c[2], reflect=c[3],
                    mcc_selected=c[4], mcc_real=c[5], mcc_created=c[6], deleted=c[7])

    def counters(self, isp: int) -> dict:
        """Event counters of a species, with the wall-injected particle count."""
        a = np.zeros(9, dtype=np.int64)
        check(lib().pic_get_counters(isp, a.ctypes.data_as(_i64p), 9))
        return dict(emit=int(a[0]), right=int(a[1]), object=int(a[2]), reflect=int(a[3]), mcc_selected=int(a[4]),
                    mcc_real=int(a[5]), mcc_created=int(a[6]), deleted=int(a[7]), injected=int(a[8]))

    def mcc_events(self, isp: int, n: int = 4) -> np.ndarray:
        a = np.zeros(n, dtype=np.int64)
        check(lib().pic_get_mcc_events(isp, a.ctypes.data_as(_i64p), n))
        return a

    def field(self, name: str) -> np.ndarray:
        s = self.sizes()
        nn = s["nnodes"]
        n = dict(COEF_PPR=6 * nn, RHO_S=4 * nn, RHO_P=4 * nn, EK_TOT=4 * nn, EK_S=4 * nn,
                 RHS=s["nunknowns"]).get(name, nn)
        a = np.zeros(n)
        check(lib().pic_get_field(FIELD[name], _d(a), n))
        return a

    def set_field(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().pic_set_field(FIELD[name], _d(a), a.size))

    # reference-named drop-ins
    def MoveAll(self, isp: int, delta: int | None = None, PushFlag: int = 0):
        check(lib().pic_MoveAll(isp, self.dom.delta if delta is None else delta, PushFlag))

    def GetParChg_2D(self, delta: int | None = None):
        check(lib().pic_GetParChg_2D(self.dom.delta if delta is None else delta))

    def IFE_Solve_2D(self, Rho=None, Phi=None, delta: int | None = None, time: float = 0.0):
        d = self.dom
        delta = d.delta if delta is None else delta
        r = None if Rho is None else _d(np.ascontiguousarray(Rho, dtype=np.float64))
        p = None
        if Phi is not None:
            assert Phi.dtype == np.float64 and Phi.flags.c_contiguous
            p = _d(Phi)
        check(lib().pic_IFE_Solve_2D(r, p, d.nnx, d.nny, delta, time))

    def last_solve(self) -> dict:
        it, rr, ou = C.c_int32(), C.c_double(), C.c_int32()
        check(lib().pic_last_solve_info(C.byref(it), C.byref(rr), C.byref(ou)))
        return dict(iterations=it.value, relres=rr.value, outer=ou.value)

    def last_solve_ms(self) -> float:
        """Device time of the solver in the last IFE_Solve_2D (RHS assembly excluded)."""
        ms = C.c_double()
        check(lib().pic_last_solve_time(C.byref(ms)))
        return ms.value

    def GetEField_SIDG_PPR(self):
        check(lib().pic_GetEField_SIDG_PPR())

    def MCCBundleInit(self, gas_mass_kg: float, gas_temperature_K: float, gas_density_m3: float,
                      electron_species: int = 0, ion_species: int = 1):
        gp = GasParams(gas_mass_kg, gas_temperature_K, gas_density_m3)
        check(lib().pic_MCCBundleInit(C.byref(gp), electron_species, ion_species))

    def mcc_table(self, isp: int, nreact: int, nsigma: int = 10001):
        a = np.zeros(nreact * nsigma)
        sm, cr = C.c_double(), C.c_double()
        check(lib().pic_mcc_get_table(isp, _d(a), a.size, C.byref(sm), C.byref(cr)))
        return a, sm.value, cr.value

    def MCC(self):
        check(lib().pic_MCC())

    def set_injection(self, isp: int, density_m3: float, temp_eV: float, drift_ms=(0.0, 0.0, 0.0),
                      r_start: float = 0.0, y_lo: float = 0.0, y_hi: float = 1.0):
        """Wall-injection source of species isp (pic.inp injection block); density 0 removes it."""
        p = Injection(density_m3, temp_eV, (C.c_double * 3)(*drift_ms), r_start, y_lo, y_hi)
        check(lib().pic_set_injection(isp, C.byref(p)))

    def InjectBeams_2D(self, it: int, delta: int = 0) -> list:
        """One injection of every registered source; returns the particles added per species."""
        added = (C.c_int64 * 4)()
        check(lib().pic_InjectBeams_2D(it, delta, added))
        return list(added)

    def set_run_params(self, sort_every: int = 20, seed: int = 12345, mcc_enabled: int = 1):
        check(lib().pic_set_run_params(sort_every, seed, mcc_enabled))

    def sort_particles(self):
        check(lib().pic_sort_particles())

    def sort_group(self) -> int:
        """Slots per group of the sort's transposed layout (256 or 512)."""
        n = C.c_int32()
        check(lib().pic_sort_group(C.byref(n)))
        return n.value

    def initial_field(self):
        check(lib().pic_initial_field())

    def step(self, n: int = 1):
        check(lib().pic_step(n))

    def get_step(self) -> int:
        s = C.c_int64()
        check(lib().pic_get_step(C.byref(s)))
        return s.value

    def set_step(self, s: int):
        check(lib().pic_set_step(s))

    def phase_times(self) -> np.ndarray:
        a = np.zeros(8)
        check(lib().pic_get_phase_times(_d(a), 8))
        return a

    def kernel_stats(self) -> dict:
        """HIP-event times of the CG iteration kernel and the fused push kernel over the
        last step() call."""
        a = np.zeros(12)
        check(lib().pic_get_phase_times(_d(a), 12))
        return {"cg_ms": a[8], "cg_launches": int(a[9]), "push_ms": a[10], "push_launches": int(a[11])}

    def Get_ParticleEField_PPR(self, x, y, EleIndex):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        e = np.ascontiguousarray(EleIndex, dtype=np.int32)
        ex, ey = np.zeros_like(x), np.zeros_like(x)
        check(lib().pic_Get_ParticleEField_PPR(len(x), _d(x), _d(y), _i(e), _d(ex), _d(ey)))
        return ex, ey

    def InjectPositioning(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(len(x), dtype=np.int32)
        check(lib().pic_InjectPositioning(len(x), _d(x), _d(y), _i(out)))
        return out

    def locate(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(len(x), dtype=np.int32)
        check(lib().pic_locate(len(x), _d(x), _d(y), _i(out)))
        return out

    def ParticlePositioning(self, x, y, n_element):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        ne = np.array(n_element, dtype=np.int32)
        tf = np.zeros(len(x), dtype=np.int32)
        check(lib().pic_ParticlePositioning(len(x), _d(x), _d(y), _i(ne), _i(tf)))
        return ne, tf

    def generate_Elementmap(self, repeat_refinement: int = 0):
        d = self.dom
        em = np.zeros((d.nnx - 1) * (d.nny - 1), dtype=np.int32)
        check(lib().pic_generate_Elementmap(_i(em), repeat_refinement))
        return em.reshape((d.nnx - 1, d.nny - 1), order="F")

    # ------------------------------------------------------- learned preconditioner
    def learn_setup(self, max_levels: int = 12, channels=(16, 16), seed: int = 1) -> None:
        """Multigrid network for the current domain; len(channels) finest levels get learned
        smoothers with channels[l] channels."""
        ch = np.ascontiguousarray(channels, dtype=np.int32)
        check(lib().pic_learn_setup(max_levels, _i(ch) if len(ch) else None, len(ch), seed))

    def learn_harvest(self, enable: bool = True, capacity: int = 256, every: int = 1) -> None:
        check(lib().pic_learn_harvest(int(enable), capacity, every))

    def learn_train(self, steps: int, batch: int = 4, lr: float = 1e-3, seed: int = 1) -> tuple:
        a, b = C.c_double(), C.c_double()
        check(lib().pic_learn_train(steps, batch, lr, seed, C.byref(a), C.byref(b)))
        return a.value, b.value

    def learn_eval(self, npairs: int) -> float:
        v = C.c_double()
        check(lib().pic_learn_eval(npairs, C.byref(v)))
        return v.value

    def learn_info(self) -> dict:
        ds, it, fb, loss = C.c_int32(), C.c_int32(), C.c_int64(), C.c_double()
        check(lib().pic_learn_info(C.byref(ds), C.byref(it), C.byref(fb), C.byref(loss)))
        return {"dataset": ds.value, "last_fcg_iterations": it.value, "fallbacks": fb.value, "last_loss": loss.value}

    def learn_data_mix(self, harvested_share: float) -> None:
        """Share of harvested pairs in each training minibatch (the rest synthetic)."""
        check(lib().pic_learn_data_mix(harvested_share))

    def learn_finetune(self, every: int, train_steps: int = 100, batch: int = 4, lr: float = 1e-3, seed: int = 1):
        """Online fine-tuning inside step(): every `every` steps (0 = off) the solve harvests
        plain-CG pairs and the network trains train_steps steps on them."""
        check(lib().pic_learn_finetune(every, train_steps, batch, lr, seed))

    def learn_finetune_info(self) -> dict:
        n, ms, loss = C.c_int64(), C.c_double(), C.c_double()
        check(lib().pic_learn_finetune_info(C.byref(n), C.byref(ms), C.byref(loss)))
        return {"count": n.value, "train_ms": ms.value, "last_loss": loss.value}

    def learn_apply(self, r) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.zeros_like(r)
        check(lib().pic_learn_apply(_f(r), _f(z), len(r)))
        return z

    def learn_params(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().pic_learn_get_params(None, 0, C.byref(n)))
        p = np.zeros(n.value, dtype=np.float32)
        check(lib().pic_learn_get_params(_f(p), len(p), C.byref(n)))
        return p

    def learn_set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().pic_learn_set_params(_f(p), len(p)))

    def learn_layout(self) -> dict:
        L = np.zeros(2, np.int32)
        H = np.zeros(12, np.int32); W = np.zeros(12, np.int32); Cc = np.zeros(12, np.int32)
        off = np.zeros(16, np.int64)
        check(lib().pic_learn_layout(_i(L), _i(H), _i(W), _i(Cc), off.ctypes.data_as(_i64p)))
        n, ls = int(L[0]), int(L[1])
        return {"levels": n, "learned": ls, "H": H[:n].tolist(), "W": W[:n].tolist(), "C": Cc[:ls].tolist(),
                "off_in": [int(o) for o in off[:ls]], "off_q": [int(o) for o in off[ls:2 * ls]],
                "nparams": int(off[2 * ls])}

    def learn_set_pairs(self, r, e) -> None:
        r = np.ascontiguousarray(r, dtype=np.float32)
        e = np.ascontiguousarray(e, dtype=np.float32)
        check(lib().pic_learn_set_pairs(r.shape[0], _f(r), _f(e)))

    def learn_get_pair(self, k: int, n: int) -> tuple:
        r = np.zeros(n, np.float32); e = np.zeros(n, np.float32); ea = C.c_double()
        check(lib().pic_learn_get_pair(k, _f(r), _f(e), C.byref(ea)))
        return r, e, ea.value

    def learn_gradient(self, idx, nparams: int) -> tuple:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.zeros(nparams, np.float32)
        loss = C.c_double()
        check(lib().pic_learn_gradient(len(idx), _i(idx), C.byref(loss), _f(g), nparams))
        return loss.value, g

    # ------------------------------------------------------- learned warm start
    def warm_setup(self, levels: int = 6, channels: int = 16, seed: int = 1) -> None:
        """Encoder-decoder warm start of the solve (after learn_setup: it trains on that ring)."""
        check(lib().pic_warm_setup(levels, channels, seed))

    def warm_train(self, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 1) -> tuple:
        a, b = C.c_double(), C.c_double()
        check(lib().pic_warm_train(steps, batch, lr, seed, C.byref(a), C.byref(b)))
        return a.value, b.value

    def warm_enable(self, enable=True) -> None:
        """False/0 off; True/1 on (one captured HIP graph per application); 2 on, eager launches."""
        check(lib().pic_warm_enable(int(enable)))

    def warm_eval(self, npairs: int) -> tuple:
        a, b = C.c_double(), C.c_double()
        check(lib().pic_warm_eval(npairs, C.byref(a), C.byref(b)))
        return a.value, b.value

    def warm_apply(self, r) -> np.ndarray:
        r = np.ascontiguousarray(r, dtype=np.float32)
        y = np.zeros_like(r)
        check(lib().pic_warm_apply(_f(r), _f(y), len(r)))
        return y

    def warm_params(self) -> np.ndarray:
        n = C.c_int64()
        check(lib().pic_warm_get_params(None, 0, C.byref(n)))
        p = np.zeros(n.value, dtype=np.float32)
        check(lib().pic_warm_get_params(_f(p), len(p), C.byref(n)))
        return p

    def warm_set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().pic_warm_set_params(_f(p), len(p)))

    def warm_layout(self) -> dict:
        L = np.zeros(1, np.int32); H = np.zeros(8, np.int32); W = np.zeros(8, np.int32); Cc = np.zeros(1, np.int32)
        check(lib().pic_warm_layout(_i(L), _i(H), _i(W), _i(Cc)))
        n = int(L[0])
        return {"levels": n, "H": H[:n].tolist(), "W": W[:n].tolist(), "C": int(Cc[0])}

    def warm_gradient(self, idx, nparams: int) -> tuple:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        g = np.zeros(nparams, np.float32)
        loss = C.c_double()
        check(lib().pic_warm_gradient(len(idx), _i(idx), C.byref(loss), _f(g), nparams))
        return loss.value, g

    def warm_info(self) -> dict:
        n, en, a, loss, ap = C.c_int32(), C.c_int32(), C.c_double(), C.c_double(), C.c_int64()
        check(lib().pic_warm_info(C.byref(n), C.byref(en), C.byref(a), C.byref(loss), C.byref(ap)))
        return {"initial_pairs": n.value, "enabled": bool(en.value), "last_alpha": a.value, "last_loss": loss.value,
                "applied": ap.value}

    def kernel_stats_full(self) -> dict:
        a = np.zeros(14)
        check(lib().pic_get_phase_times(_d(a), 14))
        return {"cg_ms": a[8], "cg_launches": int(a[9]), "push_ms": a[10], "push_launches": int(a[11]),
                "fcg_ms": a[12], "fcg_iterations": int(a[13])}

    @staticmethod
    def MY_JPCG_Solver(matrix, R, X, M, NZ: int, N: int) -> int:
        """matrix: array of Sparse (SROW row pointers in entries 1..N+1, 1-based);
        R, X updated in place. Returns ITER."""
        it = C.c_int()
        check(lib().pic_MY_JPCG_Solver(matrix, _d(R), _d(X), _d(M), NZ, N, C.byref(it)))
        return it.value

    @staticmethod
    def MY_JPCG_Solver_noalloc(matrix, R, X, M, NZ: int, N: int) -> int:
        """MY_JPCG_Solver staged in the library's fixed device scratch (no allocation;
        N <= 262144, NZ <= 2359296). R, X updated in place. Returns ITER."""
        it = C.c_int()
        check(lib().pic_MY_JPCG_Solver_noalloc(matrix, _d(R), _d(X), _d(M), NZ, N, C.byref(it)))
        return it.value


def sparse_from_csr(rowptr0, col0, val) -> tuple:
    """Build the reference's SPARSE AoS (1-based) from a 0-based CSR."""
    n = len(rowptr0) - 1
    nz = max(len(val), n + 1)
    arr = (Sparse * nz)()
    for k in range(len(val)):
        arr[k].K = float(val[k])
        arr[k].JCOL = int(col0[k]) + 1
    for i in range(n + 1):
        arr[i].SROW = int(rowptr0[i]) + 1
    return arr, nz


def species_weight(dom: Domain, ppc: int, n0: float = 1.0) -> float:
    """Real particles per macro-particle giving density n0*n_ref at ppc per element:
    rho_s = Weight*qs*ppc/(L_ref^2 n_ref h^2) (GetParChg_2D.f90:896); r-z (per element on
    average, the loader following the element volumes): Weight = n0 n_ref V / N with
    V = pi (ymax^2 - ymin^2)(xmax - xmin) L_ref^3 (GetParChg_2D.f90:899-905)."""
    L = dom.normalisation()["L_ref"]
    h = (dom.xmax - dom.xmin) / (dom.nnx - 1)
    if dom.delta == 1:
        nel = (dom.nnx - 1) * (dom.nny - 1)
        vol = np.pi * (dom.ymax ** 2 - dom.ymin ** 2) * (dom.xmax - dom.xmin) * L ** 3
        return n0 * dom.n_ref * vol / (ppc * nel)
    return n0 * dom.n_ref * (L * h) ** 2 / ppc


def format_fortran_e(v: float, w: int = 15, d: int = 6) -> str:
    """Fortran Ew.d text of v as the library's file writers produce it (host only)."""
    out = C.create_string_buffer(w + 1)
    check(lib().pic_format_fortran_e(float(v), w, d, out))
    return out.value.decode()


def build_workload(pl: PicLibrary, name: str, seed: int = 2024, ppc: int | None = None, capacity_factor: float = 1.5,
                   mcc: bool = True, gas_density: float = 3.2e20):
    """Set up a named workload on the device: mesh, e + Ar+ Maxwellians, MCC tables."""
    w = workload(name)
    dom = w["domain"]
    ppc = ppc or w["ppc"]
    pl.setup(dom, bits=w["bits"], solver=w["solver"], tol=w["tol"])
    nel = (dom.nnx - 1) * (dom.nny - 1)
    cap = int(nel * ppc * capacity_factor) + 1024
    weight = species_weight(dom, ppc)
    pl.set_species(0, ELECTRON["charge_C"], ELECTRON["mass_kg"], weight, w["Te"] * EVTOK, cap)
    pl.set_species(1, ARGON_ION["charge_C"], ARGON_ION["mass_kg"], weight, w["Ti"] * EVTOK, cap)
    pl.load_maxwellian(0, ppc, w["Te"], seed)
    pl.load_maxwellian(1, ppc, w["Ti"], seed + 1)
    if mcc:
        pl.MCCBundleInit(ARGON_ION["mass_kg"], 300.0, gas_density, 0, 1)
    if w.get("inject_drift_vth"):
        # InjectBeams_2D source at x = xmin over the whole wall, at the plasma density
        v_ref = dom.normalisation()["v_ref"]
        for isp, (sp, temp) in enumerate(((ELECTRON, w["Te"]), (ARGON_ION, w["Ti"]))):
            vt = np.sqrt(temp / dom.T_ref_eV / (sp["mass_kg"] / ELECTRON_MASS))
            pl.set_injection(isp, dom.n_ref, temp, (w["inject_drift_vth"] * vt * v_ref, 0.0, 0.0),
                             r_start=dom.xmin, y_lo=dom.ymin, y_hi=dom.ymax)
    pl.set_run_params(20, seed, 1 if mcc else 0)
    return w
