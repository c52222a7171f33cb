"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module.  It loads ``oracle/_build/libehd_oracle.so`` (built by
``oracle/Makefile``) and exposes the reference's step phases on numpy views of
the oracle's buffers, in the reference layout (scalars ``[z][y][x]``, moving
populations ``[d-1][z][y][x]``, LBM.cu:21-34).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libehd_oracle.so")

F0, F1, F2, H0, H1, H2, RHO, Q, PHI, UX, UY, UZ, EX, EY, EZ, F0BC, SEXT, FREQ, PHIEXT, KX, KY, KZ = range(22)


class OrefParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ne", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("cfl", C.c_double), ("dt", C.c_double), ("cs_square", C.c_double),
        ("rho0", C.c_double), ("charge0", C.c_double), ("voltage", C.c_double),
        ("eps", C.c_double), ("diffu", C.c_double), ("nu", C.c_double), ("K", C.c_double),
        ("uw", C.c_double), ("exf", C.c_double),
        ("V", C.c_double), ("VC", C.c_double),
        ("w0", C.c_double), ("ws", C.c_double), ("wa", C.c_double), ("wd", C.c_double),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        L.oref_params_init.argtypes = [C.POINTER(OrefParams), C.c_int, C.c_int, C.c_int,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.oref_params_init.restype = C.c_int
        L.oref_compute_parameters.argtypes = [C.POINTER(OrefParams), dp, dp, dp, dp]
        L.oref_create.argtypes = [C.POINTER(OrefParams)]
        L.oref_create.restype = vp
        L.oref_destroy.argtypes = [vp]
        L.oref_field.argtypes = [vp, C.c_int]
        L.oref_field.restype = dp
        for name in ("oref_initialization", "oref_init_equilibrium", "oref_collide_save",
                     "oref_boundary", "oref_stream", "oref_bc_charge", "oref_stream_collide_save",
                     "oref_extension", "oref_derivative", "oref_extract", "oref_efield",
                     "oref_poisson", "oref_poisson_tridiag"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.oref_set_fields.argtypes = [vp, dp]
        L.oref_init_hydrostatic.argtypes = [vp, C.c_uint64, C.c_double, C.c_int]
        L.oref_init_hydrostatic_pattern.argtypes = [vp, C.c_uint64, C.c_double, C.c_int, C.c_int]
        L.oref_init_hydrostatic_pattern.restype = None
        L.oref_pattern_velocity.argtypes = [C.c_int] + [C.c_double] * 7 + [dp]
        L.oref_pattern_velocity.restype = None
        L.oref_fft3d.argtypes = [vp, dp, dp, C.c_int]
        L.oref_step.argtypes = [vp, C.c_int]
        L.oref_step_tridiag.argtypes = [vp, C.c_int]
        L.oref_current.argtypes = [vp]
        L.oref_current.restype = C.c_double
        L.oref_fft1d.argtypes = [dp, C.c_int, C.c_int]
        L.oref_hash_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.oref_hash_uniform.restype = C.c_double
        L.oref_hydrostatic_e0.argtypes = [C.c_double]
        L.oref_hydrostatic_e0.restype = C.c_double
        L.oref_tables.argtypes = [ip, ip, dp, ip, ip, ip]
        L.oref_set_threads.argtypes = [C.c_int]
        L.oref_node_terms.argtypes = [C.POINTER(OrefParams)] + [C.c_double] * 11 + [dp, dp, dp]
        L.oref_rates.argtypes = [vp, dp]
        L.oref_set_rates.argtypes = [vp, dp]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().oref_set_threads(int(n))


def make_params(nx, ny, nz, T=200.0, M=10.0, C_=10.0, Fe=4000.0, lam=0.25, lam_c=1e-6,
                cfl=0.01, voltage=1.0e4, eps=1.0e-4, rho0=1600.0, lz=1.0, uw=0.0, exf=0.0) -> OrefParams:
    p = OrefParams()
    rc = lib().oref_params_init(C.byref(p), nx, ny, nz, T, M, C_, Fe, lam, lam_c,
                                cfl, voltage, eps, rho0, lz, uw, exf)
    if rc != 0:
        raise ValueError("invalid oracle parameters")
    return p


def compute_parameters(p: OrefParams):
    T, M, Cc, Fe = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    lib().oref_compute_parameters(C.byref(p), C.byref(T), C.byref(M), C.byref(Cc), C.byref(Fe))
    return T.value, M.value, Cc.value, Fe.value


def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().oref_fft1d(a.ctypes.data_as(C.POINTER(C.c_double)), a.size, int(sign))
    return a


def hash_uniform(seed, x, y, z, comp) -> float:
    return lib().oref_hash_uniform(seed, x, y, z, comp)


def hydrostatic_e0(Cc: float) -> float:
    return lib().oref_hydrostatic_e0(Cc)


def node_terms(p: OrefParams, rho, ux, uy, uz, q, Ex, Ey, Ez, Fx, Fy, Fz):
    fe = np.zeros(27); he = np.zeros(27); fp = np.zeros(27)
    dp = C.POINTER(C.c_double)
    lib().oref_node_terms(C.byref(p), rho, ux, uy, uz, q, Ex, Ey, Ez, Fx, Fy, Fz,
                          fe.ctypes.data_as(dp), he.ctypes.data_as(dp), fp.ctypes.data_as(dp))
    return fe, he, fp


def tables():
    c = np.zeros(81, np.int32); opp = np.zeros(27, np.int32); w = np.zeros(27)
    eqf = np.zeros(108, np.int32); eqh = np.zeros(108, np.int32); mom = np.zeros(54, np.int32)
    ip = C.POINTER(C.c_int)
    lib().oref_tables(c.ctypes.data_as(ip), opp.ctypes.data_as(ip), w.ctypes.data_as(C.POINTER(C.c_double)),
                      eqf.ctypes.data_as(ip), eqh.ctypes.data_as(ip), mom.ctypes.data_as(ip))
    return dict(c=c.reshape(27, 3), opp=opp, w=w, eqf=eqf.reshape(27, 4), eqh=eqh.reshape(27, 4),
                mom=mom.reshape(3, 18))


class Box:
    """One oracle box.  Field views are numpy arrays over the oracle's memory."""

    def __init__(self, params: OrefParams):
        self.p = params
        self._b = lib().oref_create(C.byref(self.p))
        if not self._b:
            raise MemoryError("oref_create failed")
        nx, ny, nz, ne = params.nx, params.ny, params.nz, params.ne
        self.shape = (nz, ny, nx)
        S = nx * ny * nz
        SE = nx * ny * ne
        sizes = {F0: (nz, ny, nx), H0: (nz, ny, nx), F1: (26, nz, ny, nx), F2: (26, nz, ny, nx),
                 H1: (26, nz, ny, nx), H2: (26, nz, ny, nx), F0BC: (2, ny, nx),
                 SEXT: (ne, ny, nx, 2), FREQ: (ne, ny, nx, 2), PHIEXT: (ne, ny, nx, 2),
                 KX: (nx,), KY: (ny,), KZ: (ne,)}
        self._views = {}
        for which in range(22):
            shp = sizes.get(which, (nz, ny, nx))
            ptr = lib().oref_field(self._b, which)
            n = int(np.prod(shp))
            self._views[which] = np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shp)
        del S, SE

    def __del__(self):
        b = getattr(self, "_b", None)
        if b:
            lib().oref_destroy(b)
            self._b = None

    def __getitem__(self, which: int) -> np.ndarray:
        return self._views[which]

    # phases
    def initialization(self): lib().oref_initialization(self._b)
    def init_equilibrium(self): lib().oref_init_equilibrium(self._b)
    def init_hydrostatic(self, seed: int, amp: float, hydro_density: int = 0, pattern: int = 0):
        lib().oref_init_hydrostatic_pattern(self._b, seed, amp, int(hydro_density), int(pattern))

    def set_fields(self, f9: np.ndarray):
        a = np.ascontiguousarray(f9, dtype=np.float64)
        assert a.size == 9 * int(np.prod(self.shape))
        lib().oref_set_fields(self._b, a.ctypes.data_as(C.POINTER(C.c_double)))

    def collide_save(self): lib().oref_collide_save(self._b)
    def boundary(self): lib().oref_boundary(self._b)
    def stream(self): lib().oref_stream(self._b)
    def bc_charge(self): lib().oref_bc_charge(self._b)
    def stream_collide_save(self): lib().oref_stream_collide_save(self._b)
    def extension(self): lib().oref_extension(self._b)
    def derivative(self): lib().oref_derivative(self._b)
    def extract(self): lib().oref_extract(self._b)
    def efield(self): lib().oref_efield(self._b)
    def poisson(self): lib().oref_poisson(self._b)
    def poisson_tridiag(self): lib().oref_poisson_tridiag(self._b)
    def step(self, n: int = 1): lib().oref_step(self._b, int(n))
    def step_tridiag(self, n: int = 1): lib().oref_step_tridiag(self._b, int(n))
    def current(self) -> float: return lib().oref_current(self._b)

    def rates(self):
        r = np.zeros(4)
        lib().oref_rates(self._b, r.ctypes.data_as(C.POINTER(C.c_double)))
        return r

    def set_rates(self, r) -> None:
        a = np.ascontiguousarray(r, dtype=np.float64)
        assert a.size == 4
        lib().oref_set_rates(self._b, a.ctypes.data_as(C.POINTER(C.c_double)))

    def fft3d(self, src: int, dst: int, sign: int):
        dp = C.POINTER(C.c_double)
        lib().oref_fft3d(self._b, self[src].ctypes.data_as(dp), self[dst].ctypes.data_as(dp), int(sign))


def pattern_velocity(pattern: int, amp: float, xx: float, yy: float, zz: float, lx: float, ly: float,
                     lz: float) -> np.ndarray:
    """(ux, uy, uz) of a perturbation pattern at one physical point."""
    u = np.zeros(3)
    lib().oref_pattern_velocity(int(pattern), amp, xx, yy, zz, lx, ly, lz, u.ctypes.data_as(C.POINTER(C.c_double)))
    return u
