"""Python binding of the MI355X coupled-EHD-step library (``include/ehd.h``).

A thin ctypes layer over ``lib/libehd_amd.so``: every call goes through the
C-ABI, nothing here computes.  There is no fallback: if the shared library is
missing the import of :func:`lib` raises, so a test or benchmark can never
silently run anything but the HIP kernels.

Field arrays use the reference layout (LBM.cu:21-34): scalars ``[z][y][x]``
(numpy shape ``(nz, ny, nx)``), moving populations ``[d-1][z][y][x]``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libehd_amd.so")

FP64, FP32S = 0, 1
INIT_REFERENCE, INIT_HYDROSTATIC, INIT_FIELDS = 0, 1, 2
PERTURB_NOISE, PERTURB_ROLLS, PERTURB_SQUARES, PERTURB_CROSS, PERTURB_HEXAGONS = 0, 1, 2, 3, 4
ERRORS = {0: "EHD_OK", -1: "EHD_EINVAL", -2: "EHD_EHIP", -3: "EHD_ENOMEM", -4: "EHD_ESTATE",
          -5: "EHD_ENONFINITE"}


class EhdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("T", C.c_double), ("C", C.c_double), ("M", C.c_double), ("Fe", C.c_double),
        ("Pr", C.c_double), ("lambda_f", C.c_double), ("lambda_c", C.c_double),
        ("cfl", C.c_double), ("voltage", C.c_double), ("eps", C.c_double), ("rho0", C.c_double),
        ("lz", C.c_double), ("uw", C.c_double), ("exf", C.c_double),
        ("precision", C.c_int), ("init_mode", C.c_int),
        ("seed", C.c_uint64), ("perturb_amp", C.c_double), ("hydro_density", C.c_int),
        ("perturb_pattern", C.c_int),
    ]


class Derived(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dx", "dy", "dz", "dt", "lx", "ly", "cs_square", "K", "nu", "diffu", "charge0",
        "tau_plus", "tau_minus", "tau_c_plus", "tau_c_minus")]


class Diag(C.Structure):
    _fields_ = [("current", C.c_double), ("total_charge", C.c_double), ("umax", C.c_double),
                ("nusselt", C.c_double), ("mean_rho", C.c_double), ("step", C.c_longlong),
                ("time", C.c_double), ("nonfinite", C.c_longlong)]


class Profile(C.Structure):
    _fields_ = [("lbm_ms", C.c_double), ("lbm_launches", C.c_longlong),
                ("poisson_ms", C.c_double), ("poisson_calls", C.c_longlong),
                ("step_ms", C.c_double), ("steps", C.c_longlong)]


class Steady(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("converged", C.c_int), ("failed", C.c_int), ("nusselt", C.c_double),
                ("current", C.c_double), ("delta", C.c_double)]


class _TuneGroup(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nboxes", C.c_int), ("x_lines", C.c_int),
                ("y_lines", C.c_int), ("ms", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("lbm_grid", C.c_int), ("lbm_tiles", C.c_int), ("lbm_ms", C.c_double), ("ngroups", C.c_int),
                ("group", _TuneGroup * 16), ("lbm_chunk", C.c_int), ("chains", C.c_int),
                ("step_ms", C.c_double * 2)]


BATCH_GRAPH, BATCH_TUNE, BATCH_CHAINS = 1, 2, 4

EXPORTS = (
    "ehd_params_default", "ehd_derive", "ehd_compute_parameters", "ehd_create", "ehd_destroy",
    "ehd_init", "ehd_step", "ehd_lbm_step", "ehd_poisson", "ehd_read_fields", "ehd_write_phi",
    "ehd_write_charge", "ehd_read_populations", "ehd_write_populations", "ehd_diagnostics",
    "ehd_set_nusselt_reference", "ehd_profile_enable", "ehd_profile_read", "ehd_profile_reset", "ehd_set_debug_sync",
    "ehd_sync", "ehd_last_error", "ehd_abi_version", "ehd_save_state", "ehd_load_state",
    "ehd_hydrostatic_current", "ehd_run_to_steady", "ehd_hydrostatic_reference",
    "ehd_init_consistent", "ehd_batch_init_consistent",
    "ehd_batch_create", "ehd_batch_destroy", "ehd_batch_init", "ehd_batch_step", "ehd_batch_diagnostics",
    "ehd_batch_run_to_steady", "ehd_batch_box", "ehd_batch_tuning", "ehd_copy_ceiling",
    "ehd_poisson_noalloc",
    # host-only file layer, include/ehd_io.h
    "ehd_io_tecplot", "ehd_io_data_end", "ehd_io_read_end", "ehd_io_dmd", "ehd_io_umax", "ehd_io_current",
)

_lib = None


def lib() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.ehd_params_default.argtypes = [C.POINTER(Params)]
    L.ehd_derive.argtypes = [C.POINTER(Params), C.POINTER(Derived)]
    L.ehd_compute_parameters.argtypes = [C.POINTER(Derived), C.POINTER(Params), dp, dp, dp, dp]
    L.ehd_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.ehd_destroy.argtypes = [vp]
    L.ehd_destroy.restype = None
    L.ehd_init.argtypes = [vp, dp]
    L.ehd_init_consistent.argtypes = [vp, dp, C.c_int]
    L.ehd_step.argtypes = [vp, C.c_int]
    L.ehd_lbm_step.argtypes = [vp]
    L.ehd_poisson.argtypes = [vp]
    L.ehd_read_fields.argtypes = [vp] + [dp] * 9
    L.ehd_write_phi.argtypes = [vp, dp]
    L.ehd_write_charge.argtypes = [vp, dp]
    L.ehd_read_populations.argtypes = [vp, dp, dp, dp, dp]
    L.ehd_write_populations.argtypes = [vp, dp, dp, dp, dp]
    L.ehd_diagnostics.argtypes = [vp, C.POINTER(Diag)]
    L.ehd_set_nusselt_reference.argtypes = [vp, C.c_double]
    L.ehd_profile_enable.argtypes = [vp, C.c_int]
    L.ehd_set_debug_sync.argtypes = [vp, C.c_int]
    L.ehd_profile_read.argtypes = [vp, C.POINTER(Profile)]
    L.ehd_profile_reset.argtypes = [vp]
    L.ehd_sync.argtypes = [vp]
    L.ehd_hydrostatic_current.argtypes = [C.POINTER(Params), dp]
    L.ehd_run_to_steady.argtypes = [vp, C.c_int, C.c_double, C.c_longlong, C.POINTER(Steady)]
    L.ehd_hydrostatic_reference.argtypes = [vp, C.c_int, C.c_double, C.c_longlong, dp]
    L.ehd_save_state.argtypes = [vp, C.c_char_p]
    L.ehd_load_state.argtypes = [vp, C.c_char_p]
    cs, i_ = C.c_char_p, C.c_int
    L.ehd_io_tecplot.argtypes = [cs, i_, i_, C.c_double, i_, i_, i_, C.c_double, C.c_double, C.c_double, dp]
    L.ehd_io_data_end.argtypes = [cs, C.c_double, i_, i_, i_, dp]
    L.ehd_io_read_end.argtypes = [cs, i_, i_, i_, dp, dp]
    L.ehd_io_dmd.argtypes = [cs, C.c_double, i_, i_, i_, dp]
    L.ehd_io_umax.argtypes = [cs, i_, C.c_double, i_, i_, i_, dp]
    L.ehd_io_current.argtypes = [i_, i_, i_, C.c_double, C.c_double, dp, dp]
    L.ehd_batch_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.POINTER(vp)]
    L.ehd_batch_destroy.argtypes = [vp]
    L.ehd_batch_destroy.restype = None
    L.ehd_batch_init.argtypes = [vp, C.POINTER(dp)]
    L.ehd_batch_init_consistent.argtypes = [vp, C.POINTER(dp), C.c_int]
    L.ehd_batch_step.argtypes = [vp, C.c_int]
    L.ehd_batch_diagnostics.argtypes = [vp, C.POINTER(Diag)]
    L.ehd_batch_run_to_steady.argtypes = [vp, C.c_int, C.c_double, C.c_longlong, C.POINTER(Steady)]
    L.ehd_batch_box.argtypes = [vp, C.c_int]
    L.ehd_batch_box.restype = vp
    L.ehd_batch_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.ehd_last_error.argtypes = []
    L.ehd_last_error.restype = C.c_char_p
    L.ehd_abi_version.argtypes = []
    L.ehd_copy_ceiling.argtypes = [C.c_size_t, C.c_int, dp]
    L.ehd_poisson_noalloc.argtypes = [C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_void_p]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("ehd_destroy", "ehd_last_error", "ehd_io_current", "ehd_batch_destroy", "ehd_batch_box"):
            f.restype = C.c_int
    L.ehd_io_current.restype = C.c_double
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().ehd_last_error().decode(errors="replace")
        raise EhdError(rc, msg)


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params() -> Params:
    p = Params()
    _check(lib().ehd_params_default(C.byref(p)))
    return p


def make_params(nx, ny, nz, T=200.0, C_=10.0, M=10.0, Fe=4000.0, Pr=10.0, lam=0.25, lam_c=1e-6,
                precision=FP64, init_mode=INIT_HYDROSTATIC, seed=1, amp=1e-3,
                cfl=0.01, voltage=1.0e4, eps=1.0e-4, rho0=1600.0, lz=1.0, uw=0.0, exf=0.0,
                hydro_density=0, pattern=PERTURB_NOISE) -> Params:
    p = default_params()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.T, p.C, p.M, p.Fe, p.Pr = T, C_, M, Fe, Pr
    p.lambda_f, p.lambda_c = lam, lam_c
    p.cfl, p.voltage, p.eps, p.rho0, p.lz, p.uw, p.exf = cfl, voltage, eps, rho0, lz, uw, exf
    p.precision, p.init_mode, p.seed, p.perturb_amp = precision, init_mode, seed, amp
    p.hydro_density = hydro_density
    p.perturb_pattern = pattern
    return p


def hydrostatic_current(p: Params) -> float:
    v = C.c_double()
    _check(lib().ehd_hydrostatic_current(C.byref(p), C.byref(v)))
    return v.value


def copy_ceiling(nbytes: int = 1 << 31, reps: int = 20):
    """GB/s of a plain and a nontemporal float4 device-to-device copy of
    nbytes (read + written), the practical HBM ceiling of this GPU."""
    out = np.zeros(2)
    _check(lib().ehd_copy_ceiling(int(nbytes), int(reps), _dp(out)))
    return float(out[0]), float(out[1])


def derive(p: Params) -> Derived:
    d = Derived()
    _check(lib().ehd_derive(C.byref(p), C.byref(d)))
    return d


def compute_parameters(d: Derived, p: Params):
    T, M, Cc, Fe = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    _check(lib().ehd_compute_parameters(C.byref(d), C.byref(p), C.byref(T), C.byref(M), C.byref(Cc), C.byref(Fe)))
    return T.value, M.value, Cc.value, Fe.value


class Box:
    """One simulation box on the GPU (owns every device buffer)."""

    def __init__(self, params: Params, handle=None):
        self.params = params
        self.shape = (params.nz, params.ny, params.nx)
        self.S = params.nx * params.ny * params.nz
        self._owned = handle is None
        if handle is None:
            h = C.c_void_p()
            _check(lib().ehd_create(C.byref(params), C.byref(h)))
            handle = h
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().ehd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self, fields9: np.ndarray | None = None):
        a = None
        if fields9 is not None:
            a = np.ascontiguousarray(fields9, dtype=np.float64)
            if a.size != 9 * self.S:
                raise ValueError("fields9 must hold 9*S doubles")
        _check(lib().ehd_init(self._h, _dp(a)))

    def init_consistent(self, fields9: np.ndarray, iterations: int):
        a = np.ascontiguousarray(fields9, dtype=np.float64)
        if a.size != 9 * self.S:
            raise ValueError("fields9 must hold 9*S doubles")
        _check(lib().ehd_init_consistent(self._h, _dp(a), int(iterations)))

    def step(self, n: int = 1):
        _check(lib().ehd_step(self._h, int(n)))

    def lbm_step(self):
        _check(lib().ehd_lbm_step(self._h))

    def poisson(self):
        _check(lib().ehd_poisson(self._h))

    def read_fields(self, names=("rho", "q", "phi", "ux", "uy", "uz", "ex", "ey", "ez")) -> dict:
        order = ("rho", "q", "phi", "ux", "uy", "uz", "ex", "ey", "ez")
        out = {n: np.empty(self.shape) for n in names}
        _check(lib().ehd_read_fields(self._h, *[_dp(out.get(n)) for n in order]))
        return out

    def write_phi(self, phi: np.ndarray):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        _check(lib().ehd_write_phi(self._h, _dp(a)))

    def write_charge(self, q: np.ndarray):
        a = np.ascontiguousarray(q, dtype=np.float64)
        _check(lib().ehd_write_charge(self._h, _dp(a)))

    def read_populations(self):
        nz, ny, nx = self.shape
        f0 = np.empty((nz, ny, nx)); h0 = np.empty((nz, ny, nx))
        f1 = np.empty((26, nz, ny, nx)); h1 = np.empty((26, nz, ny, nx))
        _check(lib().ehd_read_populations(self._h, _dp(f0), _dp(f1), _dp(h0), _dp(h1)))
        return f0, f1, h0, h1

    def write_populations(self, f0, f1, h0, h1):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (f0, f1, h0, h1)]
        _check(lib().ehd_write_populations(self._h, *[_dp(a) for a in arrs]))

    def diagnostics(self, allow_nonfinite: bool = False) -> Diag:
        """FP64 reductions; the watchdog's EHD_ENONFINITE raises unless
        allow_nonfinite (then d.nonfinite > 0 tells)."""
        d = Diag()
        rc = lib().ehd_diagnostics(self._h, C.byref(d))
        if not (allow_nonfinite and rc == -5):
            _check(rc)
        return d

    def set_nusselt_reference(self, current_hs: float):
        _check(lib().ehd_set_nusselt_reference(self._h, float(current_hs)))

    def set_debug_sync(self, on: bool = True):
        """Synchronise and check after every launch (faults named by phase and step)."""
        _check(lib().ehd_set_debug_sync(self._h, int(on)))

    def profile_enable(self, on: bool = True, every: int = 1):
        """Events around the LBM and Poisson launches of every ``every``-th step."""
        _check(lib().ehd_profile_enable(self._h, int(every) if on else 0))

    def profile_read(self) -> Profile:
        pr = Profile()
        _check(lib().ehd_profile_read(self._h, C.byref(pr)))
        return pr

    def profile_reset(self):
        _check(lib().ehd_profile_reset(self._h))

    def sync(self):
        _check(lib().ehd_sync(self._h))

    def hydrostatic_reference(self, check_every: int = 1000, tol: float = 1e-9, max_steps: int = 10**6) -> float:
        v = C.c_double()
        _check(lib().ehd_hydrostatic_reference(self._h, int(check_every), float(tol), int(max_steps), C.byref(v)))
        return v.value

    def run_to_steady(self, check_every: int = 1000, tol: float = 1e-6, max_steps: int = 10**6,
                      allow_nonfinite: bool = False) -> Steady:
        """Steps to |dNu| < tol per chunk; a box the watchdog flags raises
        EHD_ENONFINITE unless allow_nonfinite (then s.failed = 1)."""
        s = Steady()
        rc = lib().ehd_run_to_steady(self._h, int(check_every), float(tol), int(max_steps), C.byref(s))
        if not (allow_nonfinite and rc == -5):
            _check(rc)
        return s

    def save_state(self, path: str):
        _check(lib().ehd_save_state(self._h, os.fsencode(path)))

    def load_state(self, path: str):
        _check(lib().ehd_load_state(self._h, os.fsencode(path)))


# ---- host-only file layer (include/ehd_io.h): the driver's output files ----

def _f9(fields9, S):
    a = np.ascontiguousarray(fields9, dtype=np.float64)
    if a.size != 9 * S:
        raise ValueError("fields9 must hold 9*S doubles")
    return a


def write_tecplot(path, t, fields9, shape, spacing, first=False, append=False):
    nz, ny, nx = shape
    a = _f9(fields9, nx * ny * nz)
    _check(lib().ehd_io_tecplot(os.fsencode(path), int(append), int(first), float(t), nx, ny, nz,
                                *map(float, spacing), _dp(a)))


def write_data_end(path, t, fields9, shape):
    nz, ny, nx = shape
    a = _f9(fields9, nx * ny * nz)
    _check(lib().ehd_io_data_end(os.fsencode(path), float(t), nx, ny, nz, _dp(a)))


def read_data_end(path, shape):
    nz, ny, nx = shape
    out = np.empty((9, nz, ny, nx))
    t = C.c_double()
    _check(lib().ehd_io_read_end(os.fsencode(path), nx, ny, nz, _dp(out), C.byref(t)))
    return out, t.value


def write_dmd(path, t, uz, shape):
    nz, ny, nx = shape
    a = np.ascontiguousarray(uz, dtype=np.float64)
    _check(lib().ehd_io_dmd(os.fsencode(path), float(t), nx, ny, nz, _dp(a)))


def write_umax(path, t, uz, shape, append=False):
    nz, ny, nx = shape
    a = np.ascontiguousarray(uz, dtype=np.float64)
    _check(lib().ehd_io_umax(os.fsencode(path), int(append), float(t), nx, ny, nz, _dp(a)))


def host_current(q, ez, K, dy):
    nz, ny, nx = q.shape
    a, b = (np.ascontiguousarray(v, dtype=np.float64) for v in (q, ez))
    return lib().ehd_io_current(nx, ny, nz, float(K), float(dy), _dp(a), _dp(b))


class Batch:
    """Many boxes stepped together (ehd_batch_*): one persistent LBM launch,
    Poisson passes batched per (nx, ny), HIP-graph replay, autotuned shapes."""

    def __init__(self, params_list, flags=BATCH_GRAPH | BATCH_TUNE):
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        h = C.c_void_p()
        _check(lib().ehd_batch_create(arr, len(self.params), int(flags), C.byref(h)))
        self._h = h
        self.boxes = [Box(p, handle=C.c_void_p(lib().ehd_batch_box(h, i))) for i, p in enumerate(self.params)]

    def close(self):
        if getattr(self, "_h", None):
            for b in self.boxes:
                b._h = None
            lib().ehd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.boxes)

    def init(self, fields=None):
        if fields is None:
            _check(lib().ehd_batch_init(self._h, None))
            return
        keep = [None if f is None else np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        ptrs = (C.POINTER(C.c_double) * len(keep))(*[_dp(a) if a is not None else None for a in keep])
        _check(lib().ehd_batch_init(self._h, ptrs))

    def init_consistent(self, fields, iterations: int):
        keep = [None if f is None else np.ascontiguousarray(f, dtype=np.float64) for f in fields]
        ptrs = (C.POINTER(C.c_double) * len(keep))(*[_dp(a) if a is not None else None for a in keep])
        _check(lib().ehd_batch_init_consistent(self._h, ptrs, int(iterations)))

    def step(self, n: int = 1):
        _check(lib().ehd_batch_step(self._h, int(n)))

    def diagnostics(self, allow_nonfinite: bool = False):
        """Diagnostics of every box; with allow_nonfinite the boxes the
        watchdog flags come back with nonfinite > 0 instead of raising."""
        out = (Diag * len(self.boxes))()
        rc = lib().ehd_batch_diagnostics(self._h, out)
        if not (allow_nonfinite and rc == -5):
            _check(rc)
        return list(out)

    def run_to_steady(self, check_every: int = 1000, tol: float = 1e-6, max_steps: int = 10**6):
        out = (Steady * len(self.boxes))()
        _check(lib().ehd_batch_run_to_steady(self._h, int(check_every), float(tol), int(max_steps), out))
        return list(out)

    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().ehd_batch_tuning(self._h, C.byref(t)))
        return t
