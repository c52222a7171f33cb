"""Python binding of the Fourier-neural-operator surrogate (include/ehd_fno.h).

ctypes only: forward, gradients and Adam steps run in the HIP library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check, lib


class Config(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("X", "Y", "Z", "width", "mx", "my", "mz", "layers", "d_in", "d_hidden",
                                        "d_out", "max_batch")]


class Work(C.Structure):
    _fields_ = [("flops", C.c_double), ("bytes", C.c_double), ("launches", C.c_longlong)]


EXPORTS = ("ehd_fno_create", "ehd_fno_destroy", "ehd_fno_num_params", "ehd_fno_set_params", "ehd_fno_get_params",
           "ehd_fno_forward", "ehd_fno_gradients", "ehd_fno_train_step", "ehd_fno_gemm", "ehd_fno_gemm_strided",
           "ehd_fno_gemm_work", "ehd_fno_set_dataset", "ehd_fno_train_step_indices", "ehd_fno_cgemm_strided")

_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        fp = C.POINTER(C.c_float)
        vp = C.c_void_p
        L.ehd_fno_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.ehd_fno_create.restype = C.c_int
        L.ehd_fno_destroy.argtypes = [vp]
        L.ehd_fno_destroy.restype = None
        L.ehd_fno_num_params.argtypes = [vp]
        L.ehd_fno_num_params.restype = C.c_longlong
        L.ehd_fno_set_params.argtypes = [vp, fp]
        L.ehd_fno_get_params.argtypes = [vp, fp]
        L.ehd_fno_forward.argtypes = [vp, C.c_int, fp, fp]
        L.ehd_fno_gradients.argtypes = [vp, C.c_int, fp, fp, C.POINTER(C.c_double), fp]
        L.ehd_fno_train_step.argtypes = [vp, C.c_int, fp, fp, C.c_float, C.POINTER(C.c_double)]
        L.ehd_fno_gemm.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp]
        llp = C.POINTER(C.c_longlong)
        L.ehd_fno_gemm_strided.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp, llp, C.c_longlong, fp,
                                           llp, C.c_longlong, fp, llp, C.c_longlong, C.c_float, C.c_float]
        L.ehd_fno_gemm_work.argtypes = [vp, C.c_int, C.POINTER(Work)]
        L.ehd_fno_cgemm_strided.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, llp, C.c_longlong, fp, fp, llp,
                                            C.c_longlong, fp, fp, llp, C.c_longlong, C.c_int, C.c_int, C.c_float,
                                            C.c_float]
        L.ehd_fno_set_dataset.argtypes = [vp, C.c_int, fp, fp]
        L.ehd_fno_train_step_indices.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.c_float,
                                                 C.POINTER(C.c_double)]
        for n in ("ehd_fno_set_params", "ehd_fno_get_params", "ehd_fno_forward", "ehd_fno_gradients",
                  "ehd_fno_train_step", "ehd_fno_gemm", "ehd_fno_gemm_strided", "ehd_fno_gemm_work",
                  "ehd_fno_set_dataset", "ehd_fno_train_step_indices", "ehd_fno_cgemm_strided"):
            getattr(L, n).restype = C.c_int
        _bound = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def gemm(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """A @ B through the library's bf16 matrix-core GEMM."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.empty((M, N), np.float32)
    _check(_lib().ehd_fno_gemm(M, N, K, _fp(A), _fp(B), _fp(Cm)))
    return Cm


def gemm_strided(M, N, K, A, a_strides, B, b_strides, C_, c_strides, nb=1, ksplit=0, alpha=1.0, beta=0.0):
    """The library GEMM on flat host arrays with explicit element strides
    (ehd_fno_gemm_strided): a_strides = (m, k, batch), b_strides = (k, n,
    batch), c_strides = (m, n, batch, split-K slice).  C_ is updated in place."""
    A = np.ascontiguousarray(A, dtype=np.float32).reshape(-1)
    B = np.ascontiguousarray(B, dtype=np.float32).reshape(-1)
    assert C_.dtype == np.float32 and C_.flags.c_contiguous
    Cf = C_.reshape(-1)
    arr = lambda v: (C.c_longlong * len(v))(*[int(x) for x in v])
    _check(_lib().ehd_fno_gemm_strided(M, N, K, nb, ksplit, _fp(A), arr(a_strides), A.size, _fp(B), arr(b_strides),
                                       B.size, _fp(Cf), arr(c_strides), Cf.size, float(alpha), float(beta)))
    return C_


def cgemm_strided(M, N, K, Ar, Ai, a_strides, Br, Bi, b_strides, Cr, Ci, c_strides, nb=1, sa=1, sb=1, alpha=1.0,
                  beta=0.0):
    """The library's one-launch complex product (ehd_fno_cgemm_strided) on flat
    host planes: Cr + i Ci = alpha (Ar + i sa Ai)(Br + i sb Bi) + beta (Cr + i Ci);
    Ai, Bi or Ci None: that plane is absent.  Strides as in gemm_strided
    (c_strides = (m, n, batch)); Cr and Ci are updated in place."""
    planes = []
    for X in (Ar, Ai, Br, Bi):
        planes.append(None if X is None else np.ascontiguousarray(X, dtype=np.float32).reshape(-1))
    for X in (Cr, Ci):
        if X is not None:
            assert X.dtype == np.float32 and X.flags.c_contiguous
    ptr = lambda X: None if X is None else _fp(X.reshape(-1))
    arr = lambda v: (C.c_longlong * len(v))(*[int(x) for x in v])
    _check(_lib().ehd_fno_cgemm_strided(M, N, K, nb, ptr(planes[0]), ptr(planes[1]), arr(a_strides), planes[0].size,
                                        ptr(planes[2]), ptr(planes[3]), arr(b_strides), planes[2].size, ptr(Cr),
                                        ptr(Ci), arr(c_strides), Cr.size, int(sa), int(sb), float(alpha),
                                        float(beta)))
    return Cr, Ci


# flat parameter order of include/ehd_fno.h
def param_shapes(cfg: Config):
    w, mx, my, mz = cfg.width, cfg.mx, cfg.my, cfg.mz
    out = [("Pw", (cfg.d_in, w)), ("Pb", (w,))]
    for l in range(cfg.layers):
        out += [(f"W{l}", (w, w)), (f"b{l}", (w,)), (f"Rr{l}", (mz, 2 * my, 2 * mx, w, w)),
                (f"Ri{l}", (mz, 2 * my, 2 * mx, w, w))]
    out += [("Q1w", (w, cfg.d_hidden)), ("Q1b", (cfg.d_hidden,)), ("Q2w", (cfg.d_hidden, cfg.d_out)),
            ("Q2b", (cfg.d_out,))]
    return out


def flatten(cfg: Config, params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[n], dtype=np.float32).reshape(-1) for n, _ in param_shapes(cfg)])


def unflatten(cfg: Config, flat: np.ndarray) -> dict:
    out, o = {}, 0
    for n, s in param_shapes(cfg):
        k = int(np.prod(s))
        out[n] = flat[o:o + k].reshape(s)
        o += k
    return out


class FNO:
    def __init__(self, X, Y, Z, width=32, modes=(12, 12, 12), layers=4, d_in=6, d_hidden=128, d_out=5, max_batch=8):
        mx, my, mz = min(modes[0], X // 2), min(modes[1], Y // 2), min(modes[2], Z // 2)
        self.cfg = Config(X, Y, Z, width, mx, my, mz, layers, d_in, d_hidden, d_out, max_batch)
        h = C.c_void_p()
        _check(_lib().ehd_fno_create(C.byref(self.cfg), C.byref(h)))
        self._h = h
        self.n = _lib().ehd_fno_num_params(h)

    def close(self):
        if getattr(self, "_h", None):
            _lib().ehd_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_params(self, seed: int = 0) -> np.ndarray:
        """Initial weights: pointwise layers N(0, 1/fan_in), biases 0, spectral
        weights uniform in [0, 1/width^2) (the usual FNO initialisation)."""
        rng = np.random.default_rng(seed)
        c = self.cfg
        out = {}
        for name, shape in param_shapes(c):
            if name.startswith(("Rr", "Ri")):
                out[name] = rng.uniform(0.0, 1.0 / (c.width * c.width), shape)
            elif len(shape) == 1:
                out[name] = np.zeros(shape)
            else:
                out[name] = rng.standard_normal(shape) / np.sqrt(shape[0])
        return flatten(c, out)

    def set_params(self, params):
        flat = params if isinstance(params, np.ndarray) else flatten(self.cfg, params)
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        assert flat.size == self.n
        _check(_lib().ehd_fno_set_params(self._h, _fp(flat)))

    def get_params(self) -> np.ndarray:
        flat = np.empty(self.n, np.float32)
        _check(_lib().ehd_fno_get_params(self._h, _fp(flat)))
        return flat

    def forward(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float32)
        B = a.shape[0]
        c = self.cfg
        out = np.empty((B, c.d_out, c.Z, c.Y, c.X), np.float32)
        _check(_lib().ehd_fno_forward(self._h, B, _fp(a), _fp(out)))
        self._resident = 0
        return out

    def gradients(self, a, target):
        a = np.ascontiguousarray(a, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.empty(self.n, np.float32)
        loss = C.c_double()
        _check(_lib().ehd_fno_gradients(self._h, a.shape[0], _fp(a), _fp(t), C.byref(loss), _fp(g)))
        self._resident = a.shape[0]
        return loss.value, g

    def gemm_work(self, reset: bool = False) -> Work:
        """FLOPs, bytes and launches of the GEMMs since the last reset."""
        w = Work()
        _check(_lib().ehd_fno_gemm_work(self._h, int(reset), C.byref(w)))
        return w

    def set_dataset(self, a, target):
        """Copy a training set ([n][d_in][Z][Y][X], [n][d_out][Z][Y][X]) into HBM once."""
        a = np.ascontiguousarray(a, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        assert a.shape[0] == t.shape[0]
        _check(_lib().ehd_fno_set_dataset(self._h, a.shape[0], _fp(a), _fp(t)))

    def train_step_indices(self, idx, lr: float) -> float:
        """One Adam step on the resident samples idx (gathered on the device)."""
        ix = np.ascontiguousarray(idx, dtype=np.int32)
        loss = C.c_double()
        _check(_lib().ehd_fno_train_step_indices(self._h, ix.size, ix.ctypes.data_as(C.POINTER(C.c_int)),
                                                 float(lr), C.byref(loss)))
        self._resident = ix.size
        return loss.value

    def train_step(self, a, target, lr: float, batch: int = 0) -> float:
        """One Adam step on (a, target); a and target None: on the batch the
        previous call left on the device (`batch` samples of it, default all)."""
        loss = C.c_double()
        if a is None and target is None:
            _check(_lib().ehd_fno_train_step(self._h, int(batch or getattr(self, "_resident", 0)), None, None,
                                             float(lr), C.byref(loss)))
            return loss.value
        a = np.ascontiguousarray(a, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        _check(_lib().ehd_fno_train_step(self._h, a.shape[0], _fp(a), _fp(t), float(lr), C.byref(loss)))
        self._resident = a.shape[0]
        return loss.value
