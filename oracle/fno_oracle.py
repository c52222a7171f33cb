"""FP64 numpy oracle of the Fourier-neural-operator surrogate (test infrastructure).

The reference has no learned component; the surrogate is a build-only part
of this project: a 3-D FNO (Li et al., "Fourier Neural Operator for
Parametric PDEs", ICLR 2021) mapping the governing groups and the grid
coordinates of a box to its steady fields, used to warm-start the lattice
Boltzmann run.  This file restates the network exactly as the HIP kernels
compute it (same layer order, same truncated DFTs written as matrices, same
activation), in float64, forward and backward, so that the GPU forward pass
and every GPU gradient can be checked against it.  Only tests/, smoke() and
bench.py's CPU leg use it.

Network (activations in [B, C, Z, Y, X] layout, x fastest):
  lift      v0 = a @ Pw + Pb                          (d_in -> width per point)
  layer l   u  = K_l(v) + v @ W_l + b_l,  v' = gelu(u) (no gelu after the last)
  K_l(v)    = irfft_z( ifft_y( ifft_x( R_l * fft_x( fft_y( rfft_z(v) ))))) on the
              retained modes kz in [0, mz), ky in {0..my-1, Y-my..Y-1},
              kx in {0..mx-1, X-mx..X-1}, with complex weights R_l[kz,ky,kx,c,o]
  project   h = gelu(v @ Q1w + Q1b), out = h @ Q2w + Q2b
gelu(x) = x/2 (1 + erf(x/sqrt 2)).  Loss: mean squared error over all outputs.

Pinned by: the spectral convolution equals numpy's rfftn/irfftn on the
truncated spectrum (tests/test_fno_oracle.py), and every gradient against
central finite differences of the forward pass.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from math import erf, sqrt

import numpy as np


@dataclass
class FNOConfig:
    X: int = 64
    Y: int = 64
    Z: int = 32
    width: int = 32
    modes: tuple = (12, 12, 12)       # (mx, my, mz); capped at X/2, Y/2, Z/2
    layers: int = 4
    d_in: int = 6
    d_hidden: int = 128
    d_out: int = 5

    def kmodes(self):
        mx = min(self.modes[0], self.X // 2)
        my = min(self.modes[1], self.Y // 2)
        mz = min(self.modes[2], self.Z // 2)
        return mx, my, mz


def mode_freqs(n: int, m: int) -> np.ndarray:
    """Integer frequencies of the retained modes of a full axis: 0..m-1, n-m..n-1."""
    return np.concatenate([np.arange(m), np.arange(n - m, n)])


def dft_mats(cfg: FNOConfig):
    """Forward (grid -> modes) and inverse (modes -> grid) matrices of each axis."""
    mx, my, mz = cfg.kmodes()
    X, Y, Z = cfg.X, cfg.Y, cfg.Z
    fx, fy, fz = mode_freqs(X, mx), mode_freqs(Y, my), np.arange(mz)
    Cx = np.exp(-2j * np.pi * np.outer(np.arange(X), fx) / X)          # [X, KX]
    By = np.exp(-2j * np.pi * np.outer(np.arange(Y), fy) / Y)          # [Y, KY]
    Az = np.exp(-2j * np.pi * np.outer(np.arange(Z), fz) / Z)          # [Z, KZ]
    Cxi = np.exp(2j * np.pi * np.outer(fx, np.arange(X)) / X) / X      # [KX, X]
    Byi = np.exp(2j * np.pi * np.outer(fy, np.arange(Y)) / Y) / Y      # [KY, Y]
    cz = np.where(fz == 0, 1.0, 2.0)
    Azi = cz[:, None] * np.exp(2j * np.pi * np.outer(fz, np.arange(Z)) / Z) / Z   # [KZ, Z], real part taken
    return dict(Cx=Cx, By=By, Az=Az, Cxi=Cxi, Byi=Byi, Azi=Azi)


def init_params(cfg: FNOConfig, seed: int = 0) -> dict:
    """Float64 parameters (the GPU copy is float32).  R is stored as Re/Im arrays."""
    rng = np.random.default_rng(seed)
    mx, my, mz = cfg.kmodes()
    w, L = cfg.width, cfg.layers
    p = {"Pw": rng.standard_normal((cfg.d_in, w)) / sqrt(cfg.d_in), "Pb": np.zeros(w)}
    sc = 1.0 / (w * w)
    for l in range(L):
        p[f"W{l}"] = rng.standard_normal((w, w)) / sqrt(w)
        p[f"b{l}"] = np.zeros(w)
        p[f"Rr{l}"] = sc * rng.uniform(0, 1, (mz, 2 * my, 2 * mx, w, w))
        p[f"Ri{l}"] = sc * rng.uniform(0, 1, (mz, 2 * my, 2 * mx, w, w))
    p["Q1w"] = rng.standard_normal((w, cfg.d_hidden)) / sqrt(w)
    p["Q1b"] = np.zeros(cfg.d_hidden)
    p["Q2w"] = rng.standard_normal((cfg.d_hidden, cfg.d_out)) / sqrt(cfg.d_hidden)
    p["Q2b"] = np.zeros(cfg.d_out)
    return p


_erf = np.vectorize(erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / sqrt(2.0))) + x * np.exp(-0.5 * x * x) / sqrt(2.0 * np.pi)


def pointwise(v, Wm, b):
    """[B, C, Z, Y, X] x [C, O] + b -> [B, O, Z, Y, X]"""
    return np.einsum("bczyx,co->bozyx", v, Wm) + b[None, :, None, None, None]


def spectral(v, Rr, Ri, M):
    R = Rr + 1j * Ri
    Vz = np.einsum("bczyx,zk->bckyx", v, M["Az"])
    Vy = np.einsum("bckyx,yj->bckjx", Vz, M["By"])
    S = np.einsum("bckjx,xi->bckji", Vy, M["Cx"])
    Ym = np.einsum("bckji,kjico->bokji", S, R)
    Wx = np.einsum("bokji,ix->bokjx", Ym, M["Cxi"])
    Wy = np.einsum("bokjx,jy->bokyx", Wx, M["Byi"])
    out = np.real(np.einsum("bokyx,kz->bozyx", Wy, M["Azi"]))
    return out, S


def spectral_backward(g, S, Rr, Ri, M):
    R = Rr + 1j * Ri
    G_Wy = np.einsum("bozyx,kz->bokyx", g, np.conj(M["Azi"]))
    G_Wx = np.einsum("bokyx,jy->bokjx", G_Wy, np.conj(M["Byi"]))
    G_Ym = np.einsum("bokjx,ix->bokji", G_Wx, np.conj(M["Cxi"]))
    G_S = np.einsum("bokji,kjico->bckji", G_Ym, np.conj(R))
    G_R = np.einsum("bckji,bokji->kjico", np.conj(S), G_Ym)
    G_Vy = np.einsum("bckji,xi->bckjx", G_S, np.conj(M["Cx"]))
    G_Vz = np.einsum("bckjx,yj->bckyx", G_Vy, np.conj(M["By"]))
    gv = np.real(np.einsum("bckyx,zk->bczyx", G_Vz, np.conj(M["Az"])))
    return gv, np.real(G_R), np.imag(G_R)


def forward(cfg: FNOConfig, p: dict, a: np.ndarray, M=None):
    """a: [B, d_in, Z, Y, X] -> out [B, d_out, Z, Y, X] and the cache for backward."""
    M = M or dft_mats(cfg)
    cache = {"a": a}
    v = pointwise(a, p["Pw"], p["Pb"])
    for l in range(cfg.layers):
        cache[f"v{l}"] = v
        s, S = spectral(v, p[f"Rr{l}"], p[f"Ri{l}"], M)
        u = s + pointwise(v, p[f"W{l}"], p[f"b{l}"])
        cache[f"S{l}"] = S
        cache[f"u{l}"] = u
        v = gelu(u) if l < cfg.layers - 1 else u
    cache["vL"] = v
    z1 = pointwise(v, p["Q1w"], p["Q1b"])
    cache["z1"] = z1
    h = gelu(z1)
    cache["h"] = h
    out = pointwise(h, p["Q2w"], p["Q2b"])
    return out, cache


def backward(cfg: FNOConfig, p: dict, cache: dict, dout: np.ndarray, M=None) -> dict:
    M = M or dft_mats(cfg)
    g = {}
    h = cache["h"]
    g["Q2w"] = np.einsum("bczyx,bozyx->co", h, dout)
    g["Q2b"] = dout.sum(axis=(0, 2, 3, 4))
    dh = np.einsum("bozyx,co->bczyx", dout, p["Q2w"])
    dz1 = dh * gelu_grad(cache["z1"])
    v = cache["vL"]
    g["Q1w"] = np.einsum("bczyx,bozyx->co", v, dz1)
    g["Q1b"] = dz1.sum(axis=(0, 2, 3, 4))
    dv = np.einsum("bozyx,co->bczyx", dz1, p["Q1w"])
    for l in reversed(range(cfg.layers)):
        u = cache[f"u{l}"]
        du = dv * gelu_grad(u) if l < cfg.layers - 1 else dv
        vin = cache[f"v{l}"]
        g[f"W{l}"] = np.einsum("bczyx,bozyx->co", vin, du)
        g[f"b{l}"] = du.sum(axis=(0, 2, 3, 4))
        dvs, g[f"Rr{l}"], g[f"Ri{l}"] = spectral_backward(du, cache[f"S{l}"], p[f"Rr{l}"], p[f"Ri{l}"], M)
        dv = dvs + np.einsum("bozyx,co->bczyx", du, p[f"W{l}"])
    a = cache["a"]
    g["Pw"] = np.einsum("bczyx,bozyx->co", a, dv)
    g["Pb"] = dv.sum(axis=(0, 2, 3, 4))
    return g


def mse(out, target):
    """Mean squared error and its gradient with respect to out."""
    d = out - target
    return float(np.mean(d * d)), 2.0 * d / d.size


def adam_step(p: dict, g: dict, m: dict, v: dict, t: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in place; t counts from 1."""
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)


def features(cfg: FNOConfig, T: float, C: float, Mm: float) -> np.ndarray:
    """Network input of one box: the normalised groups broadcast over the grid
    plus the node coordinates in [0, 1) (x, y periodic) and [0, 1] (z)."""
    Z, Y, X = cfg.Z, cfg.Y, cfg.X
    a = np.empty((cfg.d_in, Z, Y, X))
    a[0] = (np.log(T) - np.log(400.0))
    a[1] = (np.log(C) - np.log(10.0))
    a[2] = (np.log(Mm) - np.log(10.0))
    a[3] = (np.arange(X) / X)[None, None, :]
    a[4] = (np.arange(Y) / Y)[None, :, None]
    a[5] = (np.arange(Z) / (Z - 1))[:, None, None]
    return a
