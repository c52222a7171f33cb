"""The FP64 FNO oracle pinned on the CPU: its spectral convolution equals
numpy's real FFTs on the truncated spectrum, and its backward pass equals
central finite differences of its forward pass for every parameter tensor."""
import numpy as np
import pytest

from oracle import fno_oracle as F

SMALL = F.FNOConfig(X=8, Y=6, Z=8, width=4, modes=(2, 2, 3), layers=2, d_in=3, d_hidden=5, d_out=2)


def _spectral_numpy(v, Rr, Ri, cfg):
    """rfftn over (x, y, z) with z halved, keep the retained corners, multiply, irfftn."""
    mx, my, mz = cfg.kmodes()
    V = np.fft.rfftn(v, axes=(4, 3, 2))            # the last axis listed (z) is the halved one
    fx, fy = F.mode_freqs(cfg.X, mx), F.mode_freqs(cfg.Y, my)
    S = V[:, :, :mz][:, :, :, fy][:, :, :, :, fx]   # [B, C, KZ, KY, KX]
    Y = np.einsum("bckji,kjico->bokji", S, Rr + 1j * Ri)
    full = np.zeros((v.shape[0], Rr.shape[-1], cfg.Z // 2 + 1, cfg.Y, cfg.X), complex)
    tmp = np.zeros_like(full[:, :, :mz])
    tmp[:, :, :, fy[:, None], fx[None, :]] = Y
    full[:, :, :mz] = tmp
    return np.fft.irfftn(full, s=(cfg.X, cfg.Y, cfg.Z), axes=(4, 3, 2))


@pytest.mark.parametrize("cfg", [SMALL, F.FNOConfig(X=16, Y=12, Z=10, width=3, modes=(5, 4, 4))])
def test_spectral_conv_equals_numpy_rfftn(cfg):
    rng = np.random.default_rng(1)
    mx, my, mz = cfg.kmodes()
    v = rng.standard_normal((2, 3, cfg.Z, cfg.Y, cfg.X))
    Rr = rng.standard_normal((mz, 2 * my, 2 * mx, 3, 3))
    Ri = rng.standard_normal((mz, 2 * my, 2 * mx, 3, 3))
    got, _ = F.spectral(v, Rr, Ri, F.dft_mats(cfg))
    ref = _spectral_numpy(v, Rr, Ri, cfg)
    assert np.allclose(got, ref, rtol=1e-10, atol=1e-12)


def test_gelu_closed_forms():
    from scipy.stats import norm
    x = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    assert F.gelu(np.array([0.0]))[0] == 0.0
    # gelu(x) = x * Phi(x), Phi the standard normal CDF (an independent implementation)
    assert np.allclose(F.gelu(x), x * norm.cdf(x), rtol=1e-14, atol=1e-16)
    assert abs(F.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    h = 1e-6
    assert np.allclose(F.gelu_grad(x), (F.gelu(x + h) - F.gelu(x - h)) / (2 * h), atol=1e-8)


def test_backward_matches_finite_differences():
    cfg = SMALL
    rng = np.random.default_rng(3)
    p = F.init_params(cfg, seed=5)
    for k in p:                      # non-zero biases so that their gradients are exercised too
        if k.endswith("b") or k.startswith("b"):
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
    a = rng.standard_normal((2, cfg.d_in, cfg.Z, cfg.Y, cfg.X))
    target = rng.standard_normal((2, cfg.d_out, cfg.Z, cfg.Y, cfg.X))
    M = F.dft_mats(cfg)

    def loss(pp):
        out, _ = F.forward(cfg, pp, a, M)
        return F.mse(out, target)[0]

    out, cache = F.forward(cfg, p, a, M)
    _, dout = F.mse(out, target)
    g = F.backward(cfg, p, cache, dout, M)
    assert set(g) == set(p)
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            save = flat[idx]
            flat[idx] = save + h
            lp = loss(p)
            flat[idx] = save - h
            lm = loss(p)
            flat[idx] = save
            fd = (lp - lm) / (2 * h)
            an = g[k].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, idx, fd, an)


def test_adam_first_step_is_signed_lr():
    p = {"w": np.array([1.0, -2.0, 3.0])}
    g = {"w": np.array([0.5, -0.1, 0.0])}
    m = {"w": np.zeros(3)}
    v = {"w": np.zeros(3)}
    F.adam_step(p, g, m, v, 1, lr=0.01)
    # bias-corrected first step: lr * g / (|g| + eps) = lr * sign(g)
    assert np.allclose(p["w"], [0.99, -1.99, 3.0], atol=1e-7)
