"""The surrogate's HIP kernels against the FP64 FNO oracle: the bf16
matrix-core GEMM (operand and result lane layout), the forward pass, every
parameter gradient, and a few Adam steps.  Tolerances are those of bf16
operands with FP32 accumulation (8 significant bits per product operand)."""
import numpy as np
import pytest

from ehd_amd import fno as G
from oracle import fno_oracle as F

pytestmark = pytest.mark.gpu


def _bf16(x):
    """Round float32 to bf16 (nearest even) and back, as the kernel stages operands."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("mnk", [(16, 16, 16), (37, 45, 29), (130, 70, 200), (64, 1, 7), (1, 129, 33)])
def test_gemm_matches_bf16_reference(mnk):
    M, N, K = mnk
    rng = np.random.default_rng(M * 1000 + N)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    got = G.gemm(A, B)
    ref = _bf16(A).astype(np.float64) @ _bf16(B).astype(np.float64)   # bf16 products are exact in FP32
    scale = np.abs(_bf16(A)).astype(np.float64) @ np.abs(_bf16(B)).astype(np.float64)
    assert np.all(np.abs(got - ref) <= 1e-5 * scale + 1e-6), np.max(np.abs(got - ref) / (scale + 1e-30))


CFG = F.FNOConfig(X=16, Y=12, Z=8, width=8, modes=(4, 3, 3), layers=3, d_in=6, d_hidden=16, d_out=5)


def _setup(seed=0, B=3):
    rng = np.random.default_rng(seed)
    p = F.init_params(CFG, seed=seed)
    for k in p:
        if k in ("Pb", "Q1b", "Q2b") or k.startswith("b"):
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
        if k.startswith("R"):
            p[k] = p[k] * 20.0           # spectral path comparable to the pointwise path
    a = rng.standard_normal((B, CFG.d_in, CFG.Z, CFG.Y, CFG.X))
    t = rng.standard_normal((B, CFG.d_out, CFG.Z, CFG.Y, CFG.X))
    net = G.FNO(CFG.X, CFG.Y, CFG.Z, width=CFG.width, modes=CFG.modes, layers=CFG.layers, d_in=CFG.d_in,
                d_hidden=CFG.d_hidden, d_out=CFG.d_out, max_batch=4)
    net.set_params(p)
    return p, a, t, net


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_oracle():
    p, a, _, net = _setup()
    ref, _ = F.forward(CFG, p, a)
    got = net.forward(a)
    err = _rel(got, ref)
    print(f"FNO forward rel-L2 vs FP64 oracle: {err:.3e}")
    assert err < 2e-2
    assert np.array_equal(net.get_params(), G.flatten(net.cfg, p))


def test_gradients_match_oracle():
    p, a, t, net = _setup(seed=1)
    out, cache = F.forward(CFG, p, a)
    loss_ref, dout = F.mse(out, t)
    gref = F.backward(CFG, p, cache, dout)
    loss, g = net.gradients(a, t)
    assert abs(loss / loss_ref - 1) < 1e-2, (loss, loss_ref)
    gg = G.unflatten(net.cfg, g)
    errs = {k: _rel(gg[k], gref[k]) for k in gref}
    print("gradient rel-L2 per tensor:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) < 5e-2, errs


def test_adam_steps_reduce_loss():
    p, a, t, net = _setup(seed=2)
    losses = [net.train_step(a, t, 3e-3) for _ in range(30)]
    print("losses:", [f"{x:.4f}" for x in losses[::5]])
    assert losses[-1] < 0.9 * losses[0]
