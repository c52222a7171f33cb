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


def _up4(v):
    return (v + 3) // 4 * 4


def _operand(layout, R, K, nb, rng):
    """Flat host array of nb (R x K) matrices (R = m for A, n for B) in the
    given layout; returns (flat, (r stride, k stride, batch stride), dense)."""
    if layout == "k":        # k unit-stride, rows 16-byte aligned
        sr, sk = _up4(K + 1), 1
    elif layout == "r":      # m / n unit-stride
        sr, sk = 1, _up4(R + 1)
    else:                    # odd strides: element staging
        sr, sk = 3, 3 * R + 1
    sb = _up4((R - 1) * sr + (K - 1) * sk + 5)
    flat = rng.standard_normal(nb * sb).astype(np.float32)
    r, k = np.meshgrid(np.arange(R), np.arange(K), indexing="ij")
    dense = np.stack([flat[b * sb + r * sr + k * sk] for b in range(nb)])
    return flat, (sr, sk, sb), dense


def _result_strides(layout, M, N, ns):
    """(cm, cn, c1, c2) of C: n unit-stride with every offset a multiple of 4
    floats ("n"), the same along m ("m"), or odd offsets ("x")."""
    if layout == "n":
        cm, cn = _up4(N + 1), 1
        c2 = M * cm + 4
        return cm, cn, ns * c2 + 8, c2
    if layout == "m":
        cm, cn = 1, _up4(M + 1)
        c2 = N * cn + 4
        return cm, cn, ns * c2 + 8, c2
    cm, cn = N + 3, 1
    c2 = M * cm + 1
    return cm, cn, ns * c2 + 7, c2


@pytest.mark.parametrize("la", ["k", "r", "x"])
@pytest.mark.parametrize("lb", ["k", "r", "x"])
@pytest.mark.parametrize("lc", ["n", "m", "x"])
@pytest.mark.parametrize("mnk", [(64, 128, 96), (36, 44, 28), (37, 45, 29), (8, 300, 4), (130, 24, 64)])
def test_gemm_strided_layouts(la, lb, lc, mnk):
    """Every staging layout of A (k, m or neither unit-stride) and B (k, n or
    neither), every result layout (16-byte stores along n or m, or element
    stores), full and ragged tiles, several k-steps, batches, alpha, beta and
    split-K slices, against FP64 products of the bf16-rounded operands."""
    M, N, K = mnk
    nb = 3
    rng = np.random.default_rng(M * 7 + N * 3 + K + ord(la) * 11 + ord(lb) + ord(lc) * 5)
    A, (am, ak, a1), Ad = _operand(la, M, K, nb, rng)
    Bt, (bn, bk, b1), Btd = _operand(lb, N, K, nb, rng)     # B stored as its transpose's layout: rows n
    for ksplit in (0, 32):
        ns = -(-K // ksplit) if ksplit else 1
        cm, cn, c1, c2 = _result_strides(lc, M, N, ns)
        C0 = rng.standard_normal(nb * c1).astype(np.float32)
        alpha, beta = (1.0, 0.0) if ksplit else (-0.75, 0.5)
        got = G.gemm_strided(M, N, K, A, (am, ak, a1), Bt, (bk, bn, b1), C0.copy(), (cm, cn, c1, c2), nb=nb,
                             ksplit=ksplit, alpha=alpha, beta=beta)
        for b in range(nb):
            Ab = _bf16(Ad[b]).astype(np.float64)
            Bb = _bf16(Btd[b]).astype(np.float64).T
            for sl in range(ns):
                k0, k1 = (sl * ksplit, min(K, (sl + 1) * ksplit)) if ksplit else (0, K)
                m, n = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
                idx = b * c1 + sl * c2 + m * cm + n * cn
                ref = alpha * (Ab[:, k0:k1] @ Bb[k0:k1]) + beta * C0[idx].astype(np.float64)
                scale = np.abs(Ab[:, k0:k1]) @ np.abs(Bb[k0:k1]) + abs(beta) * np.abs(C0[idx])
                err = np.abs(got[idx] - ref)
                assert np.all(err <= 1e-5 * scale + 1e-6), (b, sl, float(np.max(err / (scale + 1e-30))))
        # elements outside C's tiles are untouched
        mask = np.ones(C0.size, bool)
        for b in range(nb):
            for sl in range(ns):
                m, n = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
                mask[(b * c1 + sl * c2 + m * cm + n * cn).ravel()] = False
        assert np.array_equal(got[mask], C0[mask])


# (imaginary A, imaginary B, imaginary C written, sa, sb): the forms the
# spectral layers use (complex x complex, complex x real, real part only)
# and the rest
CMODES = [(True, True, True, 1, 1), (True, True, True, -1, 1), (True, True, True, 1, -1),
          (True, False, True, -1, 1), (True, True, False, -1, 1), (False, True, True, 1, -1),
          (False, False, True, 1, 1)]


@pytest.mark.parametrize("la", ["k", "r", "x"])
@pytest.mark.parametrize("lb", ["k", "r", "x"])
@pytest.mark.parametrize("lc", ["n", "m", "x"])
@pytest.mark.parametrize("mnk", [(36, 44, 40), (37, 45, 29)])
def test_cgemm_strided_forms(la, lb, lc, mnk):
    """The one-launch complex product in every staging and result layout,
    each form of it (an absent imaginary plane, the real part only,
    conjugates), alpha and beta, against FP64 complex products of the
    bf16-rounded planes; elements outside C are untouched."""
    M, N, K = mnk
    nb = 2
    rng = np.random.default_rng(M * 5 + N + K * 3 + ord(la) * 7 + ord(lb) * 13 + ord(lc) * 17)
    Ar, (am, ak, a1), Ard = _operand(la, M, K, nb, rng)
    Br, (bn, bk, b1), Brd = _operand(lb, N, K, nb, rng)
    Ai, Bi = rng.standard_normal(Ar.size).astype(np.float32), rng.standard_normal(Br.size).astype(np.float32)
    r, k = np.meshgrid(np.arange(M), np.arange(K), indexing="ij")
    Aid = np.stack([Ai[b * a1 + r * am + k * ak] for b in range(nb)])
    r, k = np.meshgrid(np.arange(N), np.arange(K), indexing="ij")
    Bid = np.stack([Bi[b * b1 + r * bn + k * bk] for b in range(nb)])
    cm, cn, c1, _ = _result_strides(lc, M, N, 1)
    m, n = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    for hai, hbi, hci, sa, sb in CMODES:
        alpha, beta = (0.75, -0.5) if hci else (1.0, 0.0)
        C0r = rng.standard_normal(nb * c1).astype(np.float32)
        C0i = rng.standard_normal(nb * c1).astype(np.float32)
        gr, gi = G.cgemm_strided(M, N, K, Ar, Ai if hai else None, (am, ak, a1), Br, Bi if hbi else None,
                                 (bk, bn, b1), C0r.copy(), C0i.copy() if hci else None, (cm, cn, c1), nb=nb, sa=sa,
                                 sb=sb, alpha=alpha, beta=beta)
        mask = np.ones(C0r.size, bool)
        for b in range(nb):
            Ab = _bf16(Ard[b]) + 1j * sa * (_bf16(Aid[b]) if hai else 0)
            Bb = (_bf16(Brd[b]) + 1j * sb * (_bf16(Bid[b]) if hbi else 0)).T
            idx = b * c1 + m * cm + n * cn
            mask[idx.ravel()] = False
            prod = Ab.astype(np.complex128) @ Bb.astype(np.complex128)
            scale = np.abs(Ab) @ np.abs(Bb) + abs(beta) * (np.abs(C0r[idx]) + np.abs(C0i[idx]))
            ref_r = alpha * prod.real + beta * C0r[idx]
            err = np.abs(gr[idx] - ref_r)
            assert np.all(err <= 1e-5 * scale + 1e-6), ("re", b, hai, hbi, hci, sa, sb, float(np.max(err / scale)))
            if hci:
                ref_i = alpha * prod.imag + beta * C0i[idx]
                err = np.abs(gi[idx] - ref_i)
                assert np.all(err <= 1e-5 * scale + 1e-6), ("im", b, hai, hbi, sa, sb, float(np.max(err / scale)))
        assert np.array_equal(gr[mask], C0r[mask])
        if hci:
            assert np.array_equal(gi[mask], C0i[mask])


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


def test_gradients_with_split_k_slices_match_oracle():
    """A grid of more points than one split-K chunk of the weight-gradient
    GEMMs (16x16x12 = 3072 > 2048), so those gradients are sums over several
    slices per sample."""
    cfg = F.FNOConfig(X=16, Y=16, Z=12, width=8, modes=(3, 4, 2), layers=2, d_in=6, d_hidden=16, d_out=5)
    rng = np.random.default_rng(5)
    p = F.init_params(cfg, seed=5)
    for k in p:
        if k in ("Pb", "Q1b", "Q2b") or k.startswith("b"):
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
        if k.startswith("R"):
            p[k] = p[k] * 20.0
    a = rng.standard_normal((2, cfg.d_in, cfg.Z, cfg.Y, cfg.X))
    t = rng.standard_normal((2, cfg.d_out, cfg.Z, cfg.Y, cfg.X))
    net = G.FNO(cfg.X, cfg.Y, cfg.Z, width=cfg.width, modes=cfg.modes, layers=cfg.layers, d_in=cfg.d_in,
                d_hidden=cfg.d_hidden, d_out=cfg.d_out, max_batch=2)
    net.set_params(p)
    out, cache = F.forward(cfg, p, a)
    loss_ref, dout = F.mse(out, t)
    gref = F.backward(cfg, p, cache, dout)
    loss, g = net.gradients(a, t)
    assert abs(loss / loss_ref - 1) < 1e-2, (loss, loss_ref)
    gg = G.unflatten(net.cfg, g)
    errs = {k: _rel(gg[k], gref[k]) for k in gref}
    print("gradient rel-L2 per tensor (split-K):", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) < 5e-2, errs


def test_adam_steps_reduce_loss():
    p, a, t, net = _setup(seed=2)
    losses = [net.train_step(a, t, 3e-3) for _ in range(30)]
    print("losses:", [f"{x:.4f}" for x in losses[::5]])
    assert losses[-1] < 0.9 * losses[0]


def test_resident_batch_steps_match_host_batches():
    """Training steps on the batch left on the device (no host transfer) give
    bit for bit what the same steps with the host arrays give; a forward call
    in between invalidates the resident batch."""
    p, a, t, n1 = _setup(seed=3)
    _, _, _, n2 = _setup(seed=3)
    l1 = [n1.train_step(a, t, 2e-3) for _ in range(3)]
    l2 = [n2.train_step(a, t, 2e-3)] + [n2.train_step(None, None, 2e-3) for _ in range(2)]
    assert l1 == l2
    assert np.array_equal(n1.get_params(), n2.get_params())
    n2.forward(a)
    with pytest.raises(Exception):
        n2.train_step(None, None, 2e-3, batch=a.shape[0])


def test_resident_dataset_steps_match_host_batches():
    """Steps on mini-batches gathered on the device from a resident training
    set give bit for bit what the same steps on host copies of those samples
    give; out-of-range indices and a batch above max_batch are refused."""
    rng = np.random.default_rng(11)
    _, _, _, n1 = _setup(seed=5)
    _, _, _, n2 = _setup(seed=5)
    ds_a = rng.standard_normal((7, CFG.d_in, CFG.Z, CFG.Y, CFG.X)).astype(np.float32)
    ds_t = rng.standard_normal((7, CFG.d_out, CFG.Z, CFG.Y, CFG.X)).astype(np.float32)
    n2.set_dataset(ds_a, ds_t)
    for idx in ([6, 0, 3], [2, 2, 5, 1], [4]):
        l1 = n1.train_step(ds_a[idx], ds_t[idx], 2e-3)
        l2 = n2.train_step_indices(idx, 2e-3)
        assert l1 == l2
    assert np.array_equal(n1.get_params(), n2.get_params())
    with pytest.raises(Exception):
        n2.train_step_indices([0, 7], 2e-3)
    with pytest.raises(Exception):
        n2.train_step_indices([0, 1, 2, 3, 4], 2e-3)
    # a new set replaces the old one
    n2.set_dataset(ds_a[:2], ds_t[:2])
    with pytest.raises(Exception):
        n2.train_step_indices([2], 2e-3)
