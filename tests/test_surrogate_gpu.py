"""What the trained surrogate gains over the method without it, on a small
grid: after training on steady states from the batched solver, the
network's prediction for held-out (T, C) draws is much closer to their steady
state than the hydrostatic start the solver otherwise begins from, and the
warm-started runs reach steady state in fewer steps in total."""
import math

import numpy as np
import pytest

from ehd_amd import surrogate as S
from ehd_amd.fno import FNO

pytestmark = pytest.mark.gpu

GRID = (32, 32, 64)


def _draws(rng, k):
    return [(float(rng.uniform(650.0, 750.0)), float(rng.uniform(8.0, 12.0)))
            for _ in range(k)]


def _steady(draws, seed0):
    params = [S.box_params(GRID, T, C, seed=seed0 + i) for i, (T, C) in enumerate(draws)]
    res, fields = S.run_batch_to_steady(params, check_every=2000, tol=1e-5, max_steps=150000, tag="test")
    assert not any(r.failed for r in res)
    ys = []
    for p, f in zip(params, fields):
        ys.append(S.canonicalize(S.targets_from_fields(p, f, S.hydrostatic_profile(p))))
    return params, res, np.stack(ys)


def test_warm_start_gain():
    rng = np.random.default_rng(7)
    train, test = _draws(rng, 12), _draws(rng, 4)
    _, _, Ytr = _steady(train, 100)
    Xtr = np.stack([S.features(GRID, T, C) for T, C in train])
    net = FNO(*GRID, width=16, modes=(8, 8, 8), layers=4, max_batch=8)
    net.set_params(net.init_params(seed=3))
    losses = []
    for it in range(300):
        idx = rng.choice(len(train), size=8, replace=False)
        lr = 1e-3 * 0.5 * (1 + math.cos(math.pi * it / 300))
        losses.append(net.train_step(Xtr[idx], Ytr[idx], lr))
    print(f"training loss {np.mean(losses[:10]):.3e} -> {np.mean(losses[-10:]):.3e}")
    assert np.mean(losses[-10:]) < 0.5 * np.mean(losses[:10])

    params, res_cold, Yte = _steady(test, 900)
    pred = net.forward(np.stack([S.features(GRID, T, C) for T, C in test]))
    # field error (canonical phase) of the network vs the hydrostatic start
    # (zero velocity, hydrostatic q and phi)
    err_net = np.linalg.norm(pred - Yte) / np.linalg.norm(Yte)
    err_hs = 1.0
    print(f"held-out rel-L2 to the steady state: network {err_net:.3f}, hydrostatic start {err_hs:.3f}")
    assert err_net < 0.5 * err_hs

    # steps to steady: warm starts from the prediction (plus the same noise) vs the cold runs above
    warm_params, fields = [], []
    for i, ((T, C), y) in enumerate(zip(test, pred)):
        pw = S.box_params(GRID, T, C, seed=900 + i, init=S.INIT_FIELDS)
        warm_params.append(pw)
        fields.append(S.fields9_from_prediction(pw, y, S.hydrostatic_profile(pw), noise_seed=900 + i, amp=1e-3))
    res_warm, _ = S.run_batch_to_steady(warm_params, fields, check_every=2000, tol=1e-5, max_steps=150000, tag="warm")
    cold = sum(r.steps for r in res_cold)
    warm = sum(r.steps for r in res_warm)
    for (T, C), c, w in zip(test, res_cold, res_warm):
        print(f"T={T:.0f} C={C:.1f}: cold {c.steps} steps Nu {c.nusselt:.4f}; warm {w.steps} steps Nu {w.nusselt:.4f}")
    print(f"total steps: cold {cold}, warm {warm}")
    assert warm < cold
