"""What the trained surrogate gains over the method without it, on 64x64x64
boxes (C = 10, T between 200 and 300, where the flow settles from noise to
steady square cells): after training on steady states from the batched
solver (canonical pattern position), the network's
prediction for held-out T is much closer to their steady state than the
hydrostatic start the solver otherwise begins from, the warm-started runs
end at the cold runs' Nu (the acceptance rule: within 0.5 %), and they
reach steady state in fewer steps in total."""
import math

import numpy as np
import pytest

from ehd_amd import surrogate as S
from ehd_amd.fno import FNO

pytestmark = pytest.mark.gpu

GRID = (64, 64, 64)
TRAIN_T = (200.0, 220.0, 240.0, 260.0, 280.0, 300.0)
TEST_T = (230.0, 270.0)
C = 10.0
RUN = dict(check_every=2000, tol=1e-5, max_steps=300000)


@pytest.mark.timeout(600)   # ~100 s: steady states of 8 boxes, 2000 training steps, warm runs
def test_warm_start_gain():
    # training states and the cold runs of the held-out T, one batch
    draws = TRAIN_T + TEST_T
    params = [S.box_params(GRID, T, C, seed=100 + i) for i, T in enumerate(draws)]
    res, fields = S.run_batch_to_steady(params, tag="cold", **RUN)
    for T, r in zip(draws, res):
        print(f"cold T={T:.0f}: {r.steps} steps, Nu {r.nusselt:.5f}, converged {r.converged}, failed {r.failed}")
    assert all(r.converged for r in res)
    Y = np.stack([S.canonicalize(S.targets_from_fields(p, f, S.hydrostatic_profile(p)))
                  for p, f in zip(params, fields)])
    X = np.stack([S.features(GRID, T, C) for T in draws])
    ntr = len(TRAIN_T)
    Xtr, Ytr, Xte, Yte = X[:ntr], Y[:ntr], X[ntr:], Y[ntr:]
    res_cold = res[ntr:]

    net = FNO(*GRID, width=16, modes=(8, 8, 16), layers=4, max_batch=ntr)
    net.set_params(net.init_params(seed=3))
    losses = []
    nit = 2000
    for it in range(nit):
        lr = 2e-3 * 0.5 * (1 + math.cos(math.pi * it / nit))
        losses.append(net.train_step(Xtr, Ytr, lr))
    print(f"training loss {np.mean(losses[:10]):.3e} -> {np.mean(losses[-10:]):.3e}")
    assert np.mean(losses[-10:]) < 0.01 * np.mean(losses[:10])

    pred = net.forward(Xte)
    # distance to the steady state (canonical phase): network vs the hydrostatic
    # start (zero velocity, hydrostatic q and phi, i.e. all targets zero)
    err_net = float(np.linalg.norm(pred - Yte) / np.linalg.norm(Yte))
    print(f"held-out rel-L2 to the steady state: network {err_net:.3f}, hydrostatic start 1.000")
    assert err_net < 0.05

    warm_params, warm_fields = [], []
    for i, (T, y) in enumerate(zip(TEST_T, pred)):
        pw = S.box_params(GRID, T, C, seed=900 + i, init=S.INIT_FIELDS)
        warm_params.append(pw)
        warm_fields.append(S.fields9_from_prediction(pw, y, S.hydrostatic_profile(pw)))
    res_warm, _ = S.run_batch_to_steady(warm_params, warm_fields, tag="warm", consistent=5000, **RUN)
    cold = sum(r.steps for r in res_cold)
    warm = sum(r.steps for r in res_warm)
    for T, c, w in zip(TEST_T, res_cold, res_warm):
        print(f"T={T:.0f}: cold {c.steps} steps Nu {c.nusselt:.5f}; warm {w.steps} steps Nu {w.nusselt:.5f}")
    print(f"total steps to steady: cold {cold}, warm {warm} (ratio {cold / warm:.2f})")
    assert all(w.converged for w in res_warm)
    for c, w in zip(res_cold, res_warm):
        assert abs(w.nusselt / c.nusselt - 1) < 5e-3
    assert warm < cold
