"""The consistent start (ehd_init_consistent / ehd_batch_init_consistent):
boxes without target fields in a batch step exactly as they do normally,
the single-box and batched forms agree bit for bit, and from the fields of a
developed steady state the consistent start returns to steady sooner than
the equilibrium start (which lacks the populations' non-equilibrium parts)."""
import numpy as np
import pytest

import ehd_amd as E

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "q", "phi", "ux", "uy", "uz", "ex", "ey", "ez")


def _nine(f):
    return np.stack([f[k] for k in FIELDS])


def _p(n, init, seed, T=300.0, cfl=0.01):
    return E.make_params(*n, T=T, C_=10.0, lam=0.25, precision=E.FP32S, init_mode=init, seed=seed, amp=1e-3,
                         cfl=cfl)


def _state(b):
    return b.read_fields(), b.read_populations(), b.diagnostics()


def _same(a, b, what):
    fa, pa, da = a
    fb, pb, db = b
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), f"{what}: {k}"
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y), f"{what}: populations"
    assert da.step == db.step and da.current == db.current


def test_untargeted_boxes_step_normally_and_batch_matches_single():
    n = (32, 24, 16)
    src = E.Box(_p(n, E.INIT_HYDROSTATIC, 3))
    src.init()
    src.step(40)
    f9 = _nine(src.read_fields())
    src.close()
    K = 25
    with E.Batch([_p(n, E.INIT_FIELDS, 5), _p(n, E.INIT_HYDROSTATIC, 7)]) as bt:
        bt.init_consistent([f9, None], K)
        bt.step(6)
        got = [_state(b) for b in bt.boxes]
    with E.Box(_p(n, E.INIT_HYDROSTATIC, 7)) as b:
        b.init()
        b.step(K + 6)
        _same(got[1], _state(b), "untargeted box vs normal stepping")
    with E.Box(_p(n, E.INIT_FIELDS, 5)) as b:
        b.init_consistent(f9, K)
        assert b.diagnostics().step == K
        b.step(6)
        _same(got[0], _state(b), "batched vs single consistent start")


def test_consistent_start_rejects_fp64_and_wrong_mode():
    n = (16, 16, 8)
    f9 = np.zeros((9,) + n[::-1])
    with E.Box(E.make_params(*n, precision=E.FP64, init_mode=E.INIT_FIELDS)) as b:
        with pytest.raises(E.EhdError):
            b.init_consistent(f9, 10)
    with E.Box(_p(n, E.INIT_HYDROSTATIC, 1)) as b:
        with pytest.raises(E.EhdError):
            b.init_consistent(f9, 10)


def _steps_to_quiet(b, tol=1e-5, every=2000, limit=120000):
    prev = b.diagnostics().nusselt
    taken = 0
    quiet = 0
    while taken < limit:
        b.step(every)
        taken += every
        d = b.diagnostics()
        quiet = quiet + 1 if abs(d.nusselt - prev) < tol else 0
        prev = d.nusselt
        if quiet >= 2:
            return taken, d.nusselt
    return taken, prev


def test_consistent_start_settles_sooner_than_equilibrium_start():
    """64x64x64, T=300, C=10, dt/dx 0.04: a developed steady convection state
    (1.5e5 steps from noise); its own fields restarted by the equilibrium
    start and by the consistent start (5000 iterations, counted)."""
    n = (64, 64, 64)
    with E.Box(_p(n, E.INIT_HYDROSTATIC, 11, cfl=0.04)) as b:
        b.init()
        b.step(150000)
        nu = b.diagnostics().nusselt
        f9 = _nine(b.read_fields())
    f9[1][0] = f9[1][0].mean()          # the injected charge on the injection wall
    with E.Box(_p(n, E.INIT_FIELDS, 12, cfl=0.04)) as b:
        b.init(f9)
        eq_steps, eq_nu = _steps_to_quiet(b)
    with E.Box(_p(n, E.INIT_FIELDS, 12, cfl=0.04)) as b:
        b.init_consistent(f9, 5000)
        cs_steps, cs_nu = _steps_to_quiet(b)
        cs_steps += 5000
    print(f"steady Nu {nu:.6f}; equilibrium start: {eq_steps} steps (Nu {eq_nu:.6f}); "
          f"consistent start: {cs_steps} steps (Nu {cs_nu:.6f})")
    assert abs(eq_nu / nu - 1) < 1e-4 and abs(cs_nu / nu - 1) < 1e-4
    assert cs_steps < eq_steps
