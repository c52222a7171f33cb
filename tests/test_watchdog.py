"""The diagnostics watchdog (EHD_ENONFINITE): a NaN written into one
population spreads through the box within a step; ehd_diagnostics then
counts the non-finite nodes and returns EHD_ENONFINITE with every field
filled, ehd_run_to_steady stops with failed = 1, and in a batch only that box
is flagged: the others keep stepping and converge.  The NaN is data written
through ehd_write_populations, not a GPU fault.  (The reference has no
watchdog: LBM.cu:39-57 checks only CUDA errors.)"""
import numpy as np
import pytest

import ehd_amd as E

pytestmark = pytest.mark.gpu


def _poison(box, node=(5, 4, 7), d=3):
    f0, f1, h0, h1 = box.read_populations()
    z, y, x = node
    f1[d - 1, z, y, x] = np.nan
    box.write_populations(f0, f1, h0, h1)


@pytest.mark.parametrize("precision", [E.FP64, E.FP32S])
def test_single_box_watchdog(precision):
    p = E.make_params(16, 12, 10, T=200.0, C_=10.0, lam=0.25, precision=precision, init_mode=E.INIT_HYDROSTATIC,
                      seed=3, amp=1e-3)
    with E.Box(p) as b:
        b.init()
        b.step(2)
        assert b.diagnostics().nonfinite == 0
        _poison(b)
        b.step(3)
        with pytest.raises(E.EhdError) as ei:
            b.diagnostics()
        assert ei.value.code == -5
        d = b.diagnostics(allow_nonfinite=True)
        print(f"non-finite nodes after 3 steps: {d.nonfinite}")
        assert d.nonfinite > 1 and d.step == 5
        s = b.run_to_steady(check_every=10, tol=1e-9, max_steps=100, allow_nonfinite=True)
        assert s.failed == 1 and s.converged == 0 and s.steps == 0


def test_batch_watchdog_flags_one_box_and_steps_the_others():
    params = [E.make_params(16, 16, 17, T=T, C_=10.0, lam=0.25, precision=E.FP32S, init_mode=E.INIT_HYDROSTATIC,
                            seed=s, amp=1e-4) for T, s in ((40.0, 1), (60.0, 2), (50.0, 3))]
    with E.Batch(params) as bt:
        bt.init()
        for b in bt.boxes:
            b.hydrostatic_reference(check_every=500, tol=1e-9)
        bt.step(4)
        # write every box's populations (the batch steps boxes in one representation), one with a NaN
        for i, b in enumerate(bt.boxes):
            if i == 1:
                _poison(b)
            else:
                b.write_populations(*b.read_populations())
        bt.step(500)
        with pytest.raises(E.EhdError) as ei:
            bt.diagnostics()
        assert ei.value.code == -5
        d = bt.diagnostics(allow_nonfinite=True)
        assert d[1].nonfinite > 0 and d[0].nonfinite == 0 and d[2].nonfinite == 0
        assert all(np.isfinite(x.nusselt) for x in (d[0], d[2]))
        res = bt.run_to_steady(check_every=500, tol=1e-5, max_steps=60000)
        for r in res:
            print(f"steps {r.steps} converged {r.converged} failed {r.failed} Nu {r.nusselt:.6f}")
        assert res[1].failed == 1 and res[1].converged == 0 and res[1].steps == 0
        for i in (0, 2):
            assert res[i].failed == 0 and res[i].converged == 1 and abs(res[i].nusselt - 1.0) < 1e-3


def test_debug_sync_steps_match_and_report_by_phase():
    """Debug mode (a synchronisation and an error check after every launch)
    gives the same state as the normal step; it only moves where a fault is
    reported."""
    out = []
    for dbg in (False, True):
        p = E.make_params(32, 32, 16, T=300.0, C_=10.0, lam=0.25, precision=E.FP32S, init_mode=E.INIT_HYDROSTATIC,
                          seed=4, amp=1e-3)
        with E.Box(p) as b:
            b.set_debug_sync(dbg)
            b.init()
            b.step(7)
            out.append((b.read_fields(), b.read_populations()))
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    for x, y in zip(out[0][1], out[1][1]):
        assert np.array_equal(x, y)
