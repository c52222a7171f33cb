"""Batched boxes (ehd_batch_*): every box of a batch gives, bit for bit, what
it gives when stepped alone, whatever launch shapes the autotuner picks and
whether the steady steps replay from the captured HIP graph or not."""
import numpy as np
import pytest

import ehd_amd as E

pytestmark = pytest.mark.gpu


def _params(precision):
    # uneven boxes, as in a lateral-wavelength / T / C sweep; two share (64, 64)
    # with different nz (one Poisson group, jobs of different length)
    spec = [(48, 48, 32, 150.0, 5.0, 11), (64, 64, 32, 300.0, 10.0, 12), (64, 64, 20, 420.0, 20.0, 13),
            (80, 80, 32, 600.0, 10.0, 14)]
    return [E.make_params(nx, ny, nz, T=T, C_=C, lam=0.25, precision=precision, init_mode=E.INIT_HYDROSTATIC,
                          seed=seed, amp=1e-3) for nx, ny, nz, T, C, seed in spec]


def _state(box):
    f = box.read_fields()
    return f, box.read_populations(), box.diagnostics()


def _assert_same_state(a, b, what):
    fa, pa, da = a
    fb, pb, db = b
    for k in fa:
        assert np.array_equal(fa[k], fb[k]), f"{what}: field {k} differs (max {np.abs(fa[k] - fb[k]).max():.3e})"
    for x, y, n in zip(pa, pb, ("f0", "f1", "h0", "h1")):
        assert np.array_equal(x, y), f"{what}: {n} differs (max {np.abs(x - y).max():.3e})"
    assert da.step == db.step and da.time == db.time and da.current == db.current


@pytest.mark.parametrize("precision,flags", [(E.FP32S, E.BATCH_GRAPH | E.BATCH_TUNE), (E.FP32S, 0),
                                             (E.FP64, E.BATCH_GRAPH),
                                             (E.FP32S, E.BATCH_GRAPH | E.BATCH_CHAINS), (E.FP32S, E.BATCH_CHAINS),
                                             (E.FP64, E.BATCH_GRAPH | E.BATCH_CHAINS)])
def test_batch_matches_single_boxes(precision, flags):
    """Also with one LBM launch and Poisson chain per (nx, ny) group on
    concurrent streams (BATCH_CHAINS), replayed from the graph or not."""
    params = _params(precision)
    nsteps = (1, 9, 14)            # first (direct) step, graph pairs, odd remainders
    with E.Batch(params, flags=flags) as bt:
        t = bt.tuning()
        print("tuning:", E_tuning_str(t))
        if flags & E.BATCH_CHAINS:
            assert t.chains == 1
        elif not flags & E.BATCH_TUNE:
            assert t.chains == 0
        bt.init()
        for n in nsteps:
            bt.step(n)
        got = [_state(b) for b in bt.boxes]
    for i, p in enumerate(params):
        with E.Box(p) as b:
            b.init()
            for n in nsteps:
                b.step(n)
            _assert_same_state(got[i], _state(b), f"box {i} ({p.nx}x{p.ny}x{p.nz})")


def E_tuning_str(t):
    g = ", ".join(f"{t.group[i].nx}x{t.group[i].ny}: {t.group[i].nboxes} boxes x{t.group[i].x_lines}/y{t.group[i].y_lines}"
                  f" {t.group[i].ms:.3f} ms" for i in range(t.ngroups))
    return (f"lbm grid {t.lbm_grid} of {t.lbm_tiles} tiles ({t.lbm_ms:.3f} ms); {g}; chains {t.chains} "
            f"(step {t.step_ms[0]:.3f} ms one launch, {t.step_ms[1]:.3f} ms chains)")


def test_batch_tuning_report_and_state_checks():
    params = _params(E.FP32S)
    with E.Batch(params) as bt:
        t = bt.tuning()
        assert t.lbm_tiles == sum((p.nx * p.ny * p.nz + 255) // 256 for p in params)
        assert 0 < t.lbm_grid <= t.lbm_tiles and t.lbm_ms > 0
        assert t.ngroups == 3 and sorted(t.group[i].nboxes for i in range(3)) == [1, 1, 2]
        # three groups: the tuner timed a step with one LBM launch and with chains, kept the faster
        assert t.step_ms[0] > 0 and t.step_ms[1] > 0
        assert t.chains == (1 if t.step_ms[1] < t.step_ms[0] else 0)
        bt.init()
        bt.step(2)
        bt.boxes[1].step(1)            # one box stepped alone: the batch is out of step
        with pytest.raises(E.EhdError):
            bt.step(1)


def test_batch_steady_detector():
    """Subcritical boxes (T below the threshold) relax to the no-flow state;
    every box reports its own convergence step."""
    params = [E.make_params(16, 16, 17, T=T, C_=10.0, lam=0.25, precision=E.FP32S, init_mode=E.INIT_HYDROSTATIC,
                            seed=s, amp=1e-4) for T, s in ((40.0, 1), (60.0, 2), (80.0, 3))]
    with E.Batch(params) as bt:
        bt.init()
        for b in bt.boxes:
            b.hydrostatic_reference(check_every=500, tol=1e-9)
        res = bt.run_to_steady(check_every=500, tol=1e-5, max_steps=60000)
        for r in res:
            print(f"steps {r.steps} converged {r.converged} Nu {r.nusselt:.6f} delta {r.delta:.2e}")
            assert r.converged == 1 and r.steps < 60000
            assert abs(r.nusselt - 1.0) < 1e-3
        d = bt.diagnostics()
        assert all(x.step == max(r.steps for r in res) for x in d)


@pytest.mark.parametrize("flags", [E.BATCH_GRAPH | E.BATCH_TUNE, E.BATCH_GRAPH | E.BATCH_CHAINS])
def test_batch_with_a_256_wide_box_matches_single_boxes(flags):
    """The bench box's plane size (256 x 256: one LBM tile per x-row, the
    compile-time FFT plans of 256) in a batch beside a box of another plane
    size: every box as alone, bit for bit, with graph replay, with the tuner
    and with chains."""
    params = [E.make_params(256, 256, 8, T=400.0, C_=10.0, lam=0.25, precision=E.FP32S,
                            init_mode=E.INIT_HYDROSTATIC, seed=21, amp=1e-3),
              E.make_params(64, 64, 20, T=300.0, C_=5.0, lam=0.25, precision=E.FP32S, init_mode=E.INIT_HYDROSTATIC,
                            seed=22, amp=1e-3)]
    nsteps = (1, 9, 14)
    with E.Batch(params, flags=flags) as bt:
        bt.init()
        for n in nsteps:
            bt.step(n)
        got = [_state(b) for b in bt.boxes]
    for i, p in enumerate(params):
        with E.Box(p) as b:
            b.init()
            for n in nsteps:
                b.step(n)
            _assert_same_state(got[i], _state(b), f"box {i} ({p.nx}x{p.ny}x{p.nz})")
