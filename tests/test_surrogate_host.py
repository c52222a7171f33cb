"""Host-side pieces of the warm-start pipeline (no GPU): network inputs, the
scaling of the targets, and the inverse map from a network output to the
nine ehd_init fields (velocity, charge and potential recovered, E the
central-difference gradient of the potential with the wall copy rule)."""
import numpy as np

import ehd_amd as E
from ehd_amd import surrogate as S


def test_features():
    a = S.features((8, 6, 5), 400.0, 20.0, 10.0)
    assert a.shape == (6, 5, 6, 8)
    assert np.all(a[0] == 0.0) and np.allclose(a[1], np.log(2.0)) and np.all(a[2] == 0.0)
    assert a[3][0, 0, 1] == 1 / 8 and a[4][0, 1, 0] == 1 / 6 and a[5][4, 0, 0] == 1.0


def test_targets_and_warm_fields_round_trip():
    n = (8, 6, 7)
    p = S.box_params(n, 300.0, 10.0, init=E.INIT_FIELDS)
    rng = np.random.default_rng(0)
    U, Q, V = S.scales(p)
    qz = np.linspace(10.0, 4.0, n[2])
    phz = np.linspace(p.voltage, 0.0, n[2])
    shape = (n[2], n[1], n[0])
    f = {"ux": U * 0.01 * rng.standard_normal(shape), "uy": U * 0.01 * rng.standard_normal(shape),
         "uz": U * 0.01 * rng.standard_normal(shape),
         "q": qz[:, None, None] + Q * 0.01 * rng.standard_normal(shape),
         "phi": phz[:, None, None] + V * 1e-3 * rng.standard_normal(shape)}
    f["phi"][0] = p.voltage
    f["phi"][-1] = 0.0
    y = S.targets_from_fields(p, f, (qz, phz))
    w = S.fields9_from_prediction(p, y.astype(np.float64), (qz, phz))
    d = E.derive(p)
    assert np.allclose(w[3][1:-1], f["ux"][1:-1], rtol=1e-6, atol=1e-9 * U)
    assert np.allclose(w[5][1:-1], f["uz"][1:-1], rtol=1e-6, atol=1e-9 * U)
    assert np.allclose(w[1][1:-1], f["q"][1:-1], rtol=1e-6)
    assert np.allclose(w[2], f["phi"], rtol=1e-6, atol=1e-6 * V)
    assert np.all(w[3:6][:, 0] == 0) and np.all(w[3:6][:, -1] == 0) and np.all(w[0] == p.rho0)
    phi = w[2]
    assert np.allclose(w[6][:, :, 3], (phi[:, :, 2] - phi[:, :, 4]) / (2 * d.dx))
    assert np.allclose(w[8][3], (phi[2] - phi[4]) / (2 * d.dz))
    assert np.array_equal(w[8][0], w[8][1]) and np.array_equal(w[8][-1], w[8][-2])


def test_canonicalize_removes_shifts_and_xy_exchange():
    """Shifted and x/y-exchanged copies of one roll pattern map to the same
    canonical field."""
    Z, Y, X = 6, 16, 16
    z = np.arange(Z)[:, None, None] / (Z - 1)
    y = np.arange(Y)[None, :, None]
    x = np.arange(X)[None, None, :]
    t = np.zeros((5, Z, Y, X))
    t[2] = np.sin(np.pi * z) * np.cos(2 * np.pi * 2 * x / X + 0.3)
    t[0] = np.cos(np.pi * z) * np.sin(2 * np.pi * 2 * x / X + 0.3)
    t[3] = 0.1 * t[2] + 0.05 * np.cos(2 * np.pi * y / Y)
    ref = S.canonicalize(t)
    for sx, sy in ((3, 0), (5, 7), (11, 2)):
        u = np.roll(np.roll(t, sx, axis=3), sy, axis=2)
        c = S.canonicalize(u)
        # the pattern part (w, u) is canonical; the y-dependent part of q keeps its own y shift
        assert np.allclose(c[[0, 2]], ref[[0, 2]], atol=1e-12)
    sw = t.swapaxes(2, 3).copy()
    sw[[0, 1]] = sw[[1, 0]]
    c = S.canonicalize(sw)
    assert np.allclose(c[[0, 2]], ref[[0, 2]], atol=1e-12)
