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
    assert np.all(w[1][0] == qz[0])          # the injection wall starts at the injected charge
    assert np.allclose(w[2], f["phi"], rtol=1e-6, atol=1e-6 * V)
    assert np.all(w[3:6][:, 0] == 0) and np.all(w[3:6][:, -1] == 0) and np.all(w[0] == p.rho0)
    phi = w[2]
    assert np.allclose(w[6][:, :, 3], (phi[:, :, 2] - phi[:, :, 4]) / (2 * d.dx))
    assert np.allclose(w[8][3], (phi[2] - phi[4]) / (2 * d.dz))
    assert np.array_equal(w[8][0], w[8][1]) and np.array_equal(w[8][-1], w[8][-2])


def test_canonicalize_removes_shifts_and_xy_exchange():
    """Shifted (by whole and fractional nodes) and x/y-exchanged copies of one
    roll pattern map to the same canonical field, with the roll's upflow at
    x = 0."""
    Z, Y, X = 6, 16, 16
    z = np.arange(Z)[:, None, None] / (Z - 1)
    x = np.arange(X)[None, None, :]
    y = np.arange(Y)[None, :, None]

    def rolls(x0):
        t = np.zeros((5, Z, Y, X))
        a = 2 * np.pi * 2 * (x - x0) / X
        t[2] = np.sin(np.pi * z) * np.cos(a) + 0.1 * np.sin(np.pi * z) * np.cos(2 * a)
        t[0] = np.cos(np.pi * z) * np.sin(a)
        t[3] = 0.1 * t[2]
        t[4] = 0.01 * np.sin(np.pi * z) * np.cos(a) + 0 * y
        return t
    ref = S.canonicalize(rolls(0.0))
    assert np.allclose(ref, rolls(0.0), atol=1e-12)
    for x0 in (3.0, 5.25, 11.7, -0.4):
        assert np.allclose(S.canonicalize(rolls(x0)), ref, atol=1e-12)
    sw = rolls(2.3).swapaxes(2, 3).copy()
    sw[[0, 1]] = sw[[1, 0]]
    assert np.allclose(S.canonicalize(sw), ref, atol=1e-12)


def test_canonicalize_cells_both_axes():
    """Square cells (sum and product forms) and hexagon-like cells shifted
    along x and y by any amounts (the position a noise start picks) map to
    one canonical field."""
    Z, Y, X = 5, 12, 12
    z = np.arange(Z)[:, None, None] / (Z - 1)
    x = np.arange(X)[None, None, :]
    y = np.arange(Y)[None, :, None]

    def cells(kind, x0, y0):
        ax, ay = 2 * np.pi * (x - x0) / X, 2 * np.pi * (y - y0) / Y
        t = np.zeros((5, Z, Y, X))
        if kind == "sum":
            w = np.cos(ax) + np.cos(ay)
        elif kind == "product":
            w = np.cos(ax) * np.cos(ay)
        else:
            w = np.cos(2 * ax) + 0.6 * np.cos(ax + ay) + 0.3 * np.cos(ax - 2 * ay)
        t[2] = np.sin(np.pi * z) * w
        t[0] = 0.5 * np.cos(np.pi * z) * (np.roll(w, -1, axis=2) - np.roll(w, 1, axis=2))
        t[1] = 0.5 * np.cos(np.pi * z) * (np.roll(w, -1, axis=1) - np.roll(w, 1, axis=1))
        t[3] = 0.2 * t[2] + 0.05 * np.sin(np.pi * z)
        t[4] = 0.01 * t[2]
        return t
    for kind in ("sum", "product", "mixed"):
        ref = S.canonicalize(cells(kind, 0.3, -1.1))
        for x0, y0 in ((1.0, 7.0), (4.6, 2.2), (10.1, 11.9)):
            assert np.allclose(S.canonicalize(cells(kind, x0, y0)), ref, atol=1e-10), kind
