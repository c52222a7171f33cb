"""Pins the FP64 CPU oracle (oracle/ehd_oracle.c) before it is trusted.

Each check comes from something other than the oracle's own formulas:
  * tables parsed from the reference source text (tests/golden/reference_tables.json,
    written by tests/golden/gen_reference_tables.py),
  * numpy's FFT, a dense brute-force solve and closed forms for the Poisson part,
  * lattice and moment identities for the equilibria, the Guo force and TRT,
  * tagged-population (pure index) tests for stream, walls and injection.
A planted mistake (a wrong sign, a swapped direction, a dropped term) breaks at
least one of them.
"""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = json.load(open(os.path.join(HERE, "golden", "reference_tables.json")))


def _box(nx=6, ny=4, nz=7, **kw):
    return O.Box(O.make_params(nx, ny, nz, **kw))


# --------------------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", list(range(1, 17)) + [31, 48, 61, 62, 80, 96, 122, 126, 127, 254, 256])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = O.fft1d(x, -1)
    i = O.fft1d(x, +1)
    scale = np.max(np.abs(np.fft.fft(x))) + 1e-300
    assert np.max(np.abs(f - np.fft.fft(x))) / scale < 5e-15
    assert np.max(np.abs(i - np.fft.ifft(x) * n)) / scale < 5e-15


def test_fft3d_matches_numpy():
    b = _box(6, 4, 5)
    ne = b.p.ne
    rng = np.random.default_rng(3)
    a = rng.standard_normal((ne, 4, 6)) + 1j * rng.standard_normal((ne, 4, 6))
    b[O.SEXT][..., 0] = a.real
    b[O.SEXT][..., 1] = a.imag
    b.fft3d(O.SEXT, O.FREQ, -1)
    got = b[O.FREQ][..., 0] + 1j * b[O.FREQ][..., 1]
    ref = np.fft.fftn(a)   # numpy's fftn over (z, y, x) = cufftPlan3d(NE, NY, NX) order
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-14


# --------------------------------------------------------------------------- parameters

def test_compute_parameters_reference_defaults():
    """LBM.h:54-62 constants give T~996, M=C=10, Fe=4000 (LBM.cu:1375-1378) and back."""
    p = O.make_params(122, 122, 101)
    p.K, p.eps, p.voltage, p.nu, p.rho0 = 2.5e-5, 1e-4, 1e4, 0.0251004, 1600.0
    p.charge0, p.diffu, p.lz = 10.0, 6.25e-5, 1.0
    T, M, C, Fe = O.compute_parameters(p)
    assert abs(T - 996.0000637) < 1e-6
    assert abs(M - 10.0) < 1e-12 and abs(C - 10.0) < 1e-12 and abs(Fe - 4000.0) < 1e-9
    q = O.make_params(122, 122, 101, T=T, M=M, C_=C, Fe=Fe)
    for name, ref in (("K", 2.5e-5), ("nu", 0.0251004), ("charge0", 10.0), ("diffu", 6.25e-5)):
        assert abs(getattr(q, name) / ref - 1) < 1e-14, name
    # grid-derived quantities of the runtime grid
    assert q.dz == 1.0 / 100 and q.dx == q.dz and abs(q.lx - 1.22) < 1e-15 and q.ne == 200
    assert q.dt == q.cfl * q.dx and abs(q.cs_square - 1.0 / 3.0 / 1e-4) < 1e-9


def test_trt_rates_identity():
    """(tau+ - 1/2)(tau- - 1/2) = Lambda for flow and charge (LBM.cu:299-302)."""
    b = _box(T=200.0, lam=0.25, lam_c=1e-6)
    op, om, ocp, ocm = b.rates()
    dt = b.p.dt
    tp, tm = 1 / (op * dt), 1 / (om * dt)
    tcp, tcm = 1 / (ocp * dt), 1 / (ocm * dt)
    assert abs((tp - 0.5) * (tm - 0.5) - 0.25) < 1e-12
    assert abs((tcp - 0.5) * (tcm - 0.5) - 1e-6) < 1e-15
    assert abs((tp - 0.5) - b.p.nu / b.p.cs_square / dt) < 1e-14
    assert abs((tcm - 0.5) - b.p.diffu / b.p.cs_square / dt) < 1e-15


# --------------------------------------------------------------------------- tables

def test_tables_match_reference_text():
    t = O.tables()
    c, opp, w = t["c"], t["opp"], t["w"]
    for d in range(1, 27):
        assert list(c[d]) == TABLES["stream_f"][str(d)], d
        assert list(c[d]) == TABLES["stream_h"][str(d)], d
        assert opp[d] == TABLES["bounce_bottom"][str(d)]
        assert opp[d] == TABLES["bounce_top"][str(d)][0]
        assert opp[d] == TABLES["inject"][str(d)][0]
        assert TABLES["h_top_copy"][str(d)] == d
        for key, mine in (("eq_collide_f", t["eqf"]), ("eq_init_f", t["eqf"]),
                          ("eq_collide_h", t["eqh"]), ("eq_init_h", t["eqh"])):
            ref = TABLES[key][str(d)]
            assert [v for v in mine[d] if v != 0] == ref, (key, d)
    wt = TABLES["weights"]
    cls = {0: "w0", 1: "ws", 2: "wa", 3: "wd"}
    for d in range(27):
        k = 0 if d == 0 else (1 if d <= 6 else (2 if d <= 18 else 3))
        a, bb = wt[cls[k]]
        assert w[d] == a / bb
        if d:
            assert TABLES["inject"][str(d)][1] == k
    assert TABLES["inject_rest_lines"] == 1
    for a, ax in enumerate("xyz"):
        pos, neg = TABLES["moments"][ax]
        assert list(t["mom"][a][:9]) == pos and list(t["mom"][a][9:]) == neg


def test_lattice_invariants():
    t = O.tables()
    c, w = t["c"].astype(float), t["w"]
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w @ c, 0, atol=1e-16)
    assert np.allclose((w[:, None, None] * c[:, :, None] * c[:, None, :]).sum(0), np.eye(3) / 3, atol=1e-16)
    for d in range(27):
        assert np.array_equal(c[t["opp"][d]], -c[d])


def test_moving_wall_signs_match_reference_text():
    """Top-plate terms of gpu_boundary (LBM.cu:1142-1171), by tagged populations."""
    b = _box(4, 3, 5, uw=0.37)
    rng = np.random.default_rng(1)
    b[O.F1][...] = rng.standard_normal(b[O.F1].shape)
    b[O.H2][...] = rng.standard_normal(b[O.H2].shape)
    b[O.F0BC][...] = rng.standard_normal(b[O.F0BC].shape)
    b.boundary()
    p = b.p
    multi = {1: 2.0 * p.rho0 * p.uw / p.cs_square * p.ws / p.cfl,
             2: 2.0 * p.rho0 * p.uw / p.cs_square * p.wa / p.cfl,
             3: 2.0 * p.rho0 * p.uw / p.cs_square * p.wd / p.cfl}
    nz = p.nz
    for d in range(1, 27):
        e, sg, cl = TABLES["bounce_top"][str(d)]
        assert np.array_equal(b[O.F2][d - 1, 0], b[O.F1][TABLES["bounce_bottom"][str(d)] - 1, 0])
        expect = b[O.F1][e - 1, nz - 1] + (sg * multi[cl] if sg else 0.0)
        assert np.array_equal(b[O.F2][d - 1, nz - 1], expect), d
        assert np.array_equal(b[O.H2][d - 1, nz - 1], b[O.H2][d - 1, nz - 2])
    assert np.array_equal(b[O.F0][0], b[O.F0BC][0]) and np.array_equal(b[O.F0][nz - 1], b[O.F0BC][1])
    assert np.array_equal(b[O.H0][nz - 1], b[O.H0][nz - 2])


def test_stream_is_the_pull_permutation():
    b = _box(5, 4, 6)
    tag = np.arange(b[O.F2].size, dtype=float).reshape(b[O.F2].shape)
    b[O.F2][...] = tag
    b[O.H2][...] = -tag
    b.stream()
    for d in range(1, 27):
        cx, cy, cz = TABLES["stream_f"][str(d)]
        ref = np.roll(tag[d - 1], shift=(cz, cy, cx), axis=(0, 1, 2))   # f1(x) = f2(x - c)
        assert np.array_equal(b[O.F1][d - 1], ref), d
        assert np.array_equal(b[O.H1][d - 1], -ref), d


def test_injection_bc():
    b = _box(4, 3, 5)
    rng = np.random.default_rng(2)
    b[O.H2][...] = rng.standard_normal(b[O.H2].shape)
    h0 = rng.standard_normal(b[O.H0].shape)
    b[O.H0][...] = h0
    b.bc_charge()
    p = b.p
    wcls = {1: p.ws, 2: p.wa, 3: p.wd}
    for d in range(1, 27):
        e, cl = TABLES["inject"][str(d)]
        assert np.array_equal(b[O.H1][d - 1, 0], -b[O.H2][e - 1, 0] + 2.0 * p.charge0 * wcls[cl])
    assert np.array_equal(b[O.H0][0], -h0[0] + 2.0 * p.charge0 * p.w0)
    assert np.array_equal(b[O.H0][1:], h0[1:])


def test_injection_wall_charge_has_period_two_memory():
    """A consequence of the injection rule (LBM.cu:1281-1353) and of a
    collision that conserves charge: the z=0 charge populations are rebuilt
    from their own post-collision values, so the wall charge follows
    q' = 2 q_inj - q and never forgets its start.  From the injected value it
    stays put; from q_inj + a it alternates q_inj - a, q_inj + a, ... (warm
    starts must therefore begin the wall at q_inj)."""
    ob = O.Box(O.make_params(4, 4, 8, T=200.0, M=10.0, C_=10.0, Fe=4000.0, lam=0.25))
    ob.init_hydrostatic(1, 1e-3, 0)
    qinj = ob[O.Q][0].copy()
    assert np.allclose(qinj, ob.p.charge0, rtol=1e-14)
    ob.init_equilibrium()
    ob.step(3)
    assert np.allclose(ob[O.Q][0], qinj, rtol=1e-12)
    ob.init_hydrostatic(1, 1e-3, 0)
    ob[O.Q][0] += 0.3 * qinj
    ob.init_equilibrium()
    seq = []
    for _ in range(6):
        ob.step(1)
        seq.append(ob[O.Q][0].copy())
    for k in range(6):
        assert np.allclose(seq[k], qinj * (0.7 if k % 2 else 1.3), rtol=1e-12)


# --------------------------------------------------------------------------- collision

def _xi(p):
    c = O.tables()["c"].astype(float)
    return c / p.cfl, O.tables()["w"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_equilibrium_moments(seed):
    p = O.make_params(4, 4, 5, T=400.0)
    rng = np.random.default_rng(seed)
    rho = 1600 * (1 + 0.01 * rng.standard_normal())
    u = 0.3 * rng.standard_normal(3)
    q = 5 + rng.standard_normal()
    E = 1e4 * rng.standard_normal(3)
    fe, he, _ = O.node_terms(p, rho, *u, q, *E, 0.0, 0.0, 0.0)
    xi, w = _xi(p)
    cs2 = p.cs_square
    for pops, dens, vel in ((fe, rho, u), (he, q, u + p.K * E)):
        assert abs(pops.sum() / dens - 1) < 1e-14
        assert np.allclose(xi.T @ pops, dens * vel, rtol=1e-12, atol=1e-12 * dens)
        m2 = (pops[:, None, None] * xi[:, :, None] * xi[:, None, :]).sum(0)
        assert np.allclose(m2, dens * (cs2 * np.eye(3) + np.outer(vel, vel)), rtol=1e-12, atol=1e-10 * dens)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_guo_force_closed_form(seed):
    p = O.make_params(4, 4, 5)
    rng = np.random.default_rng(seed)
    u = 0.5 * rng.standard_normal(3)
    F = 1e5 * rng.standard_normal(3)
    _, _, fp = O.node_terms(p, 1600.0, *u, 1.0, 0, 0, 0, *F)
    xi, w = _xi(p)
    cs2 = p.cs_square
    ref = w * (((xi - u) @ F) / cs2 + (xi @ u) * (xi @ F) / cs2 ** 2)
    assert np.max(np.abs(fp - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert abs(fp.sum()) < 1e-12 * np.abs(F).max()
    assert np.allclose(xi.T @ fp, F, rtol=1e-12)
    m2 = (fp[:, None, None] * xi[:, :, None] * xi[:, None, :]).sum(0)
    assert np.allclose(m2, np.outer(u, F) + np.outer(F, u), rtol=1e-10, atol=1e-10 * np.abs(F).max())


def _random_state(b, rng, amp=1e-3):
    """Populations near equilibrium with random deviations, E random."""
    p = b.p
    nz, ny, nx = b.shape
    w = O.tables()["w"]
    b[O.F0][...] = w[0] * p.rho0 * (1 + amp * rng.standard_normal((nz, ny, nx)))
    b[O.H0][...] = w[0] * 5.0 * (1 + 0.1 * rng.standard_normal((nz, ny, nx)))
    for d in range(1, 27):
        b[O.F1][d - 1] = w[d] * p.rho0 * (1 + amp * rng.standard_normal((nz, ny, nx)))
        b[O.H1][d - 1] = w[d] * 5.0 * (1 + 0.1 * rng.standard_normal((nz, ny, nx)))
    for k in (O.EX, O.EY, O.EZ):
        b[k][...] = 1e4 * rng.standard_normal((nz, ny, nx))


def test_collision_conserves_mass_charge_and_adds_dt_force():
    b = _box(5, 4, 6, exf=3.0)
    rng = np.random.default_rng(5)
    _random_state(b, rng)
    f_pre = np.concatenate([b[O.F0][None], b[O.F1]])
    h_pre = np.concatenate([b[O.H0][None], b[O.H1]])
    E = np.stack([b[O.EX], b[O.EY], b[O.EZ]])
    b.collide_save()
    f_post = np.concatenate([b[O.F0][None], b[O.F2]])
    h_post = np.concatenate([b[O.H0][None], b[O.H2]])
    xi, _ = _xi(b.p)
    sl = slice(1, -1)   # z=0 carries the velocity override
    assert np.max(np.abs(f_post.sum(0) - f_pre.sum(0))[sl]) < 1e-12 * b.p.rho0
    assert np.max(np.abs(h_post.sum(0) - h_pre.sum(0))[sl]) < 1e-13 * 5.0
    q = h_pre.sum(0)
    F = q[None] * E
    F[0] += b.p.exf
    j_pre = np.einsum("dk,dzyx->kzyx", xi, f_pre)
    j_post = np.einsum("dk,dzyx->kzyx", xi, f_post)
    dj = (j_post - j_pre)[:, sl]
    assert np.max(np.abs(dj - b.p.dt * F[:, sl])) < 1e-9 * np.max(np.abs(b.p.dt * F))


def test_wall_velocity_override():
    """z=0 velocity = -(momentum + F dt/2 at z=1, pre-collision) / rho(z=0) (LBM.cu:416-493)."""
    b = _box(4, 3, 6)
    rng = np.random.default_rng(9)
    _random_state(b, rng)
    f_pre = np.concatenate([b[O.F0][None], b[O.F1]])
    h_pre = np.concatenate([b[O.H0][None], b[O.H1]])
    E1 = np.stack([b[O.EX][1], b[O.EY][1], b[O.EZ][1]])
    b.collide_save()
    xi, _ = _xi(b.p)
    rho_wall = f_pre[:, 0].sum(0)
    q1 = h_pre[:, 1].sum(0)
    j1 = np.einsum("dk,dyx->kyx", xi, f_pre[:, 1])
    ref = -(j1 + q1[None] * E1 * b.p.dt * 0.5) / rho_wall[None]
    got = np.stack([b[O.UX][0], b[O.UY][0], b[O.UZ][0]])
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def _trt_taus(p):
    """Relaxation times the TRT scheme must have for the box's transport
    coefficients (Ginzburg's TRT: the even rate sets the shear viscosity
    nu = cs^2 (tau+ - 1/2) dt, the odd rate of the charge set sets its
    diffusivity D = cs^2 (tau_c- - 1/2) dt, and the magic parameter fixes the
    other rate, Lambda = (tau+ - 1/2)(tau- - 1/2); LBM.h:79-80)."""
    tp = p.nu / (p.cs_square * p.dt) + 0.5
    tcm = p.diffu / (p.cs_square * p.dt) + 0.5
    return dict(fp=tp, fm=0.5 + p.V / (tp - 0.5), cm=tcm, cp=0.5 + p.VC / (tcm - 0.5))


def _trt_moment_residuals(b):
    """Collide the box once and measure how each moment relaxed, against the
    closed forms of a TRT collision with the relaxation times of _trt_taus:
      sum xi h:        d = -(m - q v)/tau_c-           (v = u + K E)
      sum xi xi h:     d = -(m - q (cs^2 I + v v))/tau_c+
      sum xi_x xi_y xi_z h:  d = -m/tau_c-          (zero in the equilibrium)
      sum xi xi f:     d = -(m - rho (cs^2 I + u u))/tau+ + dt (1 - 1/(2 tau+)) (u F + F u)
      sum xi_x xi_y xi_z f:  d = -m/tau-            (zero in equilibrium and force)
    on interior nodes, u = (sum xi f + F dt/2)/rho, F = q E + exf x.  Returns
    the largest residual of each, relative to the largest change it predicts."""
    p = b.p
    xi, _ = _xi(p)
    tau = _trt_taus(p)
    f_pre = np.concatenate([b[O.F0][None], b[O.F1]])[:, 1:-1]
    h_pre = np.concatenate([b[O.H0][None], b[O.H1]])[:, 1:-1]
    E = np.stack([b[O.EX], b[O.EY], b[O.EZ]])[:, 1:-1]
    b.collide_save()
    f_post = np.concatenate([b[O.F0][None], b[O.F2]])[:, 1:-1]
    h_post = np.concatenate([b[O.H0][None], b[O.H2]])[:, 1:-1]
    rho, q = f_pre.sum(0), h_pre.sum(0)
    F = q[None] * E
    F[0] += p.exf
    u = (np.einsum("dk,dzyx->kzyx", xi, f_pre) + 0.5 * p.dt * F) / rho[None]
    v = u + p.K * E
    eye = np.eye(3)[:, :, None, None, None]
    x3 = xi[:, 0] * xi[:, 1] * xi[:, 2]

    def m1(a): return np.einsum("dk,dzyx->kzyx", xi, a)
    def m2(a): return np.einsum("dk,dl,dzyx->klzyx", xi, xi, a)
    def m3(a): return np.einsum("d,dzyx->zyx", x3, a)

    def resid(delta, want):
        return float(np.max(np.abs(delta - want)) / np.max(np.abs(want)))

    uu = np.einsum("kzyx,lzyx->klzyx", u, u)
    vv = np.einsum("kzyx,lzyx->klzyx", v, v)
    uF = np.einsum("kzyx,lzyx->klzyx", u, F)
    cs2 = p.cs_square
    return {
        "h1": resid(m1(h_post) - m1(h_pre), -(m1(h_pre) - q[None] * v) / tau["cm"]),
        "h2": resid(m2(h_post) - m2(h_pre), -(m2(h_pre) - q * (cs2 * eye + vv)) / tau["cp"]),
        "h3": resid(m3(h_post) - m3(h_pre), -m3(h_pre) / tau["cm"]),
        "f2": resid(m2(f_post) - m2(f_pre), -(m2(f_pre) - rho * (cs2 * eye + uu)) / tau["fp"]
                    + p.dt * (1 - 0.5 / tau["fp"]) * (uF + uF.transpose(1, 0, 2, 3, 4))),
        "f3": resid(m3(f_post) - m3(f_pre), -m3(f_pre) / tau["fm"]),
    }


@pytest.mark.parametrize("kw", [dict(), dict(lam=1.0 / 12.0, exf=3.0), dict(lam_c=0.25, Fe=400.0)])
def test_trt_relaxation_rates_closed_form(kw):
    """Each moment of the collision relaxes with the rate the transport
    coefficients fix: h's flux and odd moments with tau_c- (diffusivity),
    h's second moment with tau_c+ (magic parameter Lambda_c), f's stress with
    tau+ (viscosity), f's odd third moment with tau- (Lambda)."""
    b = _box(6, 5, 7, **kw)
    _random_state(b, np.random.default_rng(17))
    r = _trt_moment_residuals(b)
    assert max(r.values()) < 1e-9, r


@pytest.mark.parametrize("swap", ["charge", "flow"])
def test_trt_moment_checks_catch_swapped_rates(swap):
    """A planted mistake in the oracle (on the CPU): swapping the even and odd
    rates of one population set -- the same as swapping their uses in every
    relaxation line of the collision -- must fail the closed forms above."""
    b = _box(6, 5, 7)
    _random_state(b, np.random.default_rng(17))
    r = b.rates()
    r = r[[1, 0, 2, 3]] if swap == "flow" else r[[0, 1, 3, 2]]
    b.set_rates(r)
    res = _trt_moment_residuals(b)
    keys = ("f2", "f3") if swap == "flow" else ("h1", "h2", "h3")
    assert min(res[k] for k in keys) > 1e-3, res


# --------------------------------------------------------------------------- Poisson

def _numpy_poisson(q, p):
    """The reference's algorithm (poisson.cu:78-167) in numpy: odd extension, fftn,
    divide by mu, ifftn, real part / size."""
    nz, ny, nx = q.shape
    ne = 2 * (nz - 1)
    s = np.zeros((ne, ny, nx), complex)
    s[1:nz - 1] = -q[1:nz - 1] / p.eps
    s[1] -= p.voltage / p.dz / p.dz
    s[nz:ne - 1] = q[ne - np.arange(nz, ne - 1)][:, :, :] / p.eps
    s[ne - 1] = q[1] / p.eps + p.voltage / p.dz / p.dz
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=1.0 / nx) / p.lx
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=1.0 / ny) / p.ly
    kz = 2 * np.pi * np.fft.fftfreq(ne, d=1.0 / ne) / (ne * p.dz)
    mu = (4 / p.dz ** 2) * np.sin(kz * p.dz / 2)[:, None, None] ** 2 + kx[None, None, :] ** 2 + ky[None, :, None] ** 2
    mu[0, 0, 0] = 1.0
    phi_e = np.fft.ifftn(-np.fft.fftn(s) / mu).real
    phi = phi_e[:nz].copy()
    phi[0] = p.voltage
    phi[nz - 1] = 0.0
    return phi


@pytest.mark.parametrize("shape", [(6, 4, 6), (8, 6, 7), (12, 10, 9)])
def test_poisson_matches_numpy_and_tridiagonal(shape):
    nx, ny, nz = shape
    b = _box(nx, ny, nz)
    rng = np.random.default_rng(nx)
    q = 10 * rng.random((nz, ny, nx))
    b[O.Q][...] = q
    b.poisson()
    phi_fft = b[O.PHI].copy()
    ref = _numpy_poisson(q, b.p)
    assert np.max(np.abs(phi_fft - ref)) / np.max(np.abs(ref)) < 1e-13
    b.poisson_tridiag()
    assert np.max(np.abs(b[O.PHI] - phi_fft)) / np.max(np.abs(ref)) < 1e-13


def test_poisson_zero_charge_is_linear():
    b = _box(6, 4, 9)
    b[O.Q][...] = 0.0
    b.poisson()
    z = np.arange(9)
    lin = b.p.voltage * (1 - z / 8.0)
    assert np.max(np.abs(b[O.PHI] - lin[:, None, None])) < 1e-10 * b.p.voltage


def test_poisson_single_mode_closed_form():
    nx, ny, nz = 8, 6, 9
    b = _box(nx, ny, nz)
    p = b.p
    z = np.arange(nz)[:, None, None]
    x = np.arange(nx)[None, None, :]
    m, kxi = 2, 1
    q = 3.0 * np.sin(np.pi * m * z / (nz - 1)) * np.cos(2 * np.pi * kxi * x / nx) * np.ones((nz, ny, nx))
    b[O.Q][...] = q
    b.poisson()
    kx = 2 * np.pi * kxi / p.lx
    lam = 4 / p.dz ** 2 * np.sin(np.pi * m / (2 * (nz - 1))) ** 2 + kx ** 2
    ref = p.voltage * (1 - z / (nz - 1.0)) + (q / p.eps) / lam
    assert np.max(np.abs(b[O.PHI] - ref)) / np.max(np.abs(ref)) < 1e-13


def test_poisson_dense_bruteforce():
    """Dense solve of the discrete operator (spectral x,y; second differences in z)."""
    nx, ny, nz = 4, 4, 6
    b = _box(nx, ny, nz)
    p = b.p
    rng = np.random.default_rng(11)
    q = rng.random((nz, ny, nx))
    b[O.Q][...] = q
    b.poisson()

    def d2(n, L):
        F = np.fft.fft(np.eye(n), axis=0)
        k = 2 * np.pi * np.fft.fftfreq(n, d=1.0 / n) / L
        return (np.linalg.inv(F) @ np.diag(-k ** 2) @ F).real

    nI = nz - 2
    Dz = (np.diag(-2 * np.ones(nI)) + np.diag(np.ones(nI - 1), 1) + np.diag(np.ones(nI - 1), -1)) / p.dz ** 2
    A = (np.kron(np.kron(Dz, np.eye(ny)), np.eye(nx)) + np.kron(np.kron(np.eye(nI), d2(ny, p.ly)), np.eye(nx))
         + np.kron(np.kron(np.eye(nI), np.eye(ny)), d2(nx, p.lx)))
    rhs = (-q[1:nz - 1] / p.eps).reshape(-1).copy()
    rhs[:nx * ny] -= p.voltage / p.dz ** 2      # phi(0) = V moved to the right-hand side
    sol = np.linalg.solve(A, rhs).reshape(nI, ny, nx)
    assert np.max(np.abs(b[O.PHI][1:nz - 1] - sol)) / np.max(np.abs(sol)) < 1e-12


def test_efield_closed_forms():
    nx, ny, nz = 8, 6, 7
    b = _box(nx, ny, nz)
    p = b.p
    z = np.arange(nz)[:, None, None]
    x = np.arange(nx)[None, None, :]
    b[O.PHI][...] = p.voltage * (1 - z / (nz - 1.0)) + 5.0 * np.sin(2 * np.pi * x / nx)
    b.efield()
    assert np.allclose(b[O.EZ], p.voltage / ((nz - 1) * p.dz), rtol=1e-12)
    assert np.max(np.abs(b[O.EY])) < 1e-9
    ref_ex = 0.5 * 5.0 * (np.sin(2 * np.pi * (x - 1) / nx) - np.sin(2 * np.pi * (x + 1) / nx)) / p.dx
    assert np.allclose(b[O.EX], np.broadcast_to(ref_ex, b[O.EX].shape), rtol=1e-12, atol=1e-9)


def test_current_closed_form():
    nx, ny, nz = 6, 4, 7
    b = _box(nx, ny, nz)
    p = b.p
    rng = np.random.default_rng(4)
    b[O.Q][...] = rng.random((nz, ny, nx))
    b[O.EZ][...] = rng.random((nz, ny, nx))
    qtop = 2 * b[O.Q][nz - 2] - b[O.Q][nz - 3]
    ref = (qtop * b[O.EZ][nz - 1]).sum() * p.K * p.dy * p.dy
    assert abs(b.current() - ref) < 1e-13 * abs(ref)


# --------------------------------------------------------------------------- initial state

@pytest.mark.parametrize("C", [5.0, 10.0, 20.0])
def test_hydrostatic_profile(C):
    e0 = O.hydrostatic_e0(C)
    zeta = np.linspace(0, 1, 20001)
    E = np.sqrt(e0 ** 2 + 2 * C * e0 * zeta)
    integral = np.trapezoid(E, zeta)
    assert abs(integral - 1) < 1e-8                       # phi(0) - phi(1) = 1
    q = C * e0 / E
    dEdz = np.gradient(E, zeta)
    assert np.max((np.abs(dEdz - q) / q)[50:-5]) < 1e-4           # Gauss: dE/dz = q
    assert np.allclose(q * E, C * e0)                      # uniform current
    if C == 10.0:
        assert abs(e0 - 0.110742) < 5e-6


def test_hash_perturbation():
    vals = np.array([O.hash_uniform(1, x, y, z, c) for x in range(8) for y in range(8)
                     for z in range(4) for c in range(3)])
    assert np.all(vals >= -1) and np.all(vals < 1)
    assert abs(vals.mean()) < 0.1 and abs(vals.std() - 1 / np.sqrt(3)) < 0.05
    assert O.hash_uniform(1, 3, 4, 5, 2) == O.hash_uniform(1, 3, 4, 5, 2)
    assert O.hash_uniform(1, 3, 4, 5, 2) != O.hash_uniform(2, 3, 4, 5, 2)


@pytest.mark.parametrize("pattern", [1, 2, 3, 4])
def test_perturbation_patterns_divergence_free(pattern):
    """Every pattern of LBM.cu:389-411 is an incompressible cell: div u = 0
    (central differences of the analytic field), u = 0 on both walls, and the
    roll/square/cross fields repeat with the box width.  A wrong sign, a
    swapped component or a wrong wavenumber breaks one of these."""
    lx, ly, lz, amp = 1.22, 1.22 * (np.sqrt(3.0) if pattern == 4 else 1.0), 1.0, 1e-3
    rng = np.random.default_rng(pattern)
    h = 1e-5
    scale = 0.0
    for _ in range(40):
        x, y, z = rng.uniform(0, lx), rng.uniform(0, ly), rng.uniform(0.05, 0.95) * lz
        u = lambda a, b, c: O.pattern_velocity(pattern, amp, a, b, c, lx, ly, lz)
        div = ((u(x + h, y, z)[0] - u(x - h, y, z)[0]) + (u(x, y + h, z)[1] - u(x, y - h, z)[1])
               + (u(x, y, z + h)[2] - u(x, y, z - h)[2])) / (2 * h)
        grad = np.abs(u(x, y, z + h)[2] - u(x, y, z - h)[2]) / (2 * h)
        scale = max(scale, grad)
        assert abs(div) < 1e-7 * amp, (pattern, div)
        for zw in (0.0, lz):
            assert np.allclose(u(x, y, zw), 0.0, atol=1e-15)
        assert np.allclose(u(x + lx, y, z), u(x, y, z), atol=1e-13)
        assert np.allclose(u(x, y + ly, z), u(x, y, z), atol=1e-13)
    assert scale > 1e-3 * amp   # the field is not trivially zero


def test_crossed_rolls_reference_form():
    """Pattern 3 equals the active perturbation of LBM.cu:400-402 with LL = Lx
    (LBM.h:39 sets LL = 1.22, the width of the 122-node box)."""
    LL = 1.22
    for x, y, z in ((0.1, 0.3, 0.2), (0.9, 1.1, 0.7), (0.5, 0.05, 0.5)):
        u = O.pattern_velocity(3, 0.001, x, y, z, LL, LL, 1.0)
        uz = (np.cos(2 * np.pi * z) - 1) * (np.cos(2 * np.pi / LL * x) + np.cos(2 * np.pi / LL * y)) * 0.001
        ux = LL * np.sin(2 * np.pi * z) * np.sin(2 * np.pi / LL * x) * 0.001
        uy = LL * np.sin(2 * np.pi * z) * np.sin(2 * np.pi / LL * y) * 0.001
        assert np.allclose(u, [ux, uy, uz], rtol=1e-14, atol=1e-18)


def test_init_hydrostatic_state():
    b = _box(6, 4, 11, C_=10.0)
    b.init_hydrostatic(3, 1e-3, hydro_density=1)
    p = b.p
    assert np.all(b[O.PHI][0] == p.voltage) and np.all(b[O.PHI][-1] == 0.0)
    assert np.allclose(b[O.Q][0], p.charge0, rtol=1e-14)
    assert np.all(b[O.UX][0] == 0) and np.all(b[O.UZ][-1] == 0)
    assert 0 < np.max(np.abs(b[O.UX][1:-1])) <= 1e-3
    # hydrostatic pressure: dp/dz = q Ez (uniform), mean density rho0
    assert abs(b[O.RHO].mean() - p.rho0) < 1e-9 * p.rho0
    dp = np.diff(b[O.RHO][:, 0, 0]) * p.cs_square / p.dz
    qE = p.charge0 * p.voltage / p.lz * O.hydrostatic_e0(10.0)   # q(0) E(0)
    assert np.allclose(dp, qE, rtol=1e-9)
    b.init_equilibrium()
    f = np.concatenate([b[O.F0][None], b[O.F1]])
    assert np.allclose(f.sum(0), b[O.RHO], rtol=1e-14)
    b.init_hydrostatic(3, 1e-3, hydro_density=0)
    assert np.all(b[O.RHO] == p.rho0)


def test_coupled_step_physics_sanity():
    """Unperturbed hydrostatic start: u stays laterally uniform and tiny, charge
    stays positive, current positive."""
    b = _box(8, 6, 12, T=200.0)
    b.init_hydrostatic(1, 0.0)
    b.init_equilibrium()
    b.step(50)
    assert np.max(np.abs(b[O.UX][1:-1])) < 1e-9
    assert np.max(np.abs(np.diff(b[O.Q][1:-1], axis=2))) < 1e-9
    assert np.min(b[O.Q][1:-1]) > 0 and b.current() > 0


def test_force_population_structure_matches_reference_text():
    """Term order and signs of every Guo force population (LBM.cu:671-702), as parsed
    from the source text, follow the rule the oracle evaluates: for each nonzero axis a
    of c_d a term (c_a/CFL - u_a + P_a cflinv2) F_a with P_a = sum_b c_a c_b u_b over the
    nonzero axes in x,y,z order, and -u_b F_b for the zero axes; one-axis directions
    put the -u_b F_b terms first, the others put them last."""
    c = O.tables()["c"]
    for d in range(27):
        ref = TABLES["fpop"][str(d)]
        cls = 0 if d == 0 else (1 if d <= 6 else (2 if d <= 18 else 3))
        assert ref["coe"] == cls and ref["neg"] == (d == 0)
        if d == 0:
            assert [(t["kind"], t["sign"], t["axis"]) for t in ref["terms"]] == [("U", 1, 1), ("U", 1, 2), ("U", 1, 3)]
            continue
        nz_axes = [a for a in range(3) if c[d][a] != 0]
        zero_axes = [a for a in range(3) if c[d][a] == 0]
        a_terms = [{"kind": "A", "sign": 1, "axis": a + 1, "head": int(c[d][a]),
                    "P": [int(c[d][a] * c[d][b]) * (b + 1) for b in nz_axes]} for a in nz_axes]
        u_terms = [{"kind": "U", "sign": -1, "axis": a + 1} for a in zero_axes]
        expect = (u_terms + a_terms) if len(nz_axes) == 1 else (a_terms + u_terms)
        assert ref["terms"] == expect, d
