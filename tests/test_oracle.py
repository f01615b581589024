"""Pins of the FP64 oracle: closed forms, finite-difference identities,
conservation laws, and the reference's own known-answer case (two-particle
FIRE, tests/run_test/run_test.cpp:98-152 in the reference).  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import lattice_frame, uniform_frame

RSTAR = 2.0 ** (1.0 / 6.0)
VC25 = 4.0 * (2.5 ** -12 - 2.5 ** -6)   # -0.016316891136...


def test_shift_constant_value():
    # Reference value quoted for c = 2.5: Vc = -0.016316891136 eps.
    assert VC25 == pytest.approx(-0.016316891136, abs=1e-12)


@pytest.mark.parametrize("ti,tj,eps_key", [(0, 0, "eps_aa"), (1, 1, "eps_bb"), (0, 1, "eps_ab"),
                                            (1, 0, "eps_ab"), (2, 2, "eps_ab"), (0, 2, "eps_ab")])
def test_pair_minimum_and_type_selection(ti, tj, eps_key):
    p = O.make_params(10, 10, eps_aa=1.0, eps_bb=0.7, eps_ab=0.3)
    eps = getattr(p, eps_key)
    fx, fy, u, inc = O.pair(RSTAR, 0.0, ti, tj, p)
    assert inc
    assert abs(fx) < 1e-12 and fy == 0.0
    # shifted minimum: -eps - eps*Vc
    assert u == pytest.approx(-eps - eps * VC25, rel=1e-14)


def test_pair_sigma_mixing_and_zero_at_sigma():
    p = O.make_params(10, 10, sigma_aa=1.0, sigma_bb=0.8, sigma_ab=1.3, eps_ab=1.0)
    # u(sigma_ab) = 0 before the shift; F(r*) = 0 at r* = 2^(1/6) sigma_ab
    _, _, u, _ = O.pair(1.3, 0.0, 0, 1, p)
    assert u == pytest.approx(-VC25, rel=1e-13)
    fx, _, _, _ = O.pair(1.3 * RSTAR, 0.0, 0, 1, p)
    assert abs(fx) < 1e-12
    # cutoff scales with sigma: 2.5*0.8 = 2.0 for BB
    assert O.pair(1.999, 0, 1, 1, p)[3] and not O.pair(2.0, 0, 1, 1, p)[3]
    assert not O.pair(0.0, 0.0, 0, 0, p)[3]   # r2 == 0 excluded


def test_force_is_minus_energy_gradient():
    p = O.make_params(10, 10, eps_ab=0.5)
    for r in (0.95, 1.05, RSTAR, 1.5, 2.0, 2.4):
        for ang in (0.0, 0.7, 2.1):
            dx, dy = r * math.cos(ang), r * math.sin(ang)
            fx, fy, _, _ = O.pair(dx, dy, 0, 1, p)
            h = 1e-6
            gx = (O.pair(dx + h, dy, 0, 1, p)[2] - O.pair(dx - h, dy, 0, 1, p)[2]) / (2 * h)
            gy = (O.pair(dx, dy + h, 0, 1, p)[2] - O.pair(dx, dy - h, 0, 1, p)[2]) / (2 * h)
            # u is stored per ordered pair; F_i = -du/d(r_i)
            assert fx == pytest.approx(-gx, rel=1e-7, abs=1e-8)
            assert fy == pytest.approx(-gy, rel=1e-7, abs=1e-8)


def test_pbc_wrap_cases():
    assert O.pbc_wrap(-0.5, 10.0) == 9.5
    assert O.pbc_wrap(25.0, 10.0) == 5.0
    assert O.pbc_wrap(10.0, 10.0) == 0.0
    assert O.pbc_wrap(3.0, 0.0) == 3.0     # L <= 0: identity
    assert O.pbc_wrap(-3.0, -1.0) == -3.0


def _sim(frame, Lx, Ly, **kw):
    return O.OracleSim(O.make_params(Lx, Ly, **kw), frame)


def test_newton_third_law_and_cell_equals_allpairs():
    f, Lx, Ly = lattice_frame(900, 0.8, seed=3, jitter=0.2, aspect=2.0)
    s = _sim(f, Lx, Ly, eps_ab=0.5, force_mode=1)
    acc_c, u_c, w_c, l_c = s.pair_pass(True)
    s.set_force_mode(0)
    acc_a, u_a, w_a, l_a = s.pair_pass(True)
    assert np.abs(acc_c - acc_a).max() <= 1e-12 * np.abs(acc_a).max()
    assert u_c.sum() == pytest.approx(u_a.sum(), rel=1e-13)
    assert w_c.sum() == pytest.approx(w_a.sum(), rel=1e-13)
    F = acc_a  # masses 1
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * np.abs(F).max()


@pytest.mark.parametrize("Lx,Ly,periodic_y", [(6.0, 20.0, True), (7.4, 7.4, True), (30.0, 14.0, False),
                                              (13.0, 4.0, False)])
def test_cell_equals_allpairs_small_and_open_boxes(Lx, Ly, periodic_y):
    rng = np.random.default_rng(11)
    n = int(0.7 * Lx * Ly)
    f = uniform_frame(n, Lx, Ly, seed=5)
    if not periodic_y:   # piston mode: some particles outside [0, Ly)
        f["pos"][: n // 10, 1] += rng.uniform(-1.0, 1.0, n // 10)
    # separate overlapping pairs to keep forces moderate
    s = _sim(f, Lx, Ly, force_mode=1)
    acc_c, u_c, _, _ = s.pair_pass(periodic_y)
    s.set_force_mode(0)
    acc_a, u_a, _, _ = s.pair_pass(periodic_y)
    scale = np.abs(acc_a).max()
    assert np.abs(acc_c - acc_a).max() <= 1e-12 * scale
    assert u_c.sum() == pytest.approx(u_a.sum(), rel=1e-12)


def test_cell_list_canonical_properties():
    f, Lx, Ly = lattice_frame(2000, 0.75, seed=9, aspect=2.0)
    s = _sim(f, Lx, Ly)
    ncx, ncy, keys, start, srt = s.cell_list(True)
    side = 2.5
    assert ncx == max(3, math.floor(Lx / side)) and ncy == max(3, math.floor(Ly / side))
    x, y = f["pos"][:, 0], f["pos"][:, 1]
    cx = np.minimum(ncx - 1, np.floor(x * ncx / Lx)).astype(np.int64)
    cy = np.minimum(ncy - 1, np.floor(y * ncy / Ly)).astype(np.int64)
    assert np.array_equal(keys, (cy * ncx + cx).astype(np.uint32))
    assert start[0] == 0 and start[-1] == len(f) and np.all(np.diff(start.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(srt), np.arange(len(f)))
    for c in range(ncx * ncy):
        ids = srt[start[c]:start[c + 1]]
        assert np.all(np.diff(ids.astype(np.int64)) > 0)
        assert np.all(keys[ids] == c)


def _scaled(frame, sx, sy):
    g = frame.copy()
    g["pos"][:, 0] *= sx
    g["pos"][:, 1] *= sy
    return g


def test_virial_equals_minus_volume_derivative():
    f, Lx, Ly = lattice_frame(600, 0.8, seed=21, jitter=0.15)
    s = _sim(f, Lx, Ly, eps_ab=0.5)
    _, _, w, _ = s.pair_pass(True)
    W = w.sum()
    d = 1e-6
    up = _sim(_scaled(f, 1 + d, 1 + d), Lx * (1 + d), Ly * (1 + d), eps_ab=0.5).U()
    dn = _sim(_scaled(f, 1 - d, 1 - d), Lx * (1 - d), Ly * (1 - d), eps_ab=0.5).U()
    assert (up - dn) / (2 * d) == pytest.approx(-W, rel=1e-6)


def test_partial_U_lambda_is_shear_derivative():
    f, Lx, Ly = lattice_frame(600, 0.8, seed=22, jitter=0.15)
    eps = 0.3
    s = _sim(f, Lx, Ly, eps_ab=0.5)
    dudl = s.partial_U_lambda(eps)
    d = 1e-6
    a, b = math.exp(eps * d), math.exp(-eps * d)
    up = _sim(_scaled(f, a, b), Lx * a, Ly * b, eps_ab=0.5).U()
    a, b = math.exp(-eps * d), math.exp(eps * d)
    dn = _sim(_scaled(f, a, b), Lx * a, Ly * b, eps_ab=0.5).U()
    assert (up - dn) / (2 * d) == pytest.approx(dudl, rel=1e-6)


def test_ideal_gas_pressure_and_kinetic_energy():
    f, Lx, Ly = lattice_frame(500, 0.5, seed=4, T=1.3)
    s = _sim(f, Lx, Ly, eps_aa=0.0, eps_bb=0.0, eps_ab=0.0, mass_a=2.0, mass_b=0.5)
    m = np.where(f["type"] == 0, 2.0, 0.5)
    v2 = (f["vel"] ** 2).sum(1)
    assert s.K() == pytest.approx(0.5 * (m * v2).sum(), rel=1e-13)
    assert s.pressure() == pytest.approx((m * v2).sum() / (2 * Lx * Ly), rel=1e-13)


def _energy(s):
    return s.K() + s.U()


def test_nve_conserves_energy_and_is_reversible():
    f, Lx, Ly = lattice_frame(400, 0.75, seed=7, T=0.7)
    s = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002)
    s.compute_forces()
    E0 = _energy(s)
    start = s.frame()
    s.step_nve(300)
    E1 = _energy(s)
    assert abs(E1 - E0) < 2e-4 * abs(E0)
    g = s.frame()
    g["vel"] *= -1.0
    s.set_frame(g)
    s.compute_forces()
    s.step_nve(300)
    back = s.frame()
    d = back["pos"] - start["pos"]
    d[:, 0] -= Lx * np.round(d[:, 0] / Lx)
    d[:, 1] -= Ly * np.round(d[:, 1] / Ly)
    assert np.abs(d).max() < 1e-8
    assert np.abs(back["vel"] + start["vel"]).max() < 1e-8


def test_nvt_large_Q_reduces_to_nve_and_thermostat_drives_T():
    f, Lx, Ly = lattice_frame(300, 0.75, seed=8, T=0.7)
    a = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002, Q=1e300, T_target=2.0)
    b = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002)
    a.compute_forces()
    b.compute_forces()
    a.step_nvt(100)
    b.step_nve(100)
    assert np.abs(a.frame()["pos"] - b.frame()["pos"]).max() < 1e-12
    # a finite Q pulls the kinetic temperature towards T_target
    c = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002, Q=2.0, T_target=2.0)
    c.compute_forces()
    c.step_nvt(2000)
    T = c.K() / (len(f) - 1)
    assert 1.2 < T < 3.0
    assert c.obs().xi != 0.0


def test_nph_large_W_reduces_to_nve_and_barostat_moves_box():
    f, Lx, Ly = lattice_frame(300, 0.75, seed=10, T=0.7)
    a = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002, barostat_mass=1e300)
    b = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002)
    a.compute_forces()
    b.compute_forces()
    a.step_nph(50, 1.0)
    b.step_nve(50)
    assert a.obs().Lx == Lx
    assert np.abs(a.frame()["pos"] - b.frame()["pos"]).max() < 1e-12
    c = _sim(f, Lx, Ly, eps_ab=0.5, dt=0.002, barostat_mass=5.0)
    c.compute_forces()
    c.step_nph(400, 5.0)   # high external pressure compresses the box
    o = c.obs()
    assert o.Lx < Lx and o.Ly < Ly
    assert o.Lx / o.Ly == pytest.approx(Lx / Ly, rel=1e-12)   # fixed aspect ratio


def test_nph_clamps_at_height_floor():
    f, Lx, Ly = lattice_frame(100, 0.3, seed=12, T=0.1)
    c = _sim(f, Lx, Ly, dt=0.005, barostat_mass=0.01, barostat_height_min_ratio=0.95)
    c.compute_forces()
    c.step_nph(300, 50.0)
    o = c.obs()
    assert o.Ly == pytest.approx(0.95 * Ly, rel=1e-12)
    assert o.p_A >= 0.0


def test_piston_wall_bounce_closed_form():
    # one particle moving up into the wall; box held fixed by a huge piston mass
    k, m, v0 = 100.0, 1.0, 1.0
    f = np.zeros(1, O.PARTICLE_DTYPE)
    f["pos"][0] = (5.0, 9.9)
    f["vel"][0] = (0.0, v0)
    s = _sim(f, 10.0, 10.0, k_piston=k, barostat_mass=1e300, dt=1e-4)
    s.step_piston(1, 0.0)
    # integrate until it comes back below the wall
    n = 0
    max_depth = 0.0
    while True:
        s.step_piston(1, 0.0)
        n += 1
        y = s.frame()["pos"][0, 1]
        max_depth = max(max_depth, y - 10.0)
        if y < 9.9 or n > 20000:
            break
    vy = s.frame()["vel"][0, 1]
    assert vy == pytest.approx(-v0, rel=1e-4)
    assert max_depth == pytest.approx(v0 / math.sqrt(k / m), rel=1e-3)


def test_fire_two_particle_reference_case():
    # reference: two type-0 particles at (1,1), (3,1) in a 10x10 box, dt = 1e-3,
    # FireParams::from_config (dt_max = 5 dt), force tolerance 1e-3.
    f = np.zeros(2, O.PARTICLE_DTYPE)
    f["pos"][0] = (1.0, 1.0)
    f["pos"][1] = (3.0, 1.0)
    s = _sim(f, 10.0, 10.0, eps_ab=1.0, dt=1e-3)
    res, es, et = s.fire(dt_init=1e-3, dt_max=5e-3, force_tol=1e-3, max_steps=20000)
    assert res.converged == 1
    g = s.frame()
    r = abs(g["pos"][1, 0] - g["pos"][0, 0])
    assert r == pytest.approx(RSTAR, abs=1e-4)
    assert res.max_force < 1e-3
    assert res.potential_energy == pytest.approx(-1.0 - VC25, abs=1e-6)
    assert es[0] == 0 and es[-1] == res.steps + 1
    assert np.all(np.diff(et) <= 1e-12)


def test_fire_immediate_convergence_and_max_steps_semantics():
    f = np.zeros(2, O.PARTICLE_DTYPE)
    f["pos"][0] = (1.0, 1.0)
    f["pos"][1] = (3.0, 1.0)
    s = _sim(f, 10.0, 10.0, eps_ab=1.0, dt=1e-3)
    res, es, et = s.fire(dt_init=1e-3, dt_max=5e-3, force_tol=1e9, max_steps=10)
    assert res.converged == 1 and res.steps == 0
    assert list(es) == [0, 1]
    s2 = _sim(f, 10.0, 10.0, eps_ab=1.0, dt=1e-3)
    res, es, et = s2.fire(dt_init=1e-3, dt_max=5e-3, force_tol=1e-30, max_steps=7)
    assert res.converged == 0 and res.steps == 7
    assert list(es) == [0, 7]
