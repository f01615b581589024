"""Pins of the oracle's integrators and FIRE schedule that do not rest on the
oracle's own formulas.

- FIRE: an independent numpy FIRE written from Bitzek et al., PRL 97, 170201
  (2006) -- MD kick, power P = F.v, velocity mixing v <- (1 - alpha) v +
  alpha |v| F/|F|, and the adaptive rules (P > 0 for more than N_min steps:
  dt <- min(f_inc dt, dt_max), alpha <- 0.99 alpha; P <= 0: dt <- f_dec dt,
  alpha <- alpha0, v <- 0) -- taken in the reference's step order (kick with
  the current forces, then P and the schedule, then the mixing, then the
  drift; src/md_energy_minimizer.cu:170-215).  On two particles the
  reference's per-particle mixing equals the paper's global mixing (both
  particles have the same |v| and |F|), so the trajectories must agree step
  for step.  The run is checked to visit every branch of the schedule: a
  streak beyond N_min, dt capped at dt_max, alpha decay, and resets.  Two
  particles keep v parallel to F, where the mixing (and alpha) has no effect;
  an irregular four-particle cluster pins the reference's per-particle
  mixing and alpha schedule.
- Nose-Hoover NVT (src/md_env.cu:1639-1728): the conserved energy
  H = K + U + Q xi^2 / 2 + g T0 eta (eta = integral of xi dt, g = 2N - 2)
  drifts by an amount that halves with dt.  The reference applies the
  friction of each half kick explicitly with the freshly updated xi, which
  makes the scheme first order in H; a dropped 1/2, a wrong g or a swapped K
  would leave an O(1) drift that does not scale with dt.
- NPH area barostat (src/md_env.cu:1795-1918): H = K + U + p_A^2 / (2 W) +
  P_ext A is conserved to O(dt^2): its drift falls 4x per halving of dt.
"""
import math

import numpy as np

from oracle import oracle as O
from tests.helpers import lattice_frame


# ---------------------------------------------------------------------------
# FIRE
# ---------------------------------------------------------------------------

def lj_force_pair(d, eps=1.0, sig=1.0, rc=2.5):
    """Force on the particle at +d from the one at 0: F = 24 eps (2 s^12/r^13
    - s^6/r^7) d/r inside the cutoff (closed form of the LJ potential)."""
    r = math.hypot(d[0], d[1])
    if r >= rc * sig or r == 0.0:
        return np.zeros(2)
    f = 24.0 * eps * (2.0 * sig ** 12 / r ** 13 - sig ** 6 / r ** 7)
    return f * np.asarray(d) / r


def lj_forces(x, L):
    """Forces on every particle (all pairs, minimum image, one type)."""
    F = np.zeros_like(x)
    for i in range(len(x)):
        for j in range(len(x)):
            if i != j:
                d = x[i] - x[j]
                d -= L * np.round(d / L)
                F[i] += lj_force_pair(d)
    return F


def numpy_fire(pos, L, dt0, dt_max, alpha0, f_inc, f_dec, n_min, steps, per_particle=False):
    """Bitzek et al.'s FIRE (mass 1, periodic square box L) in the reference's
    step order; per_particle mixes each particle's velocity with its own force
    (the reference's variant, src/md_energy_minimizer.cu:40-60) instead of the
    paper's global vectors.  Returns per-step positions, velocities, dt, alpha
    and the branch taken."""
    x = np.array(pos, float)
    v = np.zeros_like(x)
    dt, alpha, streak = dt0, alpha0, 0
    out = []
    for _ in range(steps):
        F = lj_forces(x, L)
        v = v + dt * F                       # MD kick with the current forces
        P = float((F * v).sum())             # power
        if P > 0.0:
            streak += 1
            branch = "grow" if streak > n_min else "hold"
            if streak > n_min:
                dt = min(dt * f_inc, dt_max)
                alpha = alpha * 0.99
        else:
            branch = "reset"
            dt *= f_dec
            alpha = alpha0
            streak = 0
            v = np.zeros_like(v)
        if per_particle:
            vn = np.linalg.norm(v, axis=1, keepdims=True)
            fn = np.linalg.norm(F, axis=1, keepdims=True)
            ok = (vn > 0.0) & (fn > 0.0)
            v = np.where(ok, (1.0 - alpha) * v + alpha * vn * F / np.where(ok, fn, 1.0), v)
        else:                                # global mixing
            vn = np.linalg.norm(v)
            fn = np.linalg.norm(F)
            if vn > 0.0 and fn > 0.0:
                v = (1.0 - alpha) * v + alpha * vn * F / fn
        x = np.mod(x + dt * v, L)            # drift
        out.append((x.copy(), v.copy(), dt, alpha, branch))
    return out


def oracle_fire_frames(pos, L, dt0, dt_max, alpha0, f_inc, f_dec, n_min, steps):
    """The oracle's FIRE stopped after k = 1..steps iterations (a fresh
    oracle each time; force_tol 0 so it never converges early)."""
    frames = []
    for k in range(1, steps + 1):
        f = np.zeros(len(pos), O.PARTICLE_DTYPE)
        f["pos"] = pos
        o = O.OracleSim(O.make_params(L, L, force_mode=0, eps_ab=1.0), f)
        o.fire(dt0, dt_max, alpha0, f_inc, f_dec, n_min, max_steps=k, force_tol=0.0)
        frames.append(o.frame())
        o.close()
    return frames


def test_fire_schedule_against_numpy_fire():
    L = 20.0
    pos = [[9.55, 10.0], [10.45, 10.0]]   # 0.9 sigma apart: repulsive start
    args = (0.005, 0.012, 0.1, 1.1, 0.5, 5)
    steps = 120
    ref = numpy_fire(pos, L, *args, steps)
    got = oracle_fire_frames(pos, L, *args, steps)
    for k, ((x, v, dt, alpha, _), fr) in enumerate(zip(ref, got)):
        assert np.allclose(fr["pos"], x, rtol=0, atol=1e-12), k
        assert np.allclose(fr["vel"], v, rtol=0, atol=1e-12), k
    branches = [b for *_, b in ref]
    dts = [dt for _, _, dt, _, _ in ref]
    alphas = [a for _, _, _, a, _ in ref]
    assert "grow" in branches and "reset" in branches and "hold" in branches
    assert max(dts) == 0.012 and dts.count(0.012) > 5   # dt held at the cap
    assert min(alphas) < 0.1 * 0.99 ** 5         # alpha decayed over a streak
    assert any(b == "reset" for b in branches[10:])


def test_fire_per_particle_mixing_against_numpy():
    """Four particles of an irregular cluster: velocities turn away from
    their forces, so the mixing (and alpha) shapes the path; the reference's
    per-particle mixing is followed step for step."""
    L = 30.0
    pos = [[14.0, 15.0], [15.05, 15.2], [14.6, 16.0], [15.9, 16.3]]
    args = (0.004, 0.012, 0.1, 1.1, 0.5, 5)
    steps = 90
    ref = numpy_fire(pos, L, *args, steps, per_particle=True)
    got = oracle_fire_frames(pos, L, *args, steps)
    for k, ((x, v, _, _, _), fr) in enumerate(zip(ref, got)):
        assert np.allclose(fr["pos"], x, rtol=0, atol=1e-11), k
        assert np.allclose(fr["vel"], v, rtol=0, atol=1e-11), k
    # the mixing matters here: the paper's global mixing gives another path
    glob = numpy_fire(pos, L, *args, steps, per_particle=False)
    assert np.abs(glob[-1][0] - ref[-1][0]).max() > 1e-6
    branches = [b for *_, b in ref]
    assert "grow" in branches and "reset" in branches


def test_fire_converges_to_lj_minimum_like_numpy():
    """Both end at r = 2^(1/6) sigma when run long enough."""
    L = 20.0
    pos = [[9.55, 10.0], [10.45, 10.0]]
    ref = numpy_fire(pos, L, 0.005, 0.05, 0.1, 1.1, 0.5, 5, 3000)
    x = ref[-1][0]
    assert abs(abs(x[1, 0] - x[0, 0]) - 2.0 ** (1.0 / 6.0)) < 1e-6
    f = np.zeros(2, O.PARTICLE_DTYPE)
    f["pos"] = pos
    o = O.OracleSim(O.make_params(L, L, force_mode=0, eps_ab=1.0), f)
    res, _, _ = o.fire(0.005, 0.05, 0.1, 1.1, 0.5, 5, max_steps=3000, force_tol=1e-9)
    fr = o.frame()
    assert res.converged == 1
    assert abs(abs(fr["pos"][1, 0] - fr["pos"][0, 0]) - 2.0 ** (1.0 / 6.0)) < 1e-8


# ---------------------------------------------------------------------------
# conserved energies of NVT and NPH
# ---------------------------------------------------------------------------

def nvt_H_drift(dt, f, Lx, Ly, Q, T0, t_total):
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, eps_ab=0.5, dt=dt, T_target=T0, Q=Q), f)
    o.compute_forces()
    g = 2 * len(f) - 2

    def H():
        ob = o.obs()
        return o.K() + o.U() + 0.5 * Q * ob.xi ** 2 + g * T0 * ob.eta

    H0, d = H(), 0.0
    n = int(round(t_total / dt))
    for _ in range(10):
        o.step_nvt(n // 10)
        d = max(d, abs(H() - H0))
    o.close()
    return d / len(f)


def test_nose_hoover_conserved_energy_drift_is_first_order():
    f, Lx, Ly = lattice_frame(256, 0.8, seed=5, jitter=0.05, T=0.7)
    for Q in (10.0, 1.0):
        d = [nvt_H_drift(dt, f, Lx, Ly, Q, 1.0, 1.0) for dt in (0.004, 0.002, 0.001)]
        r1, r2 = d[0] / d[1], d[1] / d[2]
        assert 1.7 < r1 < 2.3 and 1.7 < r2 < 2.3, (Q, d)
        assert d[2] < 5e-3, (Q, d)


def nph_H_drift(dt, f, Lx, Ly, W, p_ext, t_total):
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, eps_ab=0.5, dt=dt, barostat_mass=W), f)
    o.compute_forces()

    def H():
        ob = o.obs()
        return o.K() + o.U() + ob.p_A ** 2 / (2.0 * W) + p_ext * ob.Lx * ob.Ly

    H0, d = H(), 0.0
    n = int(round(t_total / dt))
    for _ in range(10):
        o.step_nph(n // 10, p_ext)
        d = max(d, abs(H() - H0))
    o.close()
    return d / len(f)


def test_nph_conserved_energy_drift_is_second_order():
    f, Lx, Ly = lattice_frame(256, 0.8, seed=5, jitter=0.05, T=0.7)
    d = [nph_H_drift(dt, f, Lx, Ly, 50.0, 2.0, 1.0) for dt in (0.004, 0.002, 0.001)]
    r1, r2 = d[0] / d[1], d[1] / d[2]
    assert 3.5 < r1 < 4.6 and 3.5 < r2 < 4.6, d
    assert d[2] < 1e-4, d


# ---------------------------------------------------------------------------
# middle_reflect_LG
# ---------------------------------------------------------------------------

def test_middle_reflect_hand_cases():
    """src/md_env.cu:1457-1508: x is mirrored at the edges of the central slab
    of width devide_p Lx + 2 sigma_max (at most Lx) until it lies inside, y is
    wrapped.  Lx = 20, Ly = 10, devide_p 0.3, sigma 1: slab [6, 14]."""
    x = [10.0, 6.0, 14.0, 15.0, 19.0, 5.5, 0.5, 13.75]
    y = [1.0, -0.5, 10.25, 3.0, 9.0, 4.0, 5.0, 2.0]
    want_x = [10.0, 6.0, 14.0, 13.0, 9.0, 6.5, 11.5, 13.75]   # 15 -> 14 - 1; 19 -> 14 - 5; 0.5 -> 6 + 5.5
    want_y = [1.0, 9.5, 0.25, 3.0, 9.0, 4.0, 5.0, 2.0]
    f = np.zeros(len(x), O.PARTICLE_DTYPE)
    f["pos"] = np.stack([x, y], 1)

    def fold(devide_p, Lx=20.0):
        o = O.OracleSim(O.make_params(Lx, 10.0, force_mode=0, devide_p=devide_p), f)
        o.middle_reflect()
        fr = o.frame()
        o.close()
        return fr["pos"]

    p = fold(0.3)
    assert p[:, 0].tolist() == want_x and p[:, 1].tolist() == want_y
    # a slab wider than the box is the box: x in [0, Lx) stays, y still wraps
    p = fold(0.95)
    assert p[:, 0].tolist() == x and p[:, 1].tolist() == want_y
    # devide_p <= 0 switches the fold off entirely
    p = fold(0.0)
    assert p[:, 0].tolist() == x and p[:, 1].tolist() == y


def test_middle_reflect_is_a_triangle_wave():
    """On random x over several periods the fold equals the triangle wave
    x_min + w - | ((x - x_min) mod 2w) - w | of period 2w."""
    rng = np.random.default_rng(7)
    Lx, Ly, dp = 50.0, 10.0, 0.4
    w = dp * Lx + 2.0
    x_min = 0.5 * Lx - 0.5 * w
    x = rng.uniform(-3 * Lx, 3 * Lx, 5000)
    f = np.zeros(len(x), O.PARTICLE_DTYPE)
    f["pos"][:, 0] = x
    f["pos"][:, 1] = rng.uniform(0, Ly, len(x))
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=0, devide_p=dp), f)
    o.middle_reflect()
    got = o.frame()["pos"][:, 0]
    o.close()
    tri = x_min + w - np.abs(np.mod(x - x_min, 2 * w) - w)
    assert np.abs(got - tri).max() < 1e-12
    assert got.min() >= x_min and got.max() <= x_min + w


# ---------------------------------------------------------------------------
# pressure-tensor profile
# ---------------------------------------------------------------------------

def test_pressure_profile_sums_to_the_scalar_pressure_and_lambda():
    """Summed over the slabs (times each slab's area), Pxx + Pyy is 2 A P
    with P the scalar pressure (src/md_env.cu:2205-2252), whatever the number
    of slabs, and at rest Pxx - Pyy sums to -dU/dlambda at eps_lambda = 1
    (src/md_env.cu:2362-2406) -- both pinned by finite differences."""
    f, Lx, Ly = lattice_frame(256, 0.8, seed=5, jitter=0.08, T=0.7)
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, eps_ab=0.5), f)
    P = o.pressure()
    area = Lx * Ly
    for nb in (1, 7, 50):
        pp = o.pressure_profile(nb).reshape(3, nb)
        V = area / nb
        P_sum = (pp[0] + pp[1]).sum() * V / (2.0 * area)
        assert abs(P_sum - P) <= 1e-9 * abs(P), (nb, P_sum, P)
    o.close()
    g = f.copy()
    g["vel"] = 0.0
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, eps_ab=0.5), g)
    lam = o.partial_U_lambda(1.0)
    pp = o.pressure_profile(9).reshape(3, 9)
    assert abs((pp[0] - pp[1]).sum() * area / 9 + lam) <= 1e-9 * abs(lam)
    # Pxy = Pyx for central forces: sum over slabs of 1/2 dx Fy equals 1/2 dy Fx,
    # checked against a brute-force sum
    pos = np.mod(g["pos"], [Lx, Ly])
    sxy = 0.0
    for i in range(len(pos)):
        d = pos[i] - pos
        d -= np.round(d / [Lx, Ly]) * [Lx, Ly]
        for j in range(len(pos)):
            if j == i:
                continue
            eps = 1.0 if g["type"][i] == g["type"][j] else 0.5
            F = lj_force_pair(d[j], eps=eps)
            sxy += 0.5 * d[j][0] * F[1]
    assert abs(pp[2].sum() * area / 9 - sxy) <= 1e-8 * max(1.0, abs(sxy))
    o.close()


def test_pressure_profile_hand_cases():
    """Two particles, Ly = 10, 10 slabs: the pair's virial lands in the slab
    of its midpoint (twice, once per ordered pair), each kinetic term in its
    particle's slab; across the periodic edge the midpoint wraps."""
    L, nb = 10.0, 10
    for (y1, y2, want_bin) in ((1.8, 2.7, 2), (0.2, 9.7, 9)):
        f = np.zeros(2, O.PARTICLE_DTYPE)
        f["pos"] = [[5.0, y1], [5.4, y2]]
        f["vel"] = [[0.5, -0.25], [-1.0, 0.75]]
        o = O.OracleSim(O.make_params(L, L, force_mode=0, eps_ab=1.0), f)
        pp = o.pressure_profile(nb).reshape(3, nb) * (L * L / nb)
        o.close()
        d = np.array([5.0 - 5.4, y1 - y2])
        d -= np.round(d / L) * L
        F = lj_force_pair(d)
        want = np.zeros((3, nb))
        for (y, v) in ((y1, f["vel"][0]), (y2, f["vel"][1])):
            b = int(y / L * nb)
            want[:, b] += [v[0] * v[0], v[1] * v[1], v[0] * v[1]]
        want[:, want_bin] += [d[0] * F[0], d[1] * F[1], d[0] * F[1]]   # two ordered pairs, 1/2 each
        assert np.allclose(pp, want, rtol=1e-9, atol=1e-9), (y1, y2, pp, want)
