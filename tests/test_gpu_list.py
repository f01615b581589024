"""Neighbour-list mode (skin > 0): the pairs evaluated are those within the
cutoff, as in the per-step cell pass, so accelerations after many list steps
(and list rebuilds) equal a fresh cell pass at the same positions to FP32
rounding; NVE energy over 1e4 steps stays within 2x of the FP64 oracle's
drift; results are the same alone and in a batch; switching between list
steps and other passes keeps the cell list consistent."""
import sys

import numpy as np
import pytest

from md_simulation_cuda_amd import abi
from oracle import oracle as O
from tests.helpers import lattice_frame, near_cutoff_mask

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("skin,lanes", [(0.3, 1), (0.5, 2), (0.3, 4)])
def test_list_accelerations_equal_fresh_cell_pass(skin, lanes):
    f, Lx, Ly = lattice_frame(20000, 0.75, seed=31, jitter=0.05, T=0.8, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.8, Q=10.0)
    say(f"list vs cell: skin={skin} lanes={lanes}")
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.set_tuning(1.0, lanes, skin)
    assert s.get_tuning().skin == skin
    s.step(abi.NVT, 300)   # far beyond one list lifetime at T=0.8
    got = s.collect()
    s.close()
    c = abi.Simulation(abi.make_params(Lx, Ly, **kw), got)
    ref = c.compute_forces(True)
    c.close()
    acc = got["acc"]
    rms = np.sqrt((ref ** 2).sum(1).mean())
    err = np.sqrt(((acc - ref) ** 2).sum(1))
    band = near_cutoff_mask(got, Lx, Ly, 2.5, (1.0, 1.0, 1.0))
    say(f"  max rel {err[~band].max() / rms:.2e} rms rel {np.sqrt((err ** 2).mean()) / rms:.2e}")
    # two FP32 evaluations of the same pair set (different cell sides and
    # summation order); a missing or extra pair would be O(1) here
    assert err[~band].max() / rms < 1e-4
    assert np.sqrt((err ** 2).mean()) / rms < 5e-6


def test_list_nve_drift_against_oracle():
    n = 4096
    f, Lx, Ly = abi.gen_lattice(n, n // 2, 0.8, 0.05, 0.7, 1001)
    kw = dict(eps_ab=0.5, dt=0.005)
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, **kw), f)
    o.compute_forces()
    Eo = [o.K() + o.U()]
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.set_tuning(1.0, 1, 0.4)
    ob = s.observables()
    E = [ob.K + ob.U]
    for _ in range(20):
        o.step_nve(500)
        Eo.append(o.K() + o.U())
        s.step(abi.NVE, 500)
        ob = s.observables()
        E.append(ob.K + ob.U)
    s.close()
    d_o = np.max(np.abs(np.array(Eo) - Eo[0])) / n
    d = np.max(np.abs(np.array(E) - E[0])) / n
    say(f"list NVE drift {d:.3e} oracle {d_o:.3e}")
    assert d <= 2.0 * d_o


def test_list_same_alone_and_in_batch():
    f, Lx, Ly = lattice_frame(3000, 0.75, seed=5, jitter=0.05, T=0.6, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.6, Q=10.0)
    a = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    a.set_tuning(1.0, 1, 0.3)
    a.step(abi.NVT, 120)
    fa = a.collect()
    a.close()
    f2, L2x, L2y = lattice_frame(1500, 0.8, seed=6, jitter=0.05, T=0.9)
    caps = [abi.cell_capacity(abi.make_params(Lx, Ly), 1.0), abi.cell_capacity(abi.make_params(L2x, L2y), 1.0)]
    b = abi.Batch([3000, 1500], caps)
    b.load(0, abi.make_params(Lx, Ly, **kw), f)
    b.load(1, abi.make_params(L2x, L2y, eps_ab=0.4, dt=0.005), f2)
    b.set_tuning(1.0, 1, 0.3)
    b.set_kind(0, abi.NVT)
    b.set_kind(1, abi.NVE)
    b.step(120)
    fb = b.collect(0)
    b.close()
    assert np.array_equal(fa["pos"], fb["pos"]) and np.array_equal(fa["vel"], fb["vel"])


def test_list_mode_transitions_keep_cell_list_consistent():
    """List steps interleaved with energy passes (per-step cell pass) and NPH
    steps (no list), then the canonical cell list and forces are checked."""
    f, Lx, Ly = lattice_frame(5000, 0.75, seed=9, jitter=0.05, T=0.8)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.8, Q=10.0, barostat_mass=50.0)
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.set_tuning(1.0, 2, 0.3)
    for k in range(4):
        s.step(abi.NVT, 37)
        ob = s.observables()
        assert np.isfinite(ob.U) and np.isfinite(ob.K)
        s.step(abi.NVE, 23)
        if k == 2:
            s.step(abi.NPH, 10, p_ext=1.0)
    g = s.collect()
    ncx, ncy, keys, cs, ids = s.cell_list(True)
    assert len(np.unique(ids)) == len(g)
    assert np.all(np.diff(cs.astype(np.int64)) >= 0) and cs[-1] == len(g)
    for c in range(ncx * ncy):
        seg = ids[cs[c]:cs[c + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        assert np.all(keys[seg] == c)
    acc = g["acc"]
    s2 = abi.Simulation(abi.make_params(s.observables().Lx, s.observables().Ly, **kw), g)
    s.close()
    ref = s2.compute_forces(True)
    s2.close()
    rms = np.sqrt((ref ** 2).sum(1).mean())
    assert np.sqrt((((acc - ref) ** 2).sum(1)).mean()) / rms < 5e-6


def _acc_error(got, Lx, Ly, kw):
    c = abi.Simulation(abi.make_params(Lx, Ly, **kw), got)
    ref = c.compute_forces(True)
    c.close()
    rms = np.sqrt((ref ** 2).sum(1).mean())
    err = np.sqrt(((got["acc"] - ref) ** 2).sum(1))
    band = near_cutoff_mask(got, Lx, Ly, 2.5, (1.0, 1.0, 1.0))
    return err[~band].max() / rms, np.sqrt((err ** 2).mean()) / rms


@pytest.mark.parametrize("T,skin,window", [(0.8, 0.5, 8), (2.0, 0.3, 16), (2.0, 0.3, 32)])
def test_windows_keep_the_exact_pair_set(T, skin, window):
    """Rebuild windows; at T=2 with skin 0.3 particles leave the skin inside
    most 16- and 32-step windows, so the lean steps' 5x5 fallback runs."""
    f, Lx, Ly = lattice_frame(12000, 0.75, seed=41, jitter=0.05, T=T, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=T, Q=10.0)
    say(f"windows T={T} skin={skin} window={window}")
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.set_tuning(1.0, 1, skin, window)
    assert s.get_tuning().window == window
    s.step(abi.NVT, 160)
    got = s.collect()
    s.close()
    mx, rm = _acc_error(got, Lx, Ly, kw)
    say(f"  max rel {mx:.2e} rms rel {rm:.2e}")
    assert mx < 1e-4 and rm < 5e-6


@pytest.mark.parametrize("kind", [abi.NVT, abi.NVE])
def test_window_graphs_equal_eager_steps(kind):
    """A whole rebuild window replays as one graph; calls of uneven length
    mix such replays with step-by-step shorter windows, with host reads and
    synchronisations between them; every result equals eager launches."""
    f, Lx, Ly = lattice_frame(12000, 0.75, seed=43, jitter=0.05, T=0.8, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.8, Q=10.0)
    say(f"window graphs vs eager: kind={kind}")
    res = []
    for graph in (True, False):
        s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
        s.set_graph(graph)
        s.set_tuning(1.0, 1, 0.5, 8)
        obs = []
        for count in (20, 13, 8, 8, 3, 16):
            s.step(kind, count)
            s.sync()
            ob = s.observables()
            obs.append((ob.K, ob.U, ob.xi))
        res.append((s.collect(), obs))
        s.close()
    (fa, oa), (fb, ob_) = res
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(fa[k], fb[k]), k
    assert oa == ob_


def test_windows_nve_drift_against_oracle():
    n = 4096
    f, Lx, Ly = abi.gen_lattice(n, n // 2, 0.8, 0.05, 0.7, 1001)
    kw = dict(eps_ab=0.5, dt=0.005)
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, **kw), f)
    o.compute_forces()
    Eo = [o.K() + o.U()]
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.set_tuning(1.0, 1, 0.4, 16)
    ob = s.observables()
    E = [ob.K + ob.U]
    for _ in range(20):
        o.step_nve(500)
        Eo.append(o.K() + o.U())
        s.step(abi.NVE, 500)
        ob = s.observables()
        E.append(ob.K + ob.U)
    s.close()
    d_o = np.max(np.abs(np.array(Eo) - Eo[0])) / n
    d = np.max(np.abs(np.array(E) - E[0])) / n
    say(f"windowed NVE drift {d:.3e} oracle {d_o:.3e}")
    assert d <= 2.0 * d_o


def test_windows_same_alone_and_in_batch_with_a_falling_back_neighbour():
    cold = lattice_frame(3000, 0.75, seed=51, jitter=0.05, T=0.4, aspect=2.0)
    hot = lattice_frame(2000, 0.75, seed=52, jitter=0.05, T=2.5)
    kws = [dict(eps_ab=0.5, dt=0.005, T_target=0.4, Q=10.0), dict(eps_ab=0.7, dt=0.005, T_target=2.5, Q=10.0)]
    alone = []
    for (f, Lx, Ly), kw in zip((cold, hot), kws):
        s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
        s.set_tuning(1.0, 1, 0.3, 16)
        s.step(abi.NVT, 96)
        alone.append(s.collect())
        s.close()
    caps = [abi.cell_capacity(abi.make_params(Lx, Ly), 1.0) for _, Lx, Ly in (cold, hot)]
    b = abi.Batch([3000, 2000], caps)
    for r, ((f, Lx, Ly), kw) in enumerate(zip((cold, hot), kws)):
        b.load(r, abi.make_params(Lx, Ly, **kw), f)
        b.set_kind(r, abi.NVT)
    b.set_tuning(1.0, 1, 0.3, 16)
    b.step(96)
    for r in range(2):
        got = b.collect(r)
        assert np.array_equal(got["pos"], alone[r]["pos"]) and np.array_equal(got["vel"], alone[r]["vel"]), r
    b.close()


@pytest.mark.parametrize("kind", [abi.NVT, abi.NVE])
def test_windows_run_on_across_step_calls(kind):
    """With rebuild windows, 100 calls of one step (the drop-in's
    step_single_nose_hoover loop) give the trajectory of one 100-step call
    bit for bit: each call continues the window the last one left open."""
    f, Lx, Ly = lattice_frame(5000, 0.75, seed=61, jitter=0.05, T=0.9, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.9, Q=10.0)
    say(f"windows across calls: kind={kind}")
    res = []
    for calls in ([100], [1] * 100, [7, 1, 1, 30, 61]):
        s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
        s.set_tuning(1.0, 1, 0.5, 12)
        for c in calls:
            s.step(kind, c)
        s.sync()
        ob = s.observables(abi.OBS_K)
        res.append((s.collect(), ob.K, ob.xi))
        s.close()
    for fr, K, xi in res[1:]:
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(fr[k], res[0][0][k]), k
        assert K == res[0][1] and xi == res[0][2]
