"""NPH (area barostat) in neighbour-list mode: between rebuilds the barostat
scales every position by the same factor, so the list stays exact while each
particle's displacement from its scaled build position is at most
(s r_list - r_c) / 2; a change of the grid's dimensions rebuilds.  The pair
set is checked against a fresh cell pass at the same positions after the box
has shrunk through several grid changes, the box and barostat against the
FP64 oracle, and batched slots against the same systems alone."""
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


def test_nph_list_pair_set_is_exact_under_a_shrinking_box():
    f, Lx, Ly = lattice_frame(8000, 0.55, seed=41, jitter=0.05, T=0.8, aspect=2.0)
    # a light barostat, so the box shrinks by several per cent in 400 steps
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.8, Q=10.0, barostat_mass=0.005)
    say("NPH list mode, P_ext 4 from density 0.55, W 0.005")
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.set_tuning(1.0, 1, 0.5)
    ncx0 = int(Lx // 3.0)
    s.step(abi.NPH, 400, p_ext=4.0)
    ob = s.observables(0)
    got = s.collect()
    s.close()
    say(f"  Lx {Lx:.3f} -> {ob.Lx:.3f} (cells per row {ncx0} -> {int(ob.Lx // 3.0)})")
    assert ob.Lx < 0.97 * Lx                  # the grid changed dimensions on the way
    assert int(ob.Lx // 3.0) < ncx0
    c = abi.Simulation(abi.make_params(ob.Lx, ob.Ly, **kw), got)
    ref = c.compute_forces(True)
    c.close()
    acc = got["acc"]
    rms = np.sqrt((ref ** 2).sum(1).mean())
    err = np.sqrt(((acc - ref) ** 2).sum(1))
    band = near_cutoff_mask(got, ob.Lx, ob.Ly, 2.5, (1.0, 1.0, 1.0))
    say(f"  max rel {err[~band].max() / rms:.2e} rms rel {np.sqrt((err ** 2).mean()) / rms:.2e}")
    # two FP32 evaluations of the same pair set (different cell sides and
    # summation order); a missing pair would be O(1) here
    assert err[~band].max() / rms < 1e-4
    assert np.sqrt((err ** 2).mean()) / rms < 5e-6


def test_nph_list_follows_oracle():
    f, Lx, Ly = lattice_frame(4096, 0.75, seed=5, jitter=0.05, T=0.7)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.7, Q=10.0, barostat_mass=50.0)
    say("NPH list mode vs FP64 oracle, N=4096")
    g = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    g.set_tuning(1.0, 1, 0.5)
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, **kw), f)
    o.compute_forces()
    for k in range(5):
        g.step(abi.NPH, 10, p_ext=1.0)
        o.step_nph(10, 1.0)
        ob, oo = g.observables(0), o.obs()
        say(f"  step {10 * (k + 1)}: Lx {ob.Lx:.10f} vs {oo.Lx:.10f}, p_A {ob.p_A:.6e} vs {oo.p_A:.6e}")
        assert abs(ob.Lx - oo.Lx) <= 1e-9 * oo.Lx and abs(ob.Ly - oo.Ly) <= 1e-9 * oo.Ly
        assert abs(ob.p_A - oo.p_A) <= 1e-5 * max(abs(oo.p_A), 1.0)
    a, b = g.collect(), o.frame()
    g.close()
    d = a["pos"] - b["pos"]
    d[:, 0] -= oo.Lx * np.round(d[:, 0] / oo.Lx)
    d[:, 1] -= oo.Ly * np.round(d[:, 1] / oo.Ly)
    say(f"  max |dx| {np.abs(d).max():.2e}")
    assert np.abs(d).max() < 1e-6


def test_nph_list_slot_in_a_batch_equals_alone():
    f1, L1x, L1y = lattice_frame(6000, 0.7, seed=8, jitter=0.05, T=0.8)
    f2, L2x, L2y = lattice_frame(9000, 0.75, seed=9, jitter=0.05, T=0.7, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.7, Q=10.0, barostat_mass=10.0)
    p1 = abi.make_params(L1x, L1y, **kw)
    p2 = abi.make_params(L2x, L2y, **kw)
    say("NPH list slot beside an NVT list slot, against each alone")
    b = abi.Batch([len(f1), len(f2)], [abi.cell_capacity(p1, 1.5), abi.cell_capacity(p2, 1.5)])
    b.load(0, p1, f1)
    b.load(1, p2, f2)
    b.set_tuning(1.0, 1, 0.5, 0)
    b.set_kind(0, abi.NPH, p_ext=2.0)
    b.set_kind(1, abi.NVT)
    b.step(150)
    got = [b.collect(0), b.collect(1)]
    b.close()
    for (f, p, kind, pe), g in zip([(f1, p1, abi.NPH, 2.0), (f2, p2, abi.NVT, 0.0)], got):
        s = abi.Simulation(p, f)
        s.set_tuning(1.0, 1, 0.5, 0)
        s.step(kind, 150, p_ext=pe)
        a = s.collect()
        s.close()
        for k in ("pos", "vel", "acc"):
            assert np.array_equal(a[k], g[k]), k



def _run(f, p, kind, steps, skin, cell_scale=1.0, energy_first=False, pe=0.0):
    s = abi.Simulation(p, f)
    s.set_tuning(cell_scale, 1, skin, 0)
    if energy_first:
        s.observables()
    s.step(kind, steps, p_ext=pe)
    a = s.collect()
    s.close()
    return a


@pytest.mark.parametrize("kind,first", [(abi.NVT, True), (abi.NPH, False)])
def test_list_mode_survives_an_energy_pass_before_the_first_step(kind, first):
    """An energy pass before a slot's first step (observables() on a fresh
    handle; an NPH run's initial virial pass) runs in list mode with the
    slot's list parameters set, so list mode stays on for the run.  Compared
    with the cell list on the same grid (cells of r_c + skin = 3.0: cell_scale
    1.2 and no skin), whose pair sums group the terms by cell, the list run
    differs in rounding only; a slot that had fallen back to the cell list
    would equal it bit for bit."""
    f, Lx, Ly = lattice_frame(4096, 0.75, seed=12, jitter=0.05, T=0.7)
    p = abi.make_params(Lx, Ly, eps_ab=0.5, dt=0.005, T_target=0.7, Q=10.0, barostat_mass=50.0)
    pe = 1.0 if kind == abi.NPH else 0.0
    lists = _run(f, p, kind, 30, 0.5, energy_first=first, pe=pe)
    cells = _run(f, p, kind, 30, 0.0, cell_scale=1.2, energy_first=first, pe=pe)
    d = np.abs(lists["pos"] - cells["pos"]).max()
    say(f"  lists vs cell list on the same grid: max |dx| {d:.2e}")
    assert not np.array_equal(lists["acc"], cells["acc"])
    assert d < 1e-6
