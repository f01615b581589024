"""Job server (continuous batching): jobs of every kind, submitted while
others run, give the same frames and energies as the same systems run alone
through the single-system API; admission respects the HBM budget."""
import sys

import numpy as np
import pytest

from md_simulation_cuda_amd import abi

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


FP = abi.FireParams(0.005, 0.025, 0.1, 1.1, 0.5, 5, 1000, 0.0, 0.0)


def job_specs():
    """(kind, n, rho, T, eps_ab, nsteps, p_ext, seed)"""
    return [
        (abi.NVT, 1500, 0.80, 0.7, 0.5, 120, 0.0, 21),
        (abi.NVE, 900, 0.75, 0.6, 0.3, 75, 0.0, 22),
        (abi.NPH, 2000, 0.80, 0.8, 0.6, 90, 1.0, 23),
        (abi.FIRE, 700, 0.85, 0.0, 0.5, 150, 0.0, 24),
        (abi.NVT, 3000, 0.70, 0.5, 0.9, 60, 0.0, 25),
        (abi.NVE, 600, 0.82, 1.0, 0.4, 200, 0.0, 26),
        (abi.FIRE, 1100, 0.80, 0.0, 0.7, 90, 0.0, 27),
        (abi.NPH, 1200, 0.78, 0.7, 0.5, 45, 1.5, 28),
        (abi.NVT, 800, 0.80, 0.9, 0.5, 33, 0.0, 29),
    ]


def make(spec):
    kind, n, rho, T, eab, nsteps, pext, seed = spec
    if kind == abi.FIRE:
        f, Lx, Ly = abi.gen_lattice(n, n // 2, rho, 0.2, 0.0, seed)
    else:
        f, Lx, Ly = abi.gen_lattice(n, n // 2, rho, 0.05, T, seed)
    kw = dict(eps_ab=eab, dt=0.005, T_target=max(T, 0.1), Q=10.0, barostat_mass=50.0)
    return f, Lx, Ly, kw


LIST_SKIN = 0.5   # the server's default for its NVE / NVT / NPH slots


def alone(spec, f, Lx, Ly, kw, fire_skin=0.0):
    kind, n, rho, T, eab, nsteps, pext, seed = spec
    s = abi.Simulation(abi.make_params(Lx, Ly, precision=abi.PREC_COMP32, **kw), f)
    if kind in (abi.NVE, abi.NVT, abi.NPH):
        s.set_tuning(1.0, 1, LIST_SKIN)   # as the server's NVE / NVT / NPH slots
    else:
        s.set_tuning(1.0, 1, fire_skin)   # as its FIRE slots
    if kind == abi.FIRE:
        res, _, _ = s.fire(FP, 1e-6, nsteps)
        out = s.collect(), res
    else:
        s.step(kind, nsteps, p_ext=pext)
        out = s.collect(), s.observables()
    s.close()
    return out


def run_server(specs, frames, budget=None, stagger=True, fire_skin=0.0):
    kw = dict(min_class_log2=9, max_class_log2=12, md_slots_per_class=2, fire_slots_per_class=1, quantum=32,
              list_slots_per_class=2, list_skin=LIST_SKIN, fire_skin=fire_skin)
    if budget is not None:
        kw["hbm_budget_bytes"] = budget
    srv = abi.Server(**kw)
    ids = {}
    pending = list(range(len(specs)))
    ticks = 0
    done = set()
    while len(done) < len(specs):
        # submit a few jobs per tick, so later jobs arrive while others run
        k = 3 if stagger else len(pending)
        for i in pending[:k]:
            kind, n, rho, T, eab, nsteps, pext, seed = specs[i]
            f, Lx, Ly, pkw = frames[i]
            job = abi.make_job(abi.make_params(Lx, Ly, **pkw), kind, nsteps, p_ext=pext, force_tol=1e-6,
                               fire=FP)
            ids[i] = srv.submit(job, f)
        pending = pending[k:]
        nr, nq, fin = srv.tick()
        done.update(fin)
        ticks += 1
        assert ticks < 10000
    return srv, ids


def test_server_results_equal_alone_runs():
    specs = job_specs()
    frames = [make(s) for s in specs]
    say("server mix: %d jobs" % len(specs))
    srv, ids = run_server(specs, frames)
    st = srv.stats()
    assert st.completed == len(specs)
    for i, spec in enumerate(specs):
        f, Lx, Ly, kw = frames[i]
        info = srv.info(ids[i])
        assert info.state == abi.JOB_DONE
        rec = srv.records(ids[i])
        got = srv.result(ids[i])
        ref, ro = alone(spec, f, Lx, Ly, kw)
        say(f"job {i} kind={spec[0]} n={spec[1]} steps={spec[5]}")
        assert np.array_equal(got["pos"], ref["pos"]), i
        assert np.array_equal(got["vel"], ref["vel"]), i
        if spec[0] == abi.FIRE:
            assert info.fire.steps == ro.steps and info.fire.converged == ro.converged
            assert info.fire.max_force == ro.max_force
            assert info.last.U == ro.potential_energy
        else:
            assert info.steps_done == spec[5]
            assert len(rec) >= 1 and rec["step"][-1] == spec[5]
            assert np.all(np.diff(rec["step"]) > 0)
            assert info.last.U == ro.U and info.last.K == ro.K, i
    srv.close()


def test_server_fire_slots_with_lists_equal_alone_runs():
    """FIRE slots with neighbour lists (fire_skin > 0, rebuilt when an
    iteration needs it) beside an NVT job: each FIRE job equals mdx_fire with
    the same tuning, including one long enough for the alone run's energy
    sample at iteration 1000, which uses the lists and so leaves the rebuild
    points where the server's run has them."""
    specs = [(abi.FIRE, 700, 0.85, 0.0, 0.5, 150, 0.0, 24), (abi.NVT, 1500, 0.80, 0.7, 0.5, 120, 0.0, 21),
             (abi.FIRE, 900, 0.80, 0.0, 0.7, 1300, 0.0, 31)]
    frames = [make(s) for s in specs]
    say("server with FIRE lists (skin 0.3)")
    srv, ids = run_server(specs, frames, fire_skin=0.3)
    for i, spec in enumerate(specs):
        f, Lx, Ly, kw = frames[i]
        info = srv.info(ids[i])
        got = srv.result(ids[i])
        ref, ro = alone(spec, f, Lx, Ly, kw, fire_skin=0.3)
        say(f"job {i} kind={spec[0]} n={spec[1]} steps={spec[5]}")
        assert np.array_equal(got["pos"], ref["pos"]) and np.array_equal(got["vel"], ref["vel"]), i
        if spec[0] == abi.FIRE:
            assert info.fire.steps == ro.steps and info.fire.max_force == ro.max_force
            assert info.last.U == ro.potential_energy
    srv.close()


def test_server_admission_respects_budget():
    specs = job_specs()[:6]
    frames = [make(s) for s in specs]
    # room for about two of the largest slots at a time (a neighbour-list slot
    # holds about 400 bytes per particle)
    budget = 2.5 * 4096 * 400.0
    say("server budget %.0f bytes" % budget)
    srv, ids = run_server(specs, frames, budget=budget, stagger=False)
    st = srv.stats()
    assert st.completed == len(specs)
    assert 0 < st.peak_reserved_bytes <= budget
    assert st.reserved_bytes == 0
    srv.close()


def test_server_rejects_oversized_and_bad_jobs():
    srv = abi.Server(min_class_log2=9, max_class_log2=10, md_slots_per_class=1, fire_slots_per_class=0)
    f, Lx, Ly = abi.gen_lattice(2000, 1000, 0.8, 0.05, 0.7, 3)
    with pytest.raises(abi.MdxError):
        srv.submit(abi.make_job(abi.make_params(Lx, Ly), abi.NVT, 10), f)
    f, Lx, Ly = abi.gen_lattice(500, 250, 0.8, 0.05, 0.7, 3)
    with pytest.raises(abi.MdxError):
        srv.submit(abi.make_job(abi.make_params(Lx, Ly), abi.FIRE, 10), f)   # no FIRE slots
    with pytest.raises(abi.MdxError):
        srv.submit(abi.make_job(abi.make_params(Lx, Ly), abi.NVT, 0), f)
    srv.close()


def test_streamed_frames_equal_alone_runs_at_their_steps():
    """Jobs with a frame interval stream a frame at the end of every tick in
    which their step count reached or passed a multiple of it; each streamed
    frame equals the same system run alone to the step it reports.  A FIRE
    job streams while it runs; frames not taken are kept up to frame_queue,
    newest last."""
    quantum = 32
    specs = [(abi.NVT, 1500, 0.80, 0.7, 0.5, 130, 0.0, 41), (abi.NPH, 1200, 0.78, 0.7, 0.5, 90, 1.5, 42),
             (abi.NVE, 900, 0.75, 0.6, 0.3, 100, 0.0, 43), (abi.FIRE, 700, 0.85, 0.0, 0.5, 150, 0.0, 44)]
    every = [25, 40, 20, 30]
    frames = [make(s) for s in specs]
    srv = abi.Server(min_class_log2=9, max_class_log2=12, md_slots_per_class=2, fire_slots_per_class=1,
                     quantum=quantum, list_slots_per_class=2, list_skin=LIST_SKIN, frame_queue=3)
    ids = []
    for spec, (f, Lx, Ly, kw), fe in zip(specs, frames, every):
        kind, n, rho, T, eab, nsteps, pext, seed = spec
        job = abi.make_job(abi.make_params(Lx, Ly, **kw), kind, nsteps, p_ext=pext, force_tol=1e-6, fire=FP,
                           frame_interval=fe)
        ids.append(srv.submit(job, f))
    got = {i: [] for i in range(len(specs))}
    done, ticks = set(), 0
    while len(done) < len(specs):
        _, _, fin = srv.tick()
        done.update(fin)
        ticks += 1
        assert ticks < 1000
        for i in range(3):   # the MD jobs' frames are taken as they come
            while (fr := srv.frame(ids[i])) is not None:
                got[i].append(fr)
    say(f"streamed: {[[s for s, _ in g] for g in got.values()]} in {ticks} ticks")
    for i in range(3):
        spec, (f, Lx, Ly, kw), fe = specs[i], frames[i], every[i]
        kind, nsteps, pext = spec[0], spec[5], spec[6]
        steps = [s for s, _ in got[i]]
        assert steps == sorted(set(steps)) and steps[-1] <= nsteps
        # every multiple of the interval is covered by the first frame at or
        # after it, within one tick
        for m in range(fe, nsteps + 1, fe):
            first = min(s for s in steps if s >= m)
            assert first - m < quantum, (i, m, steps)
        s = abi.Simulation(abi.make_params(Lx, Ly, precision=abi.PREC_COMP32, **kw), f)
        if kind in (abi.NVE, abi.NVT, abi.NPH):
            s.set_tuning(1.0, 1, LIST_SKIN)
        prev = 0
        for step, fr in got[i]:
            s.step(kind, step - prev, p_ext=pext)
            prev = step
            ref = s.collect()
            assert np.array_equal(fr["pos"], ref["pos"]) and np.array_equal(fr["vel"], ref["vel"]), (i, step)
        s.close()
    # FIRE: frames not taken until the end; at most three remain, oldest first
    info = srv.info(ids[3])
    assert info.state == abi.JOB_DONE and info.n_frames <= 3
    fire_frames = []
    while (fr := srv.frame(ids[3])) is not None:
        fire_frames.append(fr)
    steps = [s for s, _ in fire_frames]
    assert steps == sorted(set(steps))
    assert all(np.isfinite(fr["pos"]).all() for _, fr in fire_frames)
    st = srv.stats()
    streamed = sum(len(g) for g in got.values()) + len(fire_frames)
    assert st.frames_streamed == streamed + st.frames_dropped
    srv.close()
