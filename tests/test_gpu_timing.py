"""mdx_time_pair_pass: the pair kernel timed back to back on the current
state leaves that state unchanged, so a run that calls it mid-window
continues bit for bit as a run that does not."""
import sys

import numpy as np
import pytest

from md_simulation_cuda_amd import abi
from tests.helpers import lattice_frame

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("skin,window", [(0.65, 16), (0.0, 0)])
def test_time_pair_pass_leaves_state_unchanged(skin, window):
    f, Lx, Ly = lattice_frame(20000, 0.75, seed=7, jitter=0.05, T=0.8, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.8, Q=10.0)
    say(f"time_pair_pass: skin={skin} window={window}")
    sims = []
    for _ in range(2):
        s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
        s.set_tuning(1.0, 1, skin, window)
        s.step(abi.NVT, 40)   # mid-window for windows of 16
        sims.append(s)
    a, b = sims
    ms = a.time_pair_pass(20)
    say(f"  pair pass {ms * 1e3:.1f} us")
    assert ms > 0.0
    for s in sims:
        s.step(abi.NVT, 30)
    ga, gb = a.collect(), b.collect()
    oa, ob = a.observables(), b.observables()
    for s in sims:
        s.close()
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(ga[k], gb[k]), k
    assert (oa.K, oa.U, oa.P, oa.xi) == (ob.K, ob.U, ob.P, ob.xi)


def _hip():
    import ctypes
    for name in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.skip("libamdhip64 not found")


def test_caller_stream_gives_the_same_run():
    import ctypes
    hip = _hip()
    st = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(st)) == 0
    f, Lx, Ly = lattice_frame(12000, 0.75, seed=11, jitter=0.05, T=0.7, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.7, Q=10.0)
    say("caller stream")
    a = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    b = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    for s in (a, b):
        s.set_tuning(1.0, 1, 0.5, 12)
    own = b.get_stream()
    assert own != 0
    b.set_stream(st.value)
    assert b.get_stream() == st.value
    for s in (a, b):
        s.step(abi.NVT, 30)
    b.set_stream(0)   # back to its own stream
    assert b.get_stream() == own
    for s in (a, b):
        s.step(abi.NVT, 20)
    ga, gb = a.collect(), b.collect()
    a.close()
    b.close()
    assert hip.hipStreamDestroy(st) == 0
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(ga[k], gb[k]), k
