"""get_pressure_profile (src/md_env.cu:3821-3918, kernel
src/md_cuda_common.cu:879-1030) on the GPU against the oracle: the library
forms each pair force in FP64 from the wrapped FP64 positions exactly as the
oracle does and sums fixed-point units as integers, so the profile of a frame
is the oracle's bit for bit, for the FP64 and the compensated FP32 forms
alike (the profile pass is FP64 in both), with LDS histograms (n_bins <=
2048) and with global ones."""
import math
import sys

import numpy as np
import pytest

from md_simulation_cuda_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_COMP32])
def test_pressure_profile_equals_oracle(prec):
    n = 3000
    Lx, Ly = math.sqrt(n / 0.75 * 2.0), math.sqrt(n / 0.75 / 2.0)
    f = abi.init_particles(n, n // 2, Lx, Ly, 0.6, devide_p=0.5)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.6, Q=10.0)
    say(f"pressure profile, precision {prec}: 2:1 slab of 3000, 37 NVT steps")
    g = abi.Simulation(abi.make_params(Lx, Ly, precision=prec, **kw), f)
    if prec == abi.PREC_COMP32:
        g.set_tuning(1.0, 1, 0.65, 16)   # mid-window when the profile is taken
    g.step(abi.NVT, 37)
    profiles = {nb: g.pressure_profile(nb) for nb in (1, 7, 64, 3000)}
    fr = g.collect()
    P = g.observables(abi.OBS_P).P
    g.step(abi.NVT, 5)   # stepping on after a profile pass
    g.close()
    o = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, **kw), fr)
    for nb, pp in profiles.items():
        po = o.pressure_profile(nb)
        say(f"  n_bins {nb}: max |gpu - oracle| {np.abs(pp - po).max():.3e}")
        assert np.array_equal(pp, po), nb
        # the slabs' Pxx + Pyy sum to the scalar pressure
        assert abs((pp[:nb] + pp[nb:2 * nb]).sum() / nb / 2.0 - P) <= 1e-6 * abs(P)
    o.close()


def test_pressure_profile_through_the_cpp_class_shape():
    """n_bins <= 0 gives an empty profile through the C++ class; the C ABI
    rejects out-of-range bin counts."""
    f, L, _ = abi.gen_lattice(1024, 512, 0.8, 0.05, 0.7, seed=7)
    g = abi.Simulation(abi.make_params(L, L, precision=abi.PREC_FP64, eps_ab=0.5), f)
    with pytest.raises(abi.MdxError):
        g.pressure_profile(0)
    with pytest.raises(abi.MdxError):
        g.pressure_profile(9000)
    g.close()
