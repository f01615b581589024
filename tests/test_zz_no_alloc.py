"""mdx_N_profile_noalloc (the allocation-free N profile) against mdx_N_profile
through the C ABI: two calls in a row on different systems, the second with
the smaller input (fewer particles and fewer bins), each bit-exact with the
allocating entry point; sizes beyond the fixed array are rejected."""
import ctypes as C

import numpy as np
import pytest

from md_simulation_cuda_amd import abi

pytestmark = pytest.mark.gpu


def counts(lib, sim, nb, fn):
    out = np.full(2 * nb, -7, np.int32)
    rc = getattr(lib, fn)(sim.handle, nb, out.ctypes.data)
    assert rc == abi.MDX_OK, (fn, rc, lib.mdx_last_error())
    return out


def run_system(n, nsteps, seed, nb):
    frame, L, _ = abi.gen_lattice(n, n // 3, 0.75, 0.05, 0.8, seed=seed)
    sim = abi.Simulation(abi.make_params(L, L, eps_ab=0.5, dt=0.005, T_target=0.8, Q=10.0), frame)
    sim.step(abi.NVT, nsteps)
    lib = abi.load()
    a = counts(lib, sim, nb, "mdx_N_profile_noalloc")
    b = counts(lib, sim, nb, "mdx_N_profile")
    sim.close()
    return a, b


def test_noalloc_profile_equals_allocating_profile_twice():
    print("noalloc N profile: N=3000, 128 bins", flush=True)
    a, b = run_system(3000, 50, 11, 128)
    assert np.array_equal(a, b)
    assert a.sum() == 3000 and a[:128].sum() == 1000
    print("noalloc N profile: N=1024, 32 bins", flush=True)
    a, b = run_system(1024, 20, 12, 32)
    assert np.array_equal(a, b)
    assert a.sum() == 1024 and a[:32].sum() == 1024 // 3


def test_noalloc_profile_rejects_too_many_bins():
    frame, L, _ = abi.gen_lattice(256, 128, 0.75, 0.05, 0.8, seed=3)
    sim = abi.Simulation(abi.make_params(L, L, eps_ab=0.5), frame)
    lib = abi.load()
    out = np.zeros(2 * 8193, np.int32)
    rc = lib.mdx_N_profile_noalloc(sim.handle, 8193, out.ctypes.data)
    assert rc == abi.MDX_ERR_CAPACITY
    assert b"8192" in lib.mdx_last_error()
    sim.close()
