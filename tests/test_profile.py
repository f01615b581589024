"""N profile (get_N_profile, src/md_env.cu:3920-3988; kernel
src/md_cuda_common.cu:827-877): counts of type 0 and of the other types in
slabs along x.  The oracle is pinned on hand-placed particles (bin edges, the
x = Lx clamp, x outside the box, the type split, the count sums); the GPU
counts equal the oracle's bit for bit after NVT and NPH runs."""
import sys

import numpy as np
import pytest

from md_simulation_cuda_amd import abi
from oracle import oracle as O
from tests.helpers import lattice_frame


def test_oracle_n_profile_hand_placed():
    Lx = 10.0
    # bins of width 2.5 with 4 bins: 0 -> 0, 2.5 -> 1 (an edge goes up),
    # 2.4999 -> 0, 9.99 -> 3, 10.0 -> 4 clamped to 3, 10.5 and -0.1 outside
    x = [0.0, 2.5, 2.4999, 9.99, 10.0, 10.5, -0.1, 5.0]
    t = [0, 0, 1, 1, 0, 0, 1, 7]   # any type other than 0 counts as B
    got = O.n_profile(x, t, Lx, 4)
    assert got.tolist() == [1, 1, 0, 1] + [1, 0, 1, 1]
    assert O.n_profile(x, t, Lx, 1).tolist() == [3, 3]


def test_oracle_n_profile_sums():
    rng = np.random.default_rng(3)
    Lx = 37.5
    x = rng.uniform(0.0, Lx, 5000)
    t = rng.integers(0, 2, 5000)
    for nb in (1, 3, 64, 1000):
        p = O.n_profile(x, t, Lx, nb)
        assert p[:nb].sum() == (t == 0).sum() and p[nb:].sum() == (t != 0).sum()
        # each bin holds exactly the particles in [k Lx / nb, (k + 1) Lx / nb) up to
        # the rounding of the bin formula at the edges
        edges = np.floor(x / Lx * nb).astype(int)
        assert np.array_equal(p[:nb], np.bincount(edges[t == 0], minlength=nb))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [abi.NVT, abi.NPH])
def test_n_profile_equals_oracle(kind):
    f, Lx, Ly = lattice_frame(20000, 0.75, seed=21, jitter=0.05, T=0.9, aspect=2.0)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.9, Q=10.0, barostat_mass=50.0)
    say(f"N profile kind={kind}")
    s = abi.Simulation(abi.make_params(Lx, Ly, **kw), f)
    s.step(kind, 60, 1.0 if kind == abi.NPH else 0.0)
    g = s.collect()
    L = s.observables().Lx
    for nb in (1, 7, 64, 1000, 8192):
        got = s.N_profile(nb)
        ref = O.n_profile(g["pos"][:, 0], g["type"], L, nb)
        assert np.array_equal(got, ref), nb
    with pytest.raises(abi.MdxError):
        s.N_profile(8193)
    s.close()
