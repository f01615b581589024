"""The FP32 cell pass (cell list every step) on the N=1,048,576 phase-separated
start of the reference constructor's init_particles, whose top rows of cells
are empty: the start of such a cell equals N, and the pass must not read at
that index.  Its forces agree with the library's FP64 form (bit-identical to
the FP64 oracle in the parity tests) to FP32 rounding, and a few Nose-Hoover
steps on the cell path stay finite."""
import math
import sys

import numpy as np
import pytest

from md_simulation_cuda_amd import abi

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_cell_pass_with_empty_trailing_cells():
    n = 1048576
    Ly = math.sqrt(n / 0.75 / 2.0)
    Lx = 2.0 * Ly
    frame = abi.init_particles(n, n // 2, Lx, Ly, 0.6, devide_p=0.5)
    kw = dict(eps_ab=0.5, dt=0.005, T_target=0.6, Q=10.0)
    say("cell pass, N=1048576 init_particles start")
    ref = abi.Simulation(abi.make_params(Lx, Ly, precision=abi.PREC_FP64, **kw), frame)
    a64 = ref.compute_forces(True)
    ncx, ncy, keys, cs, ids = ref.cell_list(True)
    ref.close()
    counts = np.diff(cs.astype(np.int64))
    assert counts[-1] == 0   # the last cell is empty, so its start is N
    s = abi.Simulation(abi.make_params(Lx, Ly, precision=abi.PREC_COMP32, **kw), frame)
    s.set_tuning(1.0, 1, 0.0)
    a32 = s.compute_forces(True)
    # on the near-perfect lattice the net forces nearly cancel (rms |F| ~ 0.1),
    # so the errors are stated in absolute terms (pair forces here are O(1))
    err = np.sqrt(((a32 - a64) ** 2).sum(1))
    say(f"  max abs {err.max():.2e} rms abs {np.sqrt((err ** 2).mean()):.2e}")
    assert np.sqrt((err ** 2).mean()) < 1e-4
    assert err.max() < 1e-2   # pairs at the cutoff within FP32 rounding may count apart
    s.step(abi.NVT, 20)
    ob = s.observables()
    s.close()
    assert np.isfinite(ob.K) and np.isfinite(ob.U) and np.isfinite(ob.P)
