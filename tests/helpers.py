"""Shared helpers for the tests: small seeded frames built with numpy."""
from __future__ import annotations

import numpy as np

from oracle.oracle import PARTICLE_DTYPE


def lattice_frame(n: int, density: float, seed: int, jitter: float = 0.05, T: float = 0.7,
                  frac_a: float = 0.5, aspect: float = 1.0):
    """n particles on a jittered near-square grid in an Lx:Ly = aspect box."""
    rng = np.random.default_rng(seed)
    area = n / density
    Ly = np.sqrt(area / aspect)
    Lx = aspect * Ly
    nx = int(np.ceil(np.sqrt(n * aspect)))
    ny = int(np.ceil(n / nx))
    ax, ay = Lx / nx, Ly / ny
    k = np.arange(n)
    x = (k % nx + 0.5) * ax + rng.uniform(-jitter, jitter, n)
    y = (k // nx + 0.5) * ay + rng.uniform(-jitter, jitter, n)
    f = np.zeros(n, PARTICLE_DTYPE)
    f["pos"][:, 0] = np.mod(x, Lx)
    f["pos"][:, 1] = np.mod(y, Ly)
    types = np.zeros(n, np.int32)
    types[int(round(frac_a * n)):] = 1
    rng.shuffle(types)
    f["type"] = types
    v = rng.normal(0.0, np.sqrt(T), (n, 2)) if T > 0 else np.zeros((n, 2))
    v -= v.mean(axis=0)
    f["vel"] = v
    f["theta"] = rng.uniform(0, 2 * np.pi, n)
    return f, float(Lx), float(Ly)


def uniform_frame(n: int, Lx: float, Ly: float, seed: int, T: float = 0.0):
    rng = np.random.default_rng(seed)
    f = np.zeros(n, PARTICLE_DTYPE)
    f["pos"][:, 0] = rng.uniform(0, Lx, n)
    f["pos"][:, 1] = rng.uniform(0, Ly, n)
    f["type"] = rng.integers(0, 2, n).astype(np.int32)
    if T > 0:
        f["vel"] = rng.normal(0, np.sqrt(T), (n, 2))
    return f


def min_pair_distance(frame, Lx, Ly):
    p = frame["pos"]
    d = p[:, None, :] - p[None, :, :]
    d[..., 0] -= Lx * np.round(d[..., 0] / Lx)
    d[..., 1] -= Ly * np.round(d[..., 1] / Ly)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    return r.min()


def near_cutoff_mask(frame, Lx, Ly, cutoff, sig, band=1e-6, periodic_y=True):
    """Particles with a pair whose r^2 lies within band*rc^2 of rc^2.

    sig = (sigma_aa, sigma_bb, sigma_ab).  FP32 pair arithmetic may put such a
    pair on either side of the cutoff, where the unshifted force jumps.
    """
    p = frame["pos"]
    t = frame["type"]
    n = len(p)
    mask = np.zeros(n, bool)
    for i0 in range(0, n, 512):
        sl = slice(i0, min(n, i0 + 512))
        d = p[sl, None, :] - p[None, :, :]
        d[..., 0] -= Lx * np.round(d[..., 0] / Lx)
        if periodic_y:
            d[..., 1] -= Ly * np.round(d[..., 1] / Ly)
        r2 = (d ** 2).sum(-1)
        ti = t[sl, None]
        tj = t[None, :]
        s = np.where((ti == 0) & (tj == 0), sig[0], np.where((ti == 1) & (tj == 1), sig[1], sig[2]))
        rc2 = (cutoff * s) ** 2
        hit = np.abs(r2 - rc2) < band * rc2
        mask[sl] = hit.any(axis=1)
    return mask
