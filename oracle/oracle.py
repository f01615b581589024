"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/libmdx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libmdx_oracle.so"

PARTICLE_DTYPE = np.dtype(
    {
        "names": ["pos", "vel", "acc", "type", "pad", "theta"],
        "formats": [("<f8", 2), ("<f8", 2), ("<f8", 2), "<i4", "<i4", "<f8"],
        "offsets": [0, 16, 32, 48, 52, 56],
        "itemsize": 64,
    }
)


class OParams(C.Structure):
    _fields_ = [
        ("Lx", C.c_double), ("Ly", C.c_double),
        ("sigma_aa", C.c_double), ("sigma_bb", C.c_double), ("sigma_ab", C.c_double),
        ("eps_aa", C.c_double), ("eps_bb", C.c_double), ("eps_ab", C.c_double),
        ("mass_a", C.c_double), ("mass_b", C.c_double), ("cutoff", C.c_double), ("dt", C.c_double),
        ("T_target", C.c_double), ("Q", C.c_double), ("barostat_mass", C.c_double),
        ("k_piston", C.c_double), ("barostat_area_rate_init", C.c_double),
        ("barostat_height_min_ratio", C.c_double), ("barostat_height_max_ratio", C.c_double),
        ("devide_p", C.c_double), ("cell_scale", C.c_double),
        ("force_mode", C.c_int32), ("n_threads", C.c_int32),
    ]


class OObs(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("K", "U", "W", "P", "xi", "p_A", "v_p", "F_p", "Lx", "Ly", "t", "last_P", "area_rate", "eta")]


class OFireParams(C.Structure):
    _fields_ = [("dt_init", C.c_double), ("dt_max", C.c_double), ("alpha0", C.c_double),
                ("f_inc", C.c_double), ("f_dec", C.c_double), ("N_min", C.c_int32),
                ("energy_interval", C.c_int32), ("max_step", C.c_double)]


class OFireResult(C.Structure):
    _fields_ = [("steps", C.c_int32), ("converged", C.c_int32), ("max_force", C.c_double),
                ("potential_energy", C.c_double), ("n_energy", C.c_int32)]


_P = C.c_void_p
_SIG = {
    "orc_create": (_P, [C.POINTER(OParams), _P, C.c_int64]),
    "orc_destroy": (None, [_P]),
    "orc_get_frame": (None, [_P, _P]),
    "orc_set_frame": (None, [_P, _P]),
    "orc_set_force_mode": (None, [_P, C.c_int32]),
    "orc_compute_forces": (None, [_P, C.c_int32]),
    "orc_pair_pass": (None, [_P, C.c_int32, _P, _P, _P, _P, C.c_double]),
    "orc_total_K": (C.c_double, [_P]),
    "orc_total_U": (C.c_double, [_P]),
    "orc_instant_pressure": (C.c_double, [_P]),
    "orc_partial_U_lambda": (C.c_double, [_P, C.c_double]),
    "orc_get_obs": (None, [_P, C.POINTER(OObs)]),
    "orc_step_nve": (None, [_P, C.c_int32]),
    "orc_step_nvt": (None, [_P, C.c_int32, C.c_int32]),
    "orc_middle_reflect": (None, [_P]),
    "orc_pressure_profile": (None, [_P, C.c_int32, C.c_int32, _P]),
    "orc_step_nph": (None, [_P, C.c_int32, C.c_double]),
    "orc_step_piston": (None, [_P, C.c_int32, C.c_double]),
    "orc_fire": (None, [_P, C.POINTER(OFireParams), C.c_double, C.c_int32, C.POINTER(OFireResult),
                        _P, _P, C.c_int32]),
    "orc_cell_list": (C.c_int64, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P, _P, _P]),
    "orc_pbc_wrap": (C.c_double, [C.c_double, C.c_double]),
    "orc_pair": (None, [C.c_double, C.c_double, C.c_int32, C.c_int32, C.POINTER(OParams),
                        C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                        C.POINTER(C.c_int32)]),
}

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built at {LIB_PATH}; run make -C oracle")
        lib = C.CDLL(str(LIB_PATH))
        for k, (r, a) in _SIG.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def make_params(Lx: float, Ly: float, **kw) -> OParams:
    d = dict(sigma_aa=1.0, sigma_bb=1.0, sigma_ab=1.0, eps_aa=1.0, eps_bb=1.0, eps_ab=0.25,
             mass_a=1.0, mass_b=1.0, cutoff=2.5, dt=1e-3, T_target=1.0, Q=100.0,
             barostat_mass=1.0e4, k_piston=100.0, barostat_area_rate_init=0.0,
             barostat_height_min_ratio=0.6, barostat_height_max_ratio=1.4, devide_p=0.5,
             cell_scale=1.0, force_mode=1, n_threads=0)
    d.update(kw)
    return OParams(Lx=Lx, Ly=Ly, **d)


def pair(dx: float, dy: float, ti: int, tj: int, p: OParams):
    fx, fy, u = C.c_double(), C.c_double(), C.c_double()
    inc = C.c_int32()
    load().orc_pair(dx, dy, ti, tj, C.byref(p), C.byref(fx), C.byref(fy), C.byref(u), C.byref(inc))
    return fx.value, fy.value, u.value, bool(inc.value)


def pbc_wrap(x: float, L: float) -> float:
    return load().orc_pbc_wrap(x, L)


class OracleSim:
    def __init__(self, params: OParams, frame: np.ndarray):
        self._lib = load()
        frame = np.ascontiguousarray(frame, dtype=PARTICLE_DTYPE)
        self.n = len(frame)
        self.params = params
        self._h = self._lib.orc_create(C.byref(params), frame.ctypes.data, self.n)

    def close(self):
        if self._h:
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frame(self) -> np.ndarray:
        out = np.zeros(self.n, PARTICLE_DTYPE)
        self._lib.orc_get_frame(self._h, out.ctypes.data)
        return out

    def set_frame(self, frame: np.ndarray) -> None:
        frame = np.ascontiguousarray(frame, dtype=PARTICLE_DTYPE)
        self._lib.orc_set_frame(self._h, frame.ctypes.data)

    def set_force_mode(self, mode: int) -> None:
        self._lib.orc_set_force_mode(self._h, mode)

    def compute_forces(self, periodic_y: bool = True) -> None:
        self._lib.orc_compute_forces(self._h, 1 if periodic_y else 0)

    def pair_pass(self, periodic_y: bool = True, eps_lambda: float = 1.0):
        acc = np.zeros((self.n, 2))
        u = np.zeros(self.n)
        w = np.zeros(self.n)
        lam = np.zeros(self.n)
        self._lib.orc_pair_pass(self._h, 1 if periodic_y else 0, acc.ctypes.data, u.ctypes.data,
                                w.ctypes.data, lam.ctypes.data, eps_lambda)
        return acc, u, w, lam

    def K(self) -> float:
        return self._lib.orc_total_K(self._h)

    def U(self) -> float:
        return self._lib.orc_total_U(self._h)

    def pressure(self) -> float:
        return self._lib.orc_instant_pressure(self._h)

    def partial_U_lambda(self, eps: float) -> float:
        return self._lib.orc_partial_U_lambda(self._h, eps)

    def obs(self) -> OObs:
        o = OObs()
        self._lib.orc_get_obs(self._h, C.byref(o))
        return o

    def step_nve(self, n: int = 1) -> None:
        self._lib.orc_step_nve(self._h, n)

    def step_nvt(self, n: int = 1, middle_wrap: bool = False) -> None:
        self._lib.orc_step_nvt(self._h, n, 1 if middle_wrap else 0)

    def middle_reflect(self) -> None:
        self._lib.orc_middle_reflect(self._h)

    def pressure_profile(self, n_bins: int, periodic_y: bool = True) -> np.ndarray:
        out = np.zeros(3 * n_bins)
        self._lib.orc_pressure_profile(self._h, n_bins, 1 if periodic_y else 0, out.ctypes.data)
        return out

    def step_nph(self, n: int, p_ext: float) -> None:
        self._lib.orc_step_nph(self._h, n, p_ext)

    def step_piston(self, n: int, p_ext: float) -> None:
        self._lib.orc_step_piston(self._h, n, p_ext)

    def fire(self, dt_init, dt_max, alpha0=0.1, f_inc=1.1, f_dec=0.5, N_min=5, energy_interval=1000,
             max_step=0.0, force_tol=1e-3, max_steps=1000, cap=64):
        fp = OFireParams(dt_init, dt_max, alpha0, f_inc, f_dec, N_min, energy_interval, max_step)
        res = OFireResult()
        es = np.zeros(cap, np.int32)
        et = np.zeros(cap)
        self._lib.orc_fire(self._h, C.byref(fp), force_tol, max_steps, C.byref(res),
                           es.ctypes.data, et.ctypes.data, cap)
        k = min(res.n_energy, cap)
        return res, es[:k].copy(), et[:k].copy()

    def cell_list(self, periodic_y: bool = True):
        ncx, ncy = C.c_int32(), C.c_int32()
        ncell = self._lib.orc_cell_list(self._h, 1 if periodic_y else 0, C.byref(ncx), C.byref(ncy),
                                        None, None, None)
        keys = np.zeros(self.n, np.uint32)
        start = np.zeros(ncell + 1, np.uint32)
        srt = np.zeros(self.n, np.uint32)
        self._lib.orc_cell_list(self._h, 1 if periodic_y else 0, C.byref(ncx), C.byref(ncy),
                                keys.ctypes.data, start.ctypes.data, srt.ctypes.data)
        return ncx.value, ncy.value, keys, start, srt


def n_profile(x, types, Lx, n_bins):
    """Counts per x slab, type 0 then the other types (2 n_bins values): the
    reference's local_density_profile_kernel (src/md_cuda_common.cu:827-877,
    driven by get_N_profile, src/md_env.cu:3920-3988) with one rank,
    xmin = 0, xmax = Lx: bin = int((x - xmin) / (xmax - xmin) * n_bins),
    clamped to [0, n_bins); particles outside [xmin, xmax] are not counted."""
    x = np.asarray(x, np.float64)
    types = np.asarray(types)
    xmin, xmax = 0.0, float(Lx)
    inside = (x >= xmin) & (x <= xmax)
    b = ((x[inside] - xmin) / (xmax - xmin) * float(n_bins)).astype(np.int64)   # truncation, as int()
    b = np.clip(b, 0, n_bins - 1)
    t = types[inside]
    a = np.bincount(b[t == 0], minlength=n_bins)
    o = np.bincount(b[t != 0], minlength=n_bins)
    return np.concatenate([a, o]).astype(np.int32)
