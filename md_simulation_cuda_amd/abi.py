"""ctypes binding of the C ABI in include/mdx.h (libmdx.so).

The library is loaded from md_simulation_cuda_amd/_build/.  There is no
fallback: if the built HIP library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

BUILD_DIR = Path(__file__).resolve().parent / "_build"
LIB_PATH = BUILD_DIR / "libmdx.so"

MDX_OK = 0
MDX_ERR_INVALID = 1
MDX_ERR_HIP = 2
MDX_ERR_CAPACITY = 3
MDX_ERR_BOX = 4
MDX_ERR_NONFINITE = 5
MDX_ERR_STATE = 6
MDX_ERR_NODEVICE = 7

NVE, NVT, NPH, PISTON, IDLE = 0, 1, 2, 3, -1
PREC_COMP32, PREC_ACC64, PREC_FP64 = 0, 1, 2
STEP_MIDDLE_WRAP = 1
OBS_K, OBS_U, OBS_P, OBS_VIRIAL = 1, 2, 4, 8

PARTICLE_DTYPE = np.dtype(
    {
        "names": ["pos", "vel", "acc", "type", "pad", "theta"],
        "formats": [("<f8", 2), ("<f8", 2), ("<f8", 2), "<i4", "<i4", "<f8"],
        "offsets": [0, 16, 32, 48, 52, 56],
        "itemsize": 64,
    }
)


class Params(C.Structure):
    _fields_ = [
        ("Lx", C.c_double), ("Ly", C.c_double),
        ("sigma_aa", C.c_double), ("sigma_bb", C.c_double), ("sigma_ab", C.c_double),
        ("eps_aa", C.c_double), ("eps_bb", C.c_double), ("eps_ab", C.c_double),
        ("mass_a", C.c_double), ("mass_b", C.c_double),
        ("cutoff", C.c_double), ("dt", C.c_double), ("T_target", C.c_double), ("Q", C.c_double),
        ("barostat_mass", C.c_double), ("k_piston", C.c_double),
        ("barostat_area_rate_init", C.c_double),
        ("barostat_height_min_ratio", C.c_double), ("barostat_height_max_ratio", C.c_double),
        ("devide_p", C.c_double), ("cell_scale", C.c_double),
        ("precision", C.c_int32), ("device", C.c_int32),
    ]


class Obs(C.Structure):
    _fields_ = [
        ("K", C.c_double), ("U", C.c_double), ("P", C.c_double), ("W", C.c_double),
        ("xi", C.c_double), ("p_A", C.c_double), ("area_rate", C.c_double),
        ("v_p", C.c_double), ("F_p", C.c_double), ("last_P", C.c_double),
        ("Lx", C.c_double), ("Ly", C.c_double), ("t", C.c_double),
        ("piston_active", C.c_int32), ("pad", C.c_int32), ("eta", C.c_double),
    ]


class FireParams(C.Structure):
    _fields_ = [
        ("dt_init", C.c_double), ("dt_max", C.c_double), ("alpha0", C.c_double),
        ("f_inc", C.c_double), ("f_dec", C.c_double), ("N_min", C.c_int32),
        ("energy_interval", C.c_int32), ("max_step", C.c_double), ("fp64_switch", C.c_double),
    ]


class FireResult(C.Structure):
    _fields_ = [
        ("steps", C.c_int32), ("converged", C.c_int32), ("max_force", C.c_double),
        ("potential_energy", C.c_double), ("n_energy", C.c_int32), ("pad", C.c_int32),
    ]


class Timing(C.Structure):
    _fields_ = [("force_ms", C.c_double), ("force_launches", C.c_int64),
                ("step_ms", C.c_double), ("steps", C.c_int64)]


class InitParams(C.Structure):
    _fields_ = [
        ("n_particles_global", C.c_int64), ("n_particles_type0", C.c_int64),
        ("box_w", C.c_double), ("box_h", C.c_double), ("T_init", C.c_double),
        ("mass_a", C.c_double), ("mass_b", C.c_double), ("devide_p", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [("cell_scale", C.c_double), ("skin", C.c_double), ("lanes_per_particle", C.c_int32),
                ("window", C.c_int32)]


TUNE_MAX = 32


class TuneReport(C.Structure):
    _fields_ = [("best", Tuning), ("best_ms_per_step", C.c_double), ("n_tried", C.c_int32), ("pad", C.c_int32),
                ("tried", Tuning * TUNE_MAX), ("ms_per_step", C.c_double * TUNE_MAX)]

    def table(self):
        """(cell_scale, lanes_per_particle, skin, window, ms_per_step) of every candidate."""
        return [(self.tried[i].cell_scale, self.tried[i].lanes_per_particle, self.tried[i].skin,
                 self.tried[i].window, self.ms_per_step[i]) for i in range(self.n_tried)]


FIRE = 4   # job kind of the server
JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2


class ServerParams(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_double), ("min_density", C.c_double), ("max_box_growth", C.c_double),
        ("min_class_log2", C.c_int32), ("max_class_log2", C.c_int32),
        ("md_slots_per_class", C.c_int32), ("fire_slots_per_class", C.c_int32),
        ("quantum", C.c_int32), ("record_energies", C.c_int32), ("md_precision", C.c_int32),
        ("device", C.c_int32), ("list_slots_per_class", C.c_int32), ("frame_queue", C.c_int32),
        ("list_skin", C.c_double), ("fire_skin", C.c_double),
    ]


class Job(C.Structure):
    _fields_ = [
        ("params", Params), ("kind", C.c_int32), ("nsteps", C.c_int32), ("p_ext", C.c_double),
        ("force_tol", C.c_double), ("fire", FireParams), ("t_arrival", C.c_double),
        ("frame_interval", C.c_int32), ("pad", C.c_int32),
    ]


class EnergyRecord(C.Structure):
    _fields_ = [("step", C.c_int64), ("K", C.c_double), ("U", C.c_double), ("P", C.c_double)]


class JobInfo(C.Structure):
    _fields_ = [
        ("id", C.c_int64), ("state", C.c_int32), ("kind", C.c_int32), ("n", C.c_int64),
        ("steps_done", C.c_int64), ("t_arrival", C.c_double), ("t_admit", C.c_double),
        ("t_done", C.c_double), ("last", Obs), ("fire", FireResult), ("n_records", C.c_int32),
        ("n_frames", C.c_int32),
    ]


class ServerCounters(C.Structure):
    _fields_ = [
        ("atom_steps", C.c_double), ("reserved_bytes", C.c_double), ("peak_reserved_bytes", C.c_double),
        ("ticks", C.c_int64), ("completed", C.c_int64), ("queued", C.c_int64), ("slots", C.c_int64),
        ("frames_streamed", C.c_int64), ("frames_dropped", C.c_int64),
    ]


_P = C.c_void_p
_SIGNATURES = {
    "mdx_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "mdx_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "mdx_autotune": (C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(TuneReport)]),
    "mdx_batch_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "mdx_batch_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(TuneReport)]),
    "mdx_batch_fire_begin": (C.c_int, [_P, C.c_int32, C.POINTER(FireParams), C.c_double, C.c_int32]),
    "mdx_batch_fire_poll": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(FireResult)]),
    "mdx_server_default_params": (C.c_int, [C.POINTER(ServerParams)]),
    "mdx_server_create": (C.c_int, [C.POINTER(ServerParams), C.POINTER(_P)]),
    "mdx_server_destroy": (C.c_int, [_P]),
    "mdx_server_now": (C.c_int, [_P, C.POINTER(C.c_double)]),
    "mdx_server_submit": (C.c_int, [_P, C.POINTER(Job), _P, C.c_int64, C.POINTER(C.c_int64)]),
    "mdx_server_tick": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P, C.c_int32,
                                  C.POINTER(C.c_int32)]),
    "mdx_server_info": (C.c_int, [_P, C.c_int64, C.POINTER(JobInfo)]),
    "mdx_server_records": (C.c_int, [_P, C.c_int64, _P, C.c_int32, C.POINTER(C.c_int32)]),
    "mdx_server_result": (C.c_int, [_P, C.c_int64, _P, C.c_int64, C.c_int32]),
    "mdx_server_stats": (C.c_int, [_P, C.POINTER(ServerCounters)]),
    "mdx_server_frame": (C.c_int, [_P, C.c_int64, _P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "mdx_create": (C.c_int, [C.POINTER(Params), _P, C.c_int64, C.POINTER(_P)]),
    "mdx_destroy": (C.c_int, [_P]),
    "mdx_set_state": (C.c_int, [_P, _P, C.c_int64]),
    "mdx_step": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, C.c_uint32]),
    "mdx_fire": (C.c_int, [_P, C.POINTER(FireParams), C.c_double, C.c_int32, C.POINTER(FireResult),
                           _P, _P, C.c_int32]),
    "mdx_observables": (C.c_int, [_P, C.c_uint32, C.POINTER(Obs)]),
    "mdx_partial_U_lambda": (C.c_int, [_P, C.c_double, C.POINTER(C.c_double)]),
    "mdx_collect": (C.c_int, [_P, _P, C.c_int64]),
    "mdx_compute_forces": (C.c_int, [_P, C.c_int32, _P]),
    "mdx_get_n": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "mdx_set_precision": (C.c_int, [_P, C.c_int32]),
    "mdx_set_dt": (C.c_int, [_P, C.c_double]),
    "mdx_set_graph": (C.c_int, [_P, C.c_int32]),
    "mdx_set_reorder": (C.c_int, [_P, C.c_int32]),
    "mdx_cell_capacity": (C.c_int, [C.POINTER(Params), C.c_double, C.POINTER(C.c_int32)]),
    "mdx_batch_create": (C.c_int, [C.c_int32, _P, _P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "mdx_batch_destroy": (C.c_int, [_P]),
    "mdx_batch_load": (C.c_int, [_P, C.c_int32, C.POINTER(Params), _P, C.c_int64]),
    "mdx_batch_unload": (C.c_int, [_P, C.c_int32]),
    "mdx_batch_set_kind": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, C.c_uint32]),
    "mdx_batch_step": (C.c_int, [_P, C.c_int32]),
    "mdx_batch_observables": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(Obs)]),
    "mdx_batch_collect": (C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    "mdx_batch_fire": (C.c_int, [_P, C.c_int32, C.POINTER(FireParams), C.c_double, C.c_int32,
                                 C.POINTER(FireResult), _P, _P, C.c_int32]),
    "mdx_batch_sync": (C.c_int, [_P]),
    "mdx_batch_set_graph": (C.c_int, [_P, C.c_int32]),
    "mdx_batch_set_reorder": (C.c_int, [_P, C.c_int32]),
    "mdx_batch_timing_enable": (C.c_int, [_P, C.c_int32]),
    "mdx_batch_timing_read": (C.c_int, [_P, C.POINTER(Timing)]),
    "mdx_sync": (C.c_int, [_P]),
    "mdx_N_profile": (C.c_int, [_P, C.c_int32, _P]),
    "mdx_pressure_profile": (C.c_int, [_P, C.c_int32, _P]),
    "mdx_N_profile_noalloc": (C.c_int, [_P, C.c_int32, _P]),
    "mdx_middle_reflect": (C.c_int, [_P]),
    "mdx_cell_list": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P, _P, _P,
                                C.c_int64, C.c_int64]),
    "mdx_timing_enable": (C.c_int, [_P, C.c_int32]),
    "mdx_timing_read": (C.c_int, [_P, C.POINTER(Timing)]),
    "mdx_time_pair_pass": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "mdx_get_stream": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "mdx_set_stream": (C.c_int, [_P, C.c_void_p]),
    "mdx_last_error": (C.c_char_p, []),
    "mdx_version": (C.c_char_p, []),
    "mdx_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "mdx_init_particles": (C.c_int, [C.POINTER(InitParams), _P]),
    "mdx_gen_lattice": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_uint64, _P, _P]),
    "mdx_gen_uniform": (C.c_int, [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_uint64, _P, _P]),
}

_lib = None


def load() -> C.CDLL:
    """Load libmdx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libmdx.so not built at {LIB_PATH}; run python -m md_simulation_cuda_amd.build")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class MdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mdx error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != MDX_OK:
        raise MdxError(rc, load().mdx_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_count() -> int:
    c = C.c_int32(0)
    load().mdx_device_count(C.byref(c))
    return int(c.value)


def empty_frame(n: int) -> np.ndarray:
    return np.zeros(n, dtype=PARTICLE_DTYPE)


def init_particles(n: int, n_type0: int, box_w: float, box_h: float, T_init: float,
                   mass_a: float = 1.0, mass_b: float = 1.0, devide_p: float = 0.5) -> np.ndarray:
    ip = InitParams(n, n_type0, box_w, box_h, T_init, mass_a, mass_b, devide_p)
    out = empty_frame(n)
    check(load().mdx_init_particles(C.byref(ip), _ptr(out)))
    return out


def gen_lattice(n: int, n_type0: int, density: float, jitter: float, T: float, seed: int,
                mass_a: float = 1.0, mass_b: float = 1.0):
    out = empty_frame(n)
    box = np.zeros(2)
    check(load().mdx_gen_lattice(n, n_type0, density, jitter, T, mass_a, mass_b, seed, _ptr(out), _ptr(box)))
    return out, float(box[0]), float(box[1])


def gen_uniform(n: int, n_type0: int, density: float, T: float, seed: int,
                mass_a: float = 1.0, mass_b: float = 1.0):
    out = empty_frame(n)
    box = np.zeros(2)
    check(load().mdx_gen_uniform(n, n_type0, density, T, mass_a, mass_b, seed, _ptr(out), _ptr(box)))
    return out, float(box[0]), float(box[1])


def make_params(Lx: float, Ly: float, **kw) -> Params:
    """Defaults follow MDConfig (include/md_config.hpp:22-55)."""
    d = dict(sigma_aa=1.0, sigma_bb=1.0, sigma_ab=1.0, eps_aa=1.0, eps_bb=1.0, eps_ab=0.25,
             mass_a=1.0, mass_b=1.0, cutoff=2.5, dt=1e-3, T_target=1.0, Q=100.0,
             barostat_mass=1.0e4, k_piston=100.0, barostat_area_rate_init=0.0,
             barostat_height_min_ratio=0.6, barostat_height_max_ratio=1.4, devide_p=0.5,
             cell_scale=1.0, precision=PREC_COMP32, device=-1)
    d.update(kw)
    return Params(Lx=Lx, Ly=Ly, **d)


class Simulation:
    """Thin owner of one mdx_sim handle."""

    def __init__(self, params: Params, frame: np.ndarray):
        self._lib = load()
        frame = np.ascontiguousarray(frame, dtype=PARTICLE_DTYPE)
        self.n = len(frame)
        self.params = params
        h = C.c_void_p()
        check(self._lib.mdx_create(C.byref(params), _ptr(frame), self.n, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mdx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def step(self, kind: int, nsteps: int = 1, p_ext: float = 0.0, flags: int = 0) -> None:
        check(self._lib.mdx_step(self._h, kind, nsteps, p_ext, flags))

    def sync(self) -> None:
        check(self._lib.mdx_sync(self._h))

    def collect(self) -> np.ndarray:
        out = empty_frame(self.n)
        check(self._lib.mdx_collect(self._h, _ptr(out), self.n))
        return out

    def set_state(self, frame: np.ndarray) -> None:
        frame = np.ascontiguousarray(frame, dtype=PARTICLE_DTYPE)
        check(self._lib.mdx_set_state(self._h, _ptr(frame), len(frame)))

    def observables(self, mask: int = OBS_K | OBS_U | OBS_P | OBS_VIRIAL) -> Obs:
        o = Obs()
        check(self._lib.mdx_observables(self._h, mask, C.byref(o)))
        return o

    def partial_U_lambda(self, eps: float) -> float:
        v = C.c_double()
        check(self._lib.mdx_partial_U_lambda(self._h, eps, C.byref(v)))
        return float(v.value)

    def compute_forces(self, periodic_y: bool = True) -> np.ndarray:
        acc = np.zeros((self.n, 2))
        check(self._lib.mdx_compute_forces(self._h, 1 if periodic_y else 0, _ptr(acc)))
        return acc

    def set_precision(self, prec: int) -> None:
        check(self._lib.mdx_set_precision(self._h, prec))

    def N_profile(self, n_bins: int):
        """Counts per x slab: type 0 in [0, n_bins), other types after."""
        out = np.zeros(2 * n_bins, np.int32)
        check(self._lib.mdx_N_profile(self._h, n_bins, _ptr(out)))
        return out

    def pressure_profile(self, n_bins: int):
        """Pxx, Pyy, Pxy per y slab, as one array of 3 n_bins values."""
        out = np.zeros(3 * n_bins, np.float64)
        check(self._lib.mdx_pressure_profile(self._h, n_bins, _ptr(out)))
        return out

    def middle_reflect(self) -> None:
        check(self._lib.mdx_middle_reflect(self._h))

    def cell_list(self, periodic_y: bool = True):
        ncx, ncy = C.c_int32(), C.c_int32()
        check(self._lib.mdx_cell_list(self._h, 1 if periodic_y else 0, C.byref(ncx), C.byref(ncy),
                                      None, None, None, 0, 0))
        ncell = ncx.value * ncy.value
        keys = np.zeros(self.n, np.uint32)
        start = np.zeros(ncell + 1, np.uint32)
        srt = np.zeros(self.n, np.uint32)
        check(self._lib.mdx_cell_list(self._h, 1 if periodic_y else 0, C.byref(ncx), C.byref(ncy),
                                      _ptr(keys), _ptr(start), _ptr(srt), self.n, ncell + 1))
        return ncx.value, ncy.value, keys, start, srt

    def fire(self, fp: FireParams, force_tol: float, max_steps: int, cap: int = 64):
        res = FireResult()
        es = np.zeros(cap, np.int32)
        et = np.zeros(cap, np.float64)
        check(self._lib.mdx_fire(self._h, C.byref(fp), force_tol, max_steps, C.byref(res),
                                 _ptr(es), _ptr(et), cap))
        k = min(res.n_energy, cap)
        return res, es[:k].copy(), et[:k].copy()

    def set_graph(self, on: bool) -> None:
        check(self._lib.mdx_set_graph(self._h, 1 if on else 0))

    def set_reorder(self, steps: int) -> None:
        check(self._lib.mdx_set_reorder(self._h, steps))

    def set_tuning(self, cell_scale: float = 1.0, lanes_per_particle: int = 1, skin: float = 0.0,
                   window: int = 0) -> None:
        t = Tuning(cell_scale, skin, lanes_per_particle, window)
        check(self._lib.mdx_set_tuning(self._h, C.byref(t)))

    def get_tuning(self) -> Tuning:
        t = Tuning()
        check(self._lib.mdx_get_tuning(self._h, C.byref(t)))
        return t

    def autotune(self, kind: int, steps: int = 20) -> TuneReport:
        rep = TuneReport()
        check(self._lib.mdx_autotune(self._h, kind, steps, C.byref(rep)))
        return rep

    def timing_enable(self, on: bool) -> None:
        check(self._lib.mdx_timing_enable(self._h, 1 if on else 0))

    def timing_read(self) -> Timing:
        t = Timing()
        check(self._lib.mdx_timing_read(self._h, C.byref(t)))
        return t

    def get_stream(self) -> int:
        """The handle's HIP stream (hipStream_t as an integer)."""
        out = C.c_void_p()
        check(self._lib.mdx_get_stream(self._h, C.byref(out)))
        return out.value or 0

    def set_stream(self, stream: int) -> None:
        """Issue the handle's work on the caller's HIP stream (0: its own)."""
        check(self._lib.mdx_set_stream(self._h, C.c_void_p(stream or None)))

    def time_pair_pass(self, reps: int = 100) -> float:
        """Average pair-pass duration in ms, launches back to back in one
        graph (no per-launch overhead); the state is left unchanged."""
        out = C.c_double()
        check(self._lib.mdx_time_pair_pass(self._h, reps, C.byref(out)))
        return out.value


def cell_capacity(params: Params, growth: float = 1.0) -> int:
    out = C.c_int32()
    check(load().mdx_cell_capacity(C.byref(params), growth, C.byref(out)))
    return int(out.value)


class Batch:
    """Owner of one mdx_batch: many independent systems advanced together."""

    def __init__(self, n_cap, cell_cap, precision: int = PREC_COMP32, device: int = -1):
        self._lib = load()
        self.n_cap = np.ascontiguousarray(n_cap, dtype=np.int64)
        self.cell_cap = np.ascontiguousarray(cell_cap, dtype=np.int32)
        self.precision = precision
        self.n = [0] * len(self.n_cap)
        h = C.c_void_p()
        check(self._lib.mdx_batch_create(len(self.n_cap), _ptr(self.n_cap), _ptr(self.cell_cap), precision,
                                         device, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mdx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, slot: int, params: Params, frame: np.ndarray) -> None:
        frame = np.ascontiguousarray(frame, dtype=PARTICLE_DTYPE)
        check(self._lib.mdx_batch_load(self._h, slot, C.byref(params), _ptr(frame), len(frame)))
        self.n[slot] = len(frame)

    def unload(self, slot: int) -> None:
        check(self._lib.mdx_batch_unload(self._h, slot))
        self.n[slot] = 0

    def set_kind(self, slot: int, kind: int, p_ext: float = 0.0, flags: int = 0) -> None:
        check(self._lib.mdx_batch_set_kind(self._h, slot, kind, p_ext, flags))

    def step(self, nsteps: int = 1) -> None:
        check(self._lib.mdx_batch_step(self._h, nsteps))

    def sync(self) -> None:
        check(self._lib.mdx_batch_sync(self._h))

    def observables(self, slot: int, mask: int = OBS_K | OBS_U | OBS_P | OBS_VIRIAL) -> Obs:
        o = Obs()
        check(self._lib.mdx_batch_observables(self._h, slot, mask, C.byref(o)))
        return o

    def collect(self, slot: int) -> np.ndarray:
        out = empty_frame(self.n[slot])
        check(self._lib.mdx_batch_collect(self._h, slot, _ptr(out), len(out)))
        return out

    def fire(self, slot: int, fp: FireParams, force_tol: float, max_steps: int, cap: int = 64):
        res = FireResult()
        es = np.zeros(cap, np.int32)
        et = np.zeros(cap, np.float64)
        check(self._lib.mdx_batch_fire(self._h, slot, C.byref(fp), force_tol, max_steps, C.byref(res),
                                       _ptr(es), _ptr(et), cap))
        k = min(res.n_energy, cap)
        return res, es[:k].copy(), et[:k].copy()

    def fire_begin(self, slot: int, fp: FireParams, force_tol: float, max_steps: int) -> None:
        check(self._lib.mdx_batch_fire_begin(self._h, slot, C.byref(fp), force_tol, max_steps))

    def fire_poll(self, slot: int):
        done = C.c_int32()
        res = FireResult()
        check(self._lib.mdx_batch_fire_poll(self._h, slot, C.byref(done), C.byref(res)))
        return bool(done.value), res

    def set_graph(self, on: bool) -> None:
        check(self._lib.mdx_batch_set_graph(self._h, 1 if on else 0))

    def set_reorder(self, steps: int) -> None:
        check(self._lib.mdx_batch_set_reorder(self._h, steps))

    def set_tuning(self, cell_scale: float = 1.0, lanes_per_particle: int = 1, skin: float = 0.0,
                   window: int = 0) -> None:
        t = Tuning(cell_scale, skin, lanes_per_particle, window)
        check(self._lib.mdx_batch_set_tuning(self._h, C.byref(t)))

    def autotune(self, steps: int = 20) -> TuneReport:
        rep = TuneReport()
        check(self._lib.mdx_batch_autotune(self._h, steps, C.byref(rep)))
        return rep

    def timing_enable(self, on: bool) -> None:
        check(self._lib.mdx_batch_timing_enable(self._h, 1 if on else 0))

    def timing_read(self) -> Timing:
        t = Timing()
        check(self._lib.mdx_batch_timing_read(self._h, C.byref(t)))
        return t


def server_default_params() -> ServerParams:
    p = ServerParams()
    check(load().mdx_server_default_params(C.byref(p)))
    return p


def make_job(params: Params, kind: int, nsteps: int, p_ext: float = 0.0, force_tol: float = 1e-6,
             fire: FireParams | None = None, t_arrival: float = -1.0, frame_interval: int = 0) -> Job:
    j = Job()
    j.params = params
    j.kind = kind
    j.nsteps = nsteps
    j.p_ext = p_ext
    j.force_tol = force_tol
    if fire is not None:
        j.fire = fire
    j.t_arrival = t_arrival
    j.frame_interval = frame_interval
    return j


class Server:
    """Owner of one mdx_server: continuous batching of MD jobs."""

    def __init__(self, params: ServerParams | None = None, **kw):
        self._lib = load()
        p = params if params is not None else server_default_params()
        for k, v in kw.items():
            setattr(p, k, v)
        self.params = p
        self._n = {}
        h = C.c_void_p()
        check(self._lib.mdx_server_create(C.byref(p), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mdx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self) -> float:
        t = C.c_double()
        check(self._lib.mdx_server_now(self._h, C.byref(t)))
        return t.value

    def submit(self, job: Job, frame: np.ndarray) -> int:
        frame = np.ascontiguousarray(frame, dtype=PARTICLE_DTYPE)
        jid = C.c_int64()
        check(self._lib.mdx_server_submit(self._h, C.byref(job), _ptr(frame), len(frame), C.byref(jid)))
        self._n[jid.value] = len(frame)
        return jid.value

    def tick(self, cap_done: int = 256):
        nr, nq, nd = C.c_int32(), C.c_int32(), C.c_int32()
        ids = np.zeros(cap_done, np.int64)
        check(self._lib.mdx_server_tick(self._h, C.byref(nr), C.byref(nq), _ptr(ids), cap_done, C.byref(nd)))
        return nr.value, nq.value, [int(x) for x in ids[:min(nd.value, cap_done)]]

    def info(self, jid: int) -> JobInfo:
        out = JobInfo()
        check(self._lib.mdx_server_info(self._h, jid, C.byref(out)))
        return out

    def records(self, jid: int) -> np.ndarray:
        n = C.c_int32()
        check(self._lib.mdx_server_records(self._h, jid, None, 0, C.byref(n)))
        out = (EnergyRecord * max(n.value, 1))()
        check(self._lib.mdx_server_records(self._h, jid, C.cast(out, C.c_void_p), n.value, C.byref(n)))
        return np.array([(r.step, r.K, r.U, r.P) for r in out[:n.value]],
                        dtype=[("step", "<i8"), ("K", "<f8"), ("U", "<f8"), ("P", "<f8")])

    def result(self, jid: int, release: bool = True) -> np.ndarray:
        out = empty_frame(self._n[jid])
        check(self._lib.mdx_server_result(self._h, jid, _ptr(out), len(out), 1 if release else 0))
        if release:
            self._n.pop(jid, None)
        return out

    def frame(self, jid: int):
        """The oldest streamed frame of the job and its step count, or None."""
        step, got = C.c_int64(), C.c_int32()
        check(self._lib.mdx_server_frame(self._h, jid, None, 0, None, C.byref(got)))
        if got.value == 0:
            return None
        out = empty_frame(self._n[jid])
        check(self._lib.mdx_server_frame(self._h, jid, _ptr(out), len(out), C.byref(step), C.byref(got)))
        return (step.value, out) if got.value else None

    def stats(self) -> ServerCounters:
        out = ServerCounters()
        check(self._lib.mdx_server_stats(self._h, C.byref(out)))
        return out
