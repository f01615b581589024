#!/usr/bin/env python3
"""Benchmark of the 2D binary-LJ step on one MI355X.

Workload (default): Nose-Hoover NVT at N = 1,048,576 in a 2:1 periodic box at
number density 0.75 (Lx = 1672.185, Ly = 836.0925), phase-separated start from
the reference constructor's init_particles (N/2 type-0 particles in the
central half), T = 0.6, Q = 10, dt = 0.005, sigma = 1, eps_AA = eps_BB = 1,
eps_AB = 0.5, cutoff 2.5 sigma.  Inputs are resident in HBM before timing.

Prints one JSON line (rank 0) with the throughput in atom-steps/s, the
roofline of the pair kernel (HIP events on the library's stream), and a CPU
baseline: the FP64 oracle timed on the host cores over a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_VECTOR_PEAK_TFLOPS = 157.3

# Algorithmic bytes per atom of one pair-kernel launch (Nose-Hoover mode,
# compensated FP32 form): own image 8 B + class 1 B + slot cell 4 B + slot id
# 4 B, velocity read 16 B + write 16 B, acceleration write 8 B, cell_start
# ~1 B amortised.  Neighbour images are re-reads of the same array.
PAIR_BYTES_PER_ATOM = {0: 58.0, 1: 62.0, 2: 70.0}
# Algorithmic bytes per atom-step of the whole step (drift kernel, cell build,
# pair kernel): x r/w 32, v r/w 32, a w/r 16 (FP32) or 32 (FP64), image 18,
# key 8, sorted slot 8.
STEP_BYTES_PER_ATOM = {0: 114.0, 1: 130.0, 2: 138.0}
FLOP_PER_PAIR = 30.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--precision", default="comp32", choices=["comp32", "acc64", "fp64"])
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64-accumulating comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def make_workload(n):
    from md_simulation_cuda_amd import abi

    rho, aspect = 0.75, 2.0
    Ly = math.sqrt(n / rho / aspect)
    Lx = aspect * Ly
    frame = abi.init_particles(n, n // 2, Lx, Ly, 0.6, devide_p=0.5)
    kw = dict(eps_aa=1.0, eps_bb=1.0, eps_ab=0.5, dt=0.005, T_target=0.6, Q=10.0, cutoff=2.5)
    return frame, Lx, Ly, kw


def in_cutoff_pairs_per_atom(rho):
    return rho * math.pi * 2.5 ** 2


def run_gpu(frame, Lx, Ly, kw, prec, steps, warmup, timed_steps=200):
    """Timed run (each step replayed from a captured hipGraph), then a shorter
    eager run with HIP events around every pair-kernel launch."""
    from md_simulation_cuda_amd import abi

    sim = abi.Simulation(abi.make_params(Lx, Ly, precision=prec, **kw), frame)
    sim.step(abi.NVT, warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(abi.NVT, steps)
    sim.sync()
    wall = time.perf_counter() - t0
    obs = sim.observables()
    sim.timing_enable(True)
    sim.step(abi.NVT, min(steps, timed_steps))
    tm = sim.timing_read()
    sim.timing_enable(False)
    sim.close()
    return wall, tm, obs


def cpu_baseline(frame, Lx, Ly, kw, budget_s):
    from oracle import oracle as O

    n = len(frame)
    sim = O.OracleSim(O.make_params(Lx, Ly, force_mode=1, **kw), frame)
    t0 = time.perf_counter()
    sim.compute_forces()
    setup = time.perf_counter() - t0
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.step_nvt(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 10000:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": n * steps / el,
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle (cell list, OpenMP {threads} threads), {steps} NVT steps of the same "
                  f"N={n} workload in {el:.1f} s (+{setup:.1f} s initial forces)",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)
    from md_simulation_cuda_amd import abi

    prec = {"comp32": abi.PREC_COMP32, "acc64": abi.PREC_ACC64, "fp64": abi.PREC_FP64}[args.precision]
    frame, Lx, Ly, kw = make_workload(args.n)
    n = len(frame)
    wall, tm, obs = run_gpu(frame, Lx, Ly, kw, prec, args.steps, args.warmup)
    value = n * args.steps / wall
    ms_step = wall * 1e3 / args.steps
    pair_ms = tm.force_ms / max(1, tm.force_launches)
    rho = n / (Lx * Ly)
    pair_bytes = PAIR_BYTES_PER_ATOM[prec] * n
    achieved = pair_bytes / (pair_ms * 1e-3) / 1e9
    flops = FLOP_PER_PAIR * in_cutoff_pairs_per_atom(rho) * n
    out = {
        "metric": "atom-steps/s @N=1M NVT",
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "fp32 pair + fp32 two-sum accumulation, fp64 state",
                  1: "fp32 pair + fp64 accumulation, fp64 state",
                  2: "fp64"}[prec],
        "data": "synthetic (reference init_particles, phase-separated, seeded)",
        "config": {
            "workload": "NVT Nose-Hoover, 2D binary LJ, N=1048576, 2:1 box rho=0.75, T=0.6, Q=10, dt=0.005",
            "N": n, "Lx": Lx, "Ly": Ly, "eps_AB": 0.5, "cutoff": 2.5,
            "precision": args.precision, "parallelism": "1 GPU, single periodic domain",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_pair (Nose-Hoover: forces + second kick + K)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_atom": PAIR_BYTES_PER_ATOM[prec],
            "avg_launch_ms": pair_ms,
            "launches": tm.force_launches,
            "eager_ms_per_step": tm.step_ms / max(1, tm.steps),
        },
        "roofline_valu": {
            "kernel": "k_pair",
            "achieved": flops / (pair_ms * 1e-3) / 1e12,
            "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": flops / (pair_ms * 1e-3) / 1e12 / FP32_VECTOR_PEAK_TFLOPS,
            "flop_per_atom": FLOP_PER_PAIR * in_cutoff_pairs_per_atom(rho),
        },
        "roofline_step": {
            "achieved": STEP_BYTES_PER_ATOM[prec] * value / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": STEP_BYTES_PER_ATOM[prec] * value / 1e9 / HBM_PEAK_GBS,
            "bytes_per_atom_step": STEP_BYTES_PER_ATOM[prec],
        },
        "physics": {"K_per_atom": obs.K / n, "U_per_atom": obs.U / n, "P": obs.P, "xi": obs.xi},
    }
    if not args.no_plain and prec == abi.PREC_COMP32:
        wall2, tm2, _ = run_gpu(frame, Lx, Ly, kw, abi.PREC_ACC64, args.steps, args.warmup)
        out["plain_fp64_accumulation"] = {
            "value": n * args.steps / wall2,
            "ms_per_step": wall2 * 1e3 / args.steps,
            "pair_avg_launch_ms": tm2.force_ms / max(1, tm2.force_launches),
            "speedup_of_compensated": wall2 / wall,
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frame, Lx, Ly, kw, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
