"""Build recipe for the in-tree native libraries (gfx950).

Outputs (all under md_simulation_cuda_amd/_build/):
  libmdx.so      the HIP library: kernels + the C ABI of include/mdx.h
  libmd_env.so   the header-compatible C++ classes (MDSimulation,
                 EnergyMinimizer, MDConfigManager, FileWriter/FileReader),
                 host code over the C ABI
  cpp_tests/*    C++ driver programs used by tests/ (reference-style drivers)
and the CPU oracle under oracle/_build/ (test infrastructure).

Run `python -m md_simulation_cuda_amd.build` or call build_all().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
OUT = PKG / "_build"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled for every HIP translation unit; FMA
# contraction is off so that the FP64 integrator and pair arithmetic follow the
# written operation order (the FP32 pair code contracts where it asks to).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{INCLUDE}"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(map(str, cmd))}")


def build_mdx() -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    target = OUT / "libmdx.so"
    tmp = OUT / "libmdx.so.tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", str(CSRC / "mdx_api.hip"), str(CSRC / "mdx_batch.hip"),
          str(CSRC / "mdx_init.cpp"), str(CSRC / "mdx_server.cpp"),
          "-o", str(tmp)])
    os.replace(tmp, target)
    return target


def build_host() -> Path | None:
    srcs = sorted(HOST.glob("*.cpp"))
    if not srcs:
        return None
    target = OUT / "libmd_env.so"
    tmp = OUT / "libmd_env.so.tmp"
    _run([CXX, *CXX_FLAGS, "-shared", *map(str, srcs), "-o", str(tmp),
          f"-L{OUT}", "-lmdx", "-lz", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, target)
    return target


def build_cpp_tests() -> list[Path]:
    tdir = ROOT / "tests" / "cpp"
    outs = []
    if not tdir.is_dir():
        return outs
    od = OUT / "cpp_tests"
    od.mkdir(parents=True, exist_ok=True)
    for src in sorted(tdir.glob("*.cpp")):
        exe = od / src.stem
        _run([CXX, *CXX_FLAGS, str(src), "-o", str(exe), f"-L{OUT}", "-lmd_env", "-lmdx", "-lz",
              "-Wl,-rpath,$ORIGIN/.."])
        outs.append(exe)
    return outs


def build_oracle() -> Path:
    make = shutil.which("make") or "make"
    _run([make, "-B", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "libmdx_oracle.so"


def build_all() -> None:
    build_mdx()
    build_host()
    build_cpp_tests()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
