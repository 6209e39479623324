"""Build libddns.so (HIP, gfx950) and the Fortran host driver in-tree.

Usage: python -m mpi_2d_domain_decomp_turb_particles_amd.build [--no-fortran] [-j N]

Every HIP source is compiled for gfx950 only, with packed-FP32 instruction
selection disabled.  collide.hip is compiled with -ffp-contract=off so that the
double-precision pair tests are the same IEEE operations as the reference's.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
FDIR = PKG / "fortran"
BUILD = PKG / "_build"
LIB = PKG / "libddns.so"
DRIVER = PKG / "turb_gpu"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FC", "/opt/rocm/bin/amdflang")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]
SOURCES = {
    "flow.hip": [],
    "particles.hip": [],
    "hdi.hip": [],
    "collide.hip": ["-ffp-contract=off"],
    "spectrum.hip": [],
    "ddns_api.hip": [],
}


def _compile(src: str, extra: list[str]) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build_library(jobs: int = 5) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libddns.so failed")
    os.replace(tmp, LIB)
    return LIB


def build_fortran() -> Path | None:
    """Fortran host driver (turb_gpu): reads parameter.d and calls the C ABI."""
    if not Path(FLANG).exists():
        return None
    BUILD.mkdir(exist_ok=True)
    mod = FDIR / "ddns_mod.f90"
    drv = FDIR / "turb_gpu.f90"
    cmd = [FLANG, "-O2", "-J", str(BUILD), str(mod), str(drv), "-o", str(DRIVER),
           f"-L{PKG}", "-lddns", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=str(BUILD))
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("flang build of the host driver failed")
    return DRIVER


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-fortran", action="store_true")
    ap.add_argument("-j", type=int, default=5)
    a = ap.parse_args(argv)
    print("built", build_library(a.j))
    if not a.no_fortran:
        d = build_fortran()
        print("built", d if d else "(no flang: Fortran driver skipped)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
