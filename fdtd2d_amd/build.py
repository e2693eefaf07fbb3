"""Builds the native parts in-tree.

* ``fdtd2d_amd/lib/libfdtd2d.so`` - HIP kernels for gfx950 and the C ABI
  (include/fdtd2d.h), host sources compiled by g++, kernels by hipcc;
* ``fdtd2d_amd/lib/fdtd2d`` - the command-line host driver;
* ``oracle/_build/liboracle.so`` - the CPU oracle (test infrastructure).

Run ``python -m fdtd2d_amd.build`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "fdtd2d_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = LIB / "obj"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FDTD2D_OFFLOAD_ARCH", "gfx950")

CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIPFLAGS, "-c", str(src), "-o", str(out)]
    else:
        cmd = ["g++", *CXXFLAGS, "-c", str(src), "-o", str(out)]
    _run(cmd)
    return out


def build_library(jobs: int = 8) -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    lib_srcs = sorted(p for p in CSRC.iterdir() if p.suffix in (".cpp", ".hip") and p.name != "main.cpp")
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, lib_srcs))
    so = LIB / "libfdtd2d.so"
    tmp = LIB / "libfdtd2d.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp), "-pthread"])
    os.replace(tmp, so)
    main_o = _compile(CSRC / "main.cpp")
    exe = LIB / "fdtd2d"
    _run([HIPCC, str(main_o), "-o", str(exe), f"-L{LIB}", "-lfdtd2d", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN"])
    return so


def build_oracle() -> Path:
    out = ORACLE / "_build"
    out.mkdir(parents=True, exist_ok=True)
    so = out / "liboracle.so"
    tmp = out / "liboracle.so.tmp"
    _run(["g++", *CXXFLAGS, "-shared", str(ORACLE / "oracle.cpp"), str(ORACLE / "processing.cpp"), "-o", str(tmp),
          "-pthread"])
    os.replace(tmp, so)
    return so


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f1 = ex.submit(build_library)
        f2 = ex.submit(build_oracle)
        f1.result()
        f2.result()


def clean() -> None:
    shutil.rmtree(LIB, ignore_errors=True)
    shutil.rmtree(ORACLE / "_build", ignore_errors=True)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    build_all()
    print("built", LIB / "libfdtd2d.so", ORACLE / "_build" / "liboracle.so")
