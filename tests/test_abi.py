"""The C-ABI library loads and exports every entry point include/fdtd2d.h
declares; the oracle library loads. No device work (runs without a GPU)."""

import ctypes as C
import re
from pathlib import Path

import fdtd2d_amd as F
from fdtd2d_amd import _lib as L
from oracle import oracle as O

HEADER = Path(__file__).resolve().parent.parent / "include" / "fdtd2d.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdtd2d_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("fdtd2d_plan_create", "fdtd2d_run", "fdtd2d_read_detectors", "fdtd2d_read_field",
                 "fdtd2d_plan_destroy", "fdtd2d_setup_from_ini", "fdtd2d_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    bound = {name for name, _, _ in L.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_status_strings_and_version():
    lib = F.lib()
    assert lib.fdtd2d_abi_version() == 4
    assert lib.fdtd2d_strerror(0) == b"ok"
    assert lib.fdtd2d_strerror(L.EINVAL) == b"invalid argument"
    assert lib.fdtd2d_strerror(99) == b"unknown status"


def test_problem_struct_layout_matches_c():
    # the ctypes mirror must agree with the C compiler's layout: a setup's
    # problem read back through it has the sizes the builder wrote
    s = F.Setup.box(64, 48, 5, 10, 1e-3)
    p = s.c_problem.contents
    assert (p.nx, p.ny, p.npml, p.nsteps) == (64, 48, 5, 10)
    assert p.medium_kind == L.MEDIUM_DENSITY and p.pec_material == -1
    assert abs(p.dx - 1e-3) < 1e-15 and p.dx_f == C.c_float(1e-3).value


def test_invalid_problem_is_rejected_before_device_work():
    s = F.Setup.box(64, 48, 5, 10, 1e-3)
    p = s.problem()
    p.src_i = p.src_i[:0]
    p.det_i = [70]  # outside the grid
    p.det_j = [3]
    st, keep = p.to_c()
    out = C.c_void_p()
    rc = F.lib().fdtd2d_plan_create(C.byref(st), 0, C.byref(out))
    assert rc == L.EINVAL and not out.value
    assert b"detector" in F.lib().fdtd2d_last_error()


def test_footprint_counts_both_buffer_sets():
    p = F.Setup.box(1000, 500, 12, 100, 1e-3).problem()
    fp = p.footprint()
    plane = 500 * 1024 * 4
    assert fp > 9 * plane  # 2 x (Ez, Hx, Hy, Jz) + medium
    assert fp < 11 * plane


def test_oracle_library_loads():
    lib = O.lib()
    assert hasattr(lib, "oracle_create") and hasattr(lib, "oracle_run")


C_CONSUMER = r"""
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include "fdtd2d.h"
int main(int argc, char **argv) {
  if (argc < 2) return 2;
  printf("version %d\n", fdtd2d_abi_version());
  printf("problem %zu %zu %zu %zu %zu\n", sizeof(fdtd2d_problem), offsetof(fdtd2d_problem, nx),
         offsetof(fdtd2d_problem, src_table), offsetof(fdtd2d_problem, det_i), offsetof(fdtd2d_problem, nrc));
  printf("processing %zu\n", sizeof(fdtd2d_processing));
  fdtd2d_setup *s = NULL;
  int rc = fdtd2d_setup_from_ini(argv[1], &s);
  if (rc) {
    printf("error %d %s\n", rc, fdtd2d_last_error());
    return 1;
  }
  const fdtd2d_problem *p = fdtd2d_setup_problem(s);
  uint64_t bytes = 0;
  rc = fdtd2d_plan_footprint(p, &bytes);
  printf("grid %d %d %d %llu %d\n", p->nx, p->ny, p->nsteps, (unsigned long long)bytes, rc);
  fdtd2d_setup_destroy(s);
  return 0;
}
"""


def test_header_is_plain_c_and_the_layouts_agree(tmp_path):
    """A C99 caller (the kind a cgo / FFI binding compiles) builds against
    include/fdtd2d.h with -pedantic -Werror, links the library, and sees the
    struct layouts the Python mirror uses; host-only calls, no device."""
    import shutil
    import subprocess

    import pytest

    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    src = tmp_path / "consumer.c"
    src.write_text(C_CONSUMER)
    exe = tmp_path / "consumer"
    libdir = L.LIB_PATH.parent
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{HEADER.parent}", str(src),
                    f"-L{libdir}", "-lfdtd2d", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    ini = Path(__file__).resolve().parent / "golden" / "settings_release.ini"
    out = subprocess.run([str(exe), str(ini)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "version 4"
    size, nx, src_table, det_i, nrc = (int(v) for v in out[1].split()[1:])
    assert size == C.sizeof(L.ProblemC)
    assert (nx, src_table, det_i, nrc) == (L.ProblemC.nx.offset, L.ProblemC.src_table.offset,
                                           L.ProblemC.det_i.offset, L.ProblemC.nrc.offset)
    assert int(out[2].split()[1]) == C.sizeof(L.ProcessingC)
    gx, gy, steps, bytes_, rc = (int(v) for v in out[3].split()[1:])
    s = F.Setup.from_ini(str(ini))
    p = s.c_problem.contents
    assert (gx, gy, steps, rc) == (p.nx, p.ny, p.nsteps, 0)
    assert bytes_ == s.footprint()
