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
