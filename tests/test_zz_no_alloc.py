"""fdtd2d_plan_create_in, the plan that allocates nothing on the device (its
memory is the caller's, its tile lists live in the library's static
per-device array), against fdtd2d_plan_create, through the C ABI: two plans
in a row in the same device memory, the second on a smaller problem, each
bitwise equal to a plan of the same problem made the usual way; then the
refusals (a second such plan alive on the device, too little memory, tile
lists beyond the static array)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import fdtd2d_amd as F

from helpers import all_state, random_problem

pytestmark = pytest.mark.gpu


class _DeviceBuffer:
    """Device memory the test allocates as the plan's caller would: hipMalloc
    from the HIP runtime the library itself uses (the one loaded into this
    process with it)."""

    def __init__(self, nbytes: int):
        F.lib()
        path = next(ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64.so" in ln)
        self.hip = C.CDLL(path)
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.ptr = p.value
        assert self.ptr % 256 == 0

    def free(self):
        if self.ptr:
            self.hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = 0


def _problems():
    rng = np.random.default_rng(2024)
    big = random_problem(rng, nx=300, ny=260, npml=8, nsteps=53, table=True, nsrc=3, ndet=4, ncomp=2)
    big.src_table = big.src_table * 50
    nrc = 3
    big.rc_imax = np.array([40, 120, 200], np.int32)
    big.rc_i = big.rc_imax + np.arange(nrc, dtype=np.int32)
    big.rc_jcut, big.rc_level, big.rc_edge0 = 100, 0.07, 1e-3
    big.line_size, big.line_i0, big.line_j = 90, 30, int(big.src_j[0])
    big.line_begin, big.line_end = 5, 40
    small = random_problem(rng, nx=120, ny=90, npml=5, nsteps=23, table=False)
    return big, small


def _state(plan, p):
    s = all_state(plan, p.npml)
    if p.rc_i.size:
        s["rc"] = plan.read_rc().astype(np.float64)
    if p.line_size:
        s["line"] = np.asarray(plan.read_line(), np.float64)
    return s


def test_create_in_matches_create(gpu_device):
    big, small = _problems()
    assert small.footprint() < big.footprint()
    nbytes = big.footprint()
    buf = _DeviceBuffer(nbytes)
    ptr = buf.ptr
    for label, p in (("first call", big), ("second call, smaller", small)):
        with F.Plan(p, gpu_device) as ref:
            ref.run(p.nsteps)
            ref.sync()
            want = _state(ref, p)
        print(f"{label}: {p.nx}x{p.ny} P={p.npml} steps {p.nsteps}, {p.footprint()} bytes in a "
              f"{nbytes}-byte buffer", flush=True)
        plan = F.Plan.in_memory(p, gpu_device, ptr, nbytes)
        try:
            plan.run(p.nsteps - 7)
            plan.run(7)
            plan.sync()
            got = _state(plan, p)
        finally:
            plan.close()
        for k in want:
            assert np.array_equal(got[k], want[k]), f"{label}: {k} differs"
    buf.free()


def test_create_in_refusals(gpu_device):
    big, small = _problems()
    nbytes = big.footprint()
    buf = _DeviceBuffer(nbytes)
    ptr = buf.ptr
    with pytest.raises(F.FDTD2DError) as ei:
        F.Plan.in_memory(small, gpu_device, ptr, small.footprint() - 256)
    assert ei.value.code == 1  # FDTD2D_EINVAL: too little memory
    with pytest.raises(F.FDTD2DError) as ei:
        F.Plan.in_memory(small, gpu_device, ptr + 4, nbytes - 4)
    assert ei.value.code == 1  # not 256-byte aligned
    plan = F.Plan.in_memory(small, gpu_device, ptr, nbytes)
    try:
        with pytest.raises(F.FDTD2DError) as ei:
            F.Plan.in_memory(small, gpu_device, ptr, nbytes)
        assert ei.value.code == 5  # FDTD2D_ESTATE: the device's array is taken
    finally:
        plan.close()
    # the array is free again; a launch shape whose lists exceed it is refused
    # (4096x2048 at one step per launch and one-row bands: about 133000 tiles)
    wide = F.Setup.box(4096, 2048, 8, 4, 1e-3)
    need = wide.footprint()
    wbuf = _DeviceBuffer(need)
    plan = F.Plan.in_memory(wide, gpu_device, wbuf.ptr, need)
    try:
        plan.set_option("tblock", 1)
        with pytest.raises(F.FDTD2DError) as ei:
            plan.set_option("band", 1)
        assert ei.value.code == 1 and "static list array" in str(ei.value)
        plan.set_option("band", 0)
        plan.run(4)
        plan.sync()
    finally:
        plan.close()
    buf.free()
    wbuf.free()
