"""The headers in include/ compile as C (with and without <complex.h> ahead
of them, the two fftw_complex forms) and a C program written like the
reference's drivers links against the library and plans successfully."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "decomp_mpi_amd")


@pytest.mark.parametrize("complex_first", [False, True])
def test_c_program_plans(tmp_path, complex_first):
    exe = tmp_path / "plan_example"
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_api", "plan_example.c"), "-o", str(exe),
           "-L", LIBDIR, "-ldecomp_mpi_amd", "-Wl,-rpath," + LIBDIR]
    if complex_first:
        cmd.insert(1, "-DWITH_COMPLEX_H")
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout
    assert out.strip().endswith("OK"), out
    assert "A=(8,8,8)" in out


@pytest.mark.parametrize("name", ["spmd_test_c", "split_plan_test_c", "plan_compare_c"])
def test_spmd_driver_compiles(tmp_path, name):
    """The MPI-style drivers (one thread per rank) build against the
    headers without warnings."""
    exe = tmp_path / name
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-pthread", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_api", name + ".c"), "-o", str(exe),
                    "-L", LIBDIR, "-ldecomp_mpi_amd", "-Wl,-rpath," + LIBDIR, "-lm"], check=True)


def test_spmd_session_collectives():
    """dm_spmd_run: P threads, each its own view; all-reduce and barrier
    (no device involved)."""
    import ctypes
    import decomp_mpi_amd as dm
    C = dm.lib()
    C.dm_spmd_allreduce_sum.restype = ctypes.c_longlong
    C.dm_spmd_allreduce_sum.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    C.dm_spmd_barrier.argtypes = [ctypes.c_void_p]
    BODY = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)
    C.dm_spmd_run.argtypes = [ctypes.c_void_p, BODY, ctypes.c_void_p]
    P = 5
    seen = {}

    def body(me, arg):
        r = ctypes.c_int()
        C.dm_comm_rank(me, ctypes.byref(r))
        total = C.dm_spmd_allreduce_sum(me, r.value + 1)
        C.dm_spmd_barrier(me)
        again = C.dm_spmd_allreduce_sum(me, 1)
        seen[r.value] = (total, again)

    world = C.dm_world_create(P)
    cb = BODY(body)
    assert C.dm_spmd_run(world, cb, None) == 0
    C.dm_world_free(world)
    assert seen == {r: (P * (P + 1) // 2, P) for r in range(P)}


def test_spmd_collective_arguments_must_agree():
    """Inside a dm_spmd_run session a collective whose ranks pass different
    arguments (here: different offsets into their slots) runs nothing and
    raises an error on every rank (no device involved: the check comes
    first)."""
    import ctypes
    import numpy as np
    import decomp_mpi_amd as dm
    C = dm.lib()
    BODY = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)
    C.dm_spmd_run.argtypes = [ctypes.c_void_p, BODY, ctypes.c_void_p]
    P = 3
    world = C.dm_world_create(P)
    slot = 4096
    mem = np.zeros(P * slot * 2 // 8, np.float64)
    base = mem.ctypes.data
    assert C.dm_register(world, base, slot) == 0
    msgs = {}

    def body(me, arg):
        r = ctypes.c_int()
        C.dm_comm_rank(me, ctypes.byref(r))
        d = C.makePackedDecomp(me, 1, dm.ptrs([C.makeBlockDecomp(12, P)]))
        plan = C.makeReshapePlan(me, dm.mpi_type("int"), d, d)
        mine = base + r.value * slot
        C.dm_clear_error()
        # rank 1 passes its destination 16 bytes further in than the others
        C.runReshapePlan(me, plan, mine, mine + 2048 + (16 if r.value == 1 else 0))
        msgs[r.value] = C.dm_last_error()
        C.freeReshapePlan(plan)
        C.freeDecomp(d)

    C.dm_set_error_mode(1)
    try:
        cb = BODY(body)
        assert C.dm_spmd_run(world, cb, None) == 0
    finally:
        C.dm_set_error_mode(0)
    C.dm_unregister(base)
    C.dm_world_free(world)
    assert sorted(msgs) == list(range(P))
    for r in range(P):
        assert b"runReshapePlan: the ranks' arguments differ" in msgs[r], msgs
