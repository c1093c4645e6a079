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


def test_spmd_driver_compiles(tmp_path):
    """The MPI-style driver (one thread per rank) builds against the headers."""
    exe = tmp_path / "spmd_test_c"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-pthread", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_api", "spmd_test_c.c"), "-o", str(exe),
                    "-L", LIBDIR, "-ldecomp_mpi_amd", "-Wl,-rpath," + LIBDIR], check=True)


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
