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
