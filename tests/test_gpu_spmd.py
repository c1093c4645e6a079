"""The reference's test.c logic as an MPI-style program (one host thread per
rank over dm_spmd_run): each rank builds its own decompositions, passes its
own local arrays to the collective runReshapePlan and checks its own
elements (test.c:16-45 known answer); error counts summed over ranks."""
import os
import subprocess

import pytest

import decomp_mpi_amd as dm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "decomp_mpi_amd")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not dm.lib().dm_device_available():
        pytest.skip("no HIP device")
    path = tmp_path_factory.mktemp("spmd") / "spmd_test_c"
    subprocess.run(["gcc", "-std=c11", "-O1", "-pthread", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_api", "spmd_test_c.c"), "-o", str(path),
                    "-L", LIBDIR, "-ldecomp_mpi_amd", "-Wl,-rpath," + LIBDIR], check=True)
    return path


@pytest.mark.parametrize("npe", [1, 2, 3, 4, 6])
def test_spmd_test_c(exe, npe):
    out = subprocess.run([str(exe), str(npe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "errors=0 OK" in out.stdout
