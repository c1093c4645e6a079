"""The reference's driver logic as MPI-style programs (one host thread per
rank over dm_spmd_run), through the C ABI:

* test.c (test.c:92-231): each rank builds its own decompositions, passes
  its own local arrays to the collective runReshapePlan and checks its own
  elements (the test.c:16-45 known answer); error counts summed over ranks.
* split_plan_test.c (split_plan_test.c:240-330): makeSplitPlan on the
  processor grid makeProcessorGrid picks, the array of global indices in,
  forward, the Timers line, backward, the round trip checked on every
  rank; the initial and transformed arrays written in global row-major
  order (dm_file_write) and checked here against numpy's FP64 DFT.
* plan_compare.c (plan_compare.c:270-391): a slab plan and an autotuned
  plan (autotuneParallelPlan, collective) timed with the largest time over
  the ranks, the comparison line with A_factors and every axis' proc_size
  of both plans' A and B decompositions, and the autotuned plan's output
  written through getTransformed and checked against numpy.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import decomp_mpi_amd as dm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "decomp_mpi_amd")


def build(tmp_path_factory, name):
    if not dm.lib().dm_device_available():
        pytest.skip("no HIP device")
    path = tmp_path_factory.mktemp("spmd") / name
    subprocess.run(["gcc", "-std=c11", "-O1", "-pthread", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_api", name + ".c"), "-o", str(path),
                    "-L", LIBDIR, "-ldecomp_mpi_amd", "-Wl,-rpath," + LIBDIR, "-lm"], check=True)
    return path


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return build(tmp_path_factory, "spmd_test_c")


@pytest.fixture(scope="module")
def split_exe(tmp_path_factory):
    return build(tmp_path_factory, "split_plan_test_c")


@pytest.fixture(scope="module")
def compare_exe(tmp_path_factory):
    return build(tmp_path_factory, "plan_compare_c")


@pytest.mark.parametrize("npe", [1, 2, 3, 4, 6])
def test_spmd_test_c(exe, npe):
    out = subprocess.run([str(exe), str(npe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "errors=0 OK" in out.stdout


def spectrum_of_indices(n):
    return np.fft.fftn(np.arange(n ** 3, dtype=np.float64).reshape(n, n, n))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("size,npe,slab,phase", [(16, 4, 0, 0), (16, 4, 0, 1), (32, 8, 0, 0), (32, 8, 1, 1),
                                                 (24, 6, 0, 1)])
def test_split_plan_test_c(split_exe, tmp_path, size, npe, slab, phase):
    print(f"split_plan_test_c size {size} npe {npe} slab {slab} phase {phase}", flush=True)
    out = subprocess.run([str(split_exe), str(size), str(npe), str(slab), str(phase), str(tmp_path)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"size={size} npe={npe} errors=0 OK" in out.stdout
    m = re.search(r"^split_timings (\S+) (\S+) " + " ".join([r"(\S+)"] * 6), out.stdout, re.M)
    assert m, out.stdout
    t_all, r1, f1, tw, r2, f2 = (float(v) for v in m.groups()[2:])
    assert t_all > 0
    if phase:   # every phase timed on its own; a slab plan (A_d or B_d = 1 on
        # every axis) has only trivial twiddles, which the reference drops
        # with their phase (split_fft.c:396-420)
        assert min(r1, f1, r2, f2) > 0 and (tw > 0 or slab), m.group(0)
        assert abs(t_all - (r1 + f1 + tw + r2 + f2)) <= 1e-6 * t_all + 1e-9
    N = size ** 3
    init = np.fromfile(tmp_path / f"initial_split_{size}l_{npe}p.dat", dtype=np.complex128)
    assert np.array_equal(init, np.arange(N, dtype=np.float64).astype(np.complex128))
    spec = np.fromfile(tmp_path / f"parallel_split_{size}l_{npe}p.dat", dtype=np.complex128)
    assert rel(spec.reshape(size, size, size), spectrum_of_indices(size)) < 1e-12


@pytest.mark.parametrize("size,npe", [(16, 4), (32, 8)])
def test_plan_compare_c(compare_exe, tmp_path, size, npe):
    print(f"plan_compare_c size {size} npe {npe}", flush=True)
    out = subprocess.run([str(compare_exe), str(size), str(npe), str(tmp_path)], capture_output=True, text=True,
                         timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    assert f"size={size} npe={npe} OK" in out.stdout
    m = re.search(r"^compare npe (\d+) initial-grid .* slab-A (\d+) (\d+) (\d+) slab-grid-a .* auto-A (\d+) (\d+) (\d+) "
                  r"auto-grid-A (\d+) (\d+) (\d+) auto-grid-B (\d+) (\d+) (\d+) auto-time (\S+) (\S+)", out.stdout,
                  re.M)
    assert m, out.stdout
    g = [int(v) for v in m.groups()[:13]]
    assert g[0] == npe
    assert g[1:4] == [size, 1, 1]                      # slab: B = (1, size, size)
    auto_A = g[4:7]
    assert all(size % a == 0 for a in auto_A)
    assert 1 <= np.prod(g[7:10]) <= npe and 1 <= np.prod(g[10:13]) <= npe
    assert float(m.group(14)) > 0 and float(m.group(15)) > 0
    name = f"parallel_split_{auto_A[0]}-{auto_A[1]}-{auto_A[2]}_{size}l_{npe}p.dat"
    spec = np.fromfile(tmp_path / name, dtype=np.complex128)
    assert rel(spec.reshape(size, size, size), spectrum_of_indices(size)) < 1e-12
