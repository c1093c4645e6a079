"""The host C++ layer built with sanitizers and run without a GPU.

Every host source of the library (decomposition strategies, layouts, the
reshape planner, the split-plan planners, the SPMD rendezvous, the virtual
MPI layer) is compiled again with GCC's address + undefined-behaviour
sanitizers, and separately with the thread sanitizer, then linked with the
library's own device objects (not instrumented: the sanitizers cover host
code only) into the C drivers under tests/c_api:

- plan_example.c and host_plans.c (address + undefined): thousands of
  reshape plans between packed-block, pencil and slab decompositions of
  several shapes, rank counts and element sizes, split plans in both
  precisions, a refused plan in error mode 1, every object freed; a leak,
  an out-of-bounds access or undefined behaviour fails the run;
- host_spmd.c (thread): six rank threads through the all-reduces, the
  barrier, per-rank planning and a refused collective; a data race fails
  the run.

The sanitizer runtimes are linked statically so the drivers do not depend
on the order of preloaded libraries.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from decomp_mpi_amd import build as B

CAPI = os.path.join(B.ROOT, "tests", "c_api")

SANITIZERS = {
    "address": (["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
                ["-static-libasan", "-static-libubsan"]),
    "thread": (["-fsanitize=thread"], ["-static-libtsan"]),
}
ENV = {"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1",
       "TSAN_OPTIONS": "halt_on_error=1"}


def _device_objects():
    if not os.path.isdir(B.OBJDIR):
        return []
    return sorted(os.path.join(B.OBJDIR, f) for f in os.listdir(B.OBJDIR)
                  if f.endswith(".o") and not f.endswith(".cpp.o"))


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout + r.stderr
    return r


def _host_build(name, tmp_path_factory):
    """(sanitizer, directory, the host objects built with it)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    if not _device_objects():
        pytest.skip("the library's device objects are not built (python -m decomp_mpi_amd.build)")
    cflags, _ = SANITIZERS[name]
    out = tmp_path_factory.mktemp(f"san_{name}")
    cpp, _ = B.sources()
    jobs = [["g++", *B.HOST_FLAGS, "-O1", "-w", *cflags, "-c", "-o", str(out / (os.path.basename(s) + ".o")), s]
            for s in cpp]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    return name, out, [j[j.index("-o") + 1] for j in jobs]


@pytest.fixture(scope="module")
def address_build(tmp_path_factory):
    return _host_build("address", tmp_path_factory)


@pytest.fixture(scope="module")
def thread_build(tmp_path_factory):
    return _host_build("thread", tmp_path_factory)


def _driver(host_build, source, pthread=False):
    name, out, objs = host_build
    cflags, lflags = SANITIZERS[name]
    exe = out / source.replace(".c", "")
    obj = str(exe) + ".o"
    _run(["gcc", "-std=c11", "-Wall", "-Werror", "-g", *cflags, *(["-pthread"] if pthread else []),
          "-I", B.INCLUDE, "-c", "-o", obj, os.path.join(CAPI, source)])
    _run(["g++", *cflags, *lflags, "-o", str(exe), obj, *objs, *_device_objects(),
          f"-L{B.ROCM}/lib", f"-Wl,-rpath,{B.ROCM}/lib", "-lamdhip64", "-lpthread"])
    env = dict(os.environ, **ENV)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr[-4000:]
    return r.stdout


@pytest.mark.parametrize("source", ["plan_example.c", "host_plans.c"])
def test_planning_under_address_sanitizer(address_build, source):
    out = _driver(address_build, source)
    if source == "host_plans.c":
        assert int(out.split("plans=")[1].split()[0]) > 2000, out


def test_spmd_under_thread_sanitizer(thread_build):
    _driver(thread_build, "host_spmd.c", pthread=True)
