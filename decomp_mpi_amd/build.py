"""Build libdecomp_mpi_amd.so in-tree: host C++ with g++, kernels with hipcc
for gfx950, linked by hipcc.  Run as `python -m decomp_mpi_amd.build`."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdecomp_mpi_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("DM_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-O2", "-g", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{INCLUDE}", f"-I{CSRC}"]
# Packed FP32 instructions are switched off for all device code.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             f"-I{INCLUDE}", f"-I{CSRC}"]


# FFT pass kernel families: launcher name, real type, inverse, powers of two
# only, direct-sum stages (primes above 13)
FFT_FAMILIES = [(f"launch_{r[0]}_{d}_{k}", r, "true" if d == "inv" else "false", "true" if k == "p2" else "false",
                 "true" if k == "any" else "false")
                for r in ("double", "float") for d in ("fwd", "inv") for k in ("p2", "mx", "any")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def sources():
    srcs = sorted(os.listdir(CSRC))
    return ([os.path.join(CSRC, s) for s in srcs if s.endswith(".cpp")],
            [os.path.join(CSRC, s) for s in srcs if s.endswith(".hip")])


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    cpp, hip = sources()
    jobs = []
    for s in cpp:
        o = os.path.join(OBJDIR, os.path.basename(s) + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", "-o", o, s])
    for s in hip:
        if os.path.basename(s) == "fft_inst.hip":
            # one translation unit per kernel family (they build in parallel)
            for name, real, inv, pow2, anyp in FFT_FAMILIES:
                o = os.path.join(OBJDIR, f"fft_inst_{name}.o")
                defs = [f"-DDM_INST_NAME={name}", f"-DDM_INST_R={real}", f"-DDM_INST_INV={inv}",
                        f"-DDM_INST_POW2={pow2}", f"-DDM_INST_ANY={anyp}"]
                jobs.append([f"{ROCM}/bin/hipcc", *HIP_FLAGS, *defs, "-c", "-o", o, s])
            continue
        o = os.path.join(OBJDIR, os.path.basename(s) + ".o")
        jobs.append([f"{ROCM}/bin/hipcc", *HIP_FLAGS, "-c", "-o", o, s])
    # the heaviest (complex-double) families first
    jobs.sort(key=lambda j: 0 if any("DM_INST_R=double" in a for a in j) else 1)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[j.index("-o") + 1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([f"{ROCM}/bin/hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          f"-Wl,-rpath,{ROCM}/lib", "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
