#!/usr/bin/env python3
"""Benchmark: 512^3 complex 3D FFT through the split-plan API on one MI355X.

Workload: a 512^3 complex-double array in a pencil decomposition over a 4x8
grid of virtual ranks ({Local(512), Block(512,4) rank stride 8,
Block(512,8) rank stride 1}, one 64 MiB slot per rank in one HBM buffer),
uniform[-1,1) real and imaginary parts plus 8 planted modes of amplitude 10.
A step is runSplitPlan(FORWARD) then runSplitPlan(BACKWARD) on a plan from
autotuneParallelPlan.  The FP32 form (autotuneParallelPlanf on the same
input cast to complex float) is timed the same way and reported beside it.

value = algorithmic HBM bytes of a step / step time, with
algorithmic bytes = 6 * N * 16 per transform (one read and one write of the
array per transformed axis), N = 512^3.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0   # MI355X HBM3E
L = 512
P = 32
GRID = (1, 4, 8)


def make_input(seed=7):
    """Per-rank pencil blocks (512 x 128 x 64, z fastest) of the global
    array: uniform[-1,1) re/im plus 8 modes 10*exp(+2 pi i k.n / 512)."""
    rng = np.random.default_rng(seed)
    modes = rng.integers(0, L, (8, 3))
    ny, nz = L // GRID[1], L // GRID[2]
    per = L * ny * nz
    x = np.empty((P, per), np.complex128)
    e = [np.exp(2j * np.pi * np.outer(modes[:, d], np.arange(L)) / L) for d in range(3)]  # [8, L]
    for r in range(P):
        ry, rz = r // GRID[2], r % GRID[2]
        blk = rng.uniform(-1, 1, (L, ny, nz)) + 1j * rng.uniform(-1, 1, (L, ny, nz))
        ey = e[1][:, ry * ny:(ry + 1) * ny]
        ez = e[2][:, rz * nz:(rz + 1) * nz]
        for m in range(8):
            blk += 10.0 * e[0][m][:, None, None] * ey[m][None, :, None] * ez[m][None, None, :]
        x[r] = blk.ravel()
    return x, modes


def assemble(host_x):
    """The global 512^3 array from the per-rank pencil blocks."""
    ny, nz = L // GRID[1], L // GRID[2]
    glob = np.empty((L, L, L), np.complex128)
    for r in range(P):
        ry, rz = r // GRID[2], r % GRID[2]
        glob[:, ry * ny:(ry + 1) * ny, rz * nz:(rz + 1) * nz] = host_x[r].reshape(L, ny, nz)
    return glob


def host_threads():
    """Host threads for the CPU legs: the GPU's share of the box (16) at most."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def fp64_reference(glob, reps=2):
    """scipy.fft (pocketfft) in FP64 on the host threads: the independent
    oracle spectrum of the global array, and the CPU stand-in for the
    reference's MPI+FFTW path (FFTW is not installed), timed as forward +
    inverse per step.  Returns (spectrum, seconds per step, threads)."""
    import scipy.fft
    w = host_threads()
    X = None
    t0 = time.perf_counter()
    for _ in range(reps):
        X = scipy.fft.fftn(glob, workers=w)
        back = scipy.fft.ifftn(X, workers=w)
        del back
    dt = (time.perf_counter() - t0) / reps
    return X, dt, w


def spectrum_checks(spec, X, modes, tol):
    """Relative L2 against the FP64 oracle spectrum and the planted peaks
    (|X[k]| = 10 N at each planted wavenumber, within 1%)."""
    N = L ** 3
    err = float(np.linalg.norm(spec - X) / np.linalg.norm(X))
    peaks = [abs(abs(spec[tuple(k)]) / (10.0 * N) - 1) for k in modes]
    ok = err < tol and max(peaks) < 0.01
    return err, float(max(peaks)), ok


def setup(C, dm, comm, dbl, host_x, factors=None):
    dims = [C.makeLocalDecomp(L), C.makeBlockDecomp(L, GRID[1]), C.makeBlockDecomp(L, GRID[2])]
    start = C.makePackedDecomp(comm, 3, dm.ptrs(dims))
    per = C.extent(start)
    es = 16 if dbl else 8
    length = 2 * per + 1024
    dat = dm.DistBuffer(comm, per * es)
    work = dm.DistBuffer(comm, length * es)
    dat.put_all(host_x if dbl else host_x.astype(np.complex64))
    if factors is None:
        tune = C.autotuneParallelPlan if dbl else C.autotuneParallelPlanf
        plan = tune(start, dat.ptr, length, work.ptr)
    else:
        make = C.makeParallelPlan if dbl else C.makeParallelPlanf
        plan = make(start, dm.ints(factors), dat.ptr, length, work.ptr)
    if not plan:
        raise RuntimeError("autotune produced no plan")
    return start, dat, work, plan, per


def reshape_chain(C, dm, comm, reps=10):
    """The 512^3 int64 reshape chain on the same 4x8 grid: block ->
    rotated (seeded shifts, memory order z slowest, x, y fastest) ->
    interleave (stride 4 on axis 0).  Returns ms per hop and GB/s at the
    algorithmic 2 * N * 8 bytes per hop."""
    shifts = [int(v) for v in np.random.default_rng(11).integers(0, L, 3)]
    block = C.makeDecomp(3, dm.ptrs([C.makeAxisDecomp(comm, C.makeBlockDecomp(L, 4), 64 * L, 8),
                                     C.makeAxisDecomp(comm, C.makeBlockDecomp(L, 8), L, 1),
                                     C.makeAxisDecomp(comm, C.makeLocalDecomp(L), 1, 1)]))
    rot = []
    for sh, nested in zip(shifts, [C.makeBlockDecomp(L, 4), C.makeBlockDecomp(L, 8), C.makeLocalDecomp(L)]):
        rot.append(C.makeRotatedDecomp(sh, nested))
        C.freeDimDecomp(nested)
    rotated = C.makeDecomp(3, dm.ptrs([C.makeAxisDecomp(comm, rot[0], 64, 8), C.makeAxisDecomp(comm, rot[1], 1, 1),
                                       C.makeAxisDecomp(comm, rot[2], 128 * 64, 1)]))
    inner = C.makeBlockDecomp(4, 4)
    il = C.makeInterleaveDecomp(inner, 128)
    C.freeDimDecomp(inner)
    inter = C.makePackedDecomp(comm, 3, dm.ptrs([il, C.makeBlockDecomp(L, 8), C.makeLocalDecomp(L)]))
    t = dm.mpi_type("long_long")
    plans = [C.makeReshapePlan(comm, t, block, rotated), C.makeReshapePlan(comm, t, rotated, inter)]
    slots = [C.extent(d) for d in (block, rotated, inter)]
    bufs = [dm.DistBuffer(comm, s_ * 8) for s_ in slots]
    for b, s_ in zip(bufs, slots):
        b.fill(0, s_ * 8)
    C.runReshapePlan(comm, plans[0], bufs[0].ptr, bufs[1].ptr)
    C.runReshapePlan(comm, plans[1], bufs[1].ptr, bufs[2].ptr)
    hop_ms = []
    for k, (a, b) in enumerate(((0, 1), (1, 2))):
        C.dm_synchronize(comm)
        t0 = time.perf_counter()
        for _ in range(reps):
            C.runReshapePlan(comm, plans[k], bufs[a].ptr, bufs[b].ptr)
        C.dm_synchronize(comm)
        hop_ms.append((time.perf_counter() - t0) / reps * 1e3)
    for p_ in plans:
        C.freeReshapePlan(p_)
    for b in bufs:
        b.free()
    for d in (block, rotated, inter):
        C.freeDecomp(d)
    alg = 2 * L ** 3 * 8
    gbs = [alg / (ms * 1e-3) / 1e9 for ms in hop_ms]
    return {"workload": "512^3 int64 reshape chain block -> rotate(z,x,y) -> interleave(4) on 4x8 virtual ranks",
            "shifts": shifts, "ms_per_hop": [round(v, 4) for v in hop_ms],
            "GB_per_s": [round(v, 1) for v in gbs],
            "frac_of_peak": [round(v / PEAK_HBM_GBS, 4) for v in gbs]}


def pencil_bench(C, dm, steps=20, warmup=3, n=64, P=4):
    """The 64^3 complex-double pencil pipeline on a 2x2 grid (x-pencil ->
    FFT x -> y-pencil -> FFT y -> z-pencil -> FFT z, each stage one fused
    pass through dm_plan_axis_fft), beside a split plan of the same start:
    ms per forward 3D transform and relative L2 against numpy's FP64 DFT.
    The 4 MiB array stays in the memory-side cache, so no HBM fraction."""
    from decomp_mpi_amd.layout import flat_index
    comm = C.dm_world_create(P)

    def pencil(ax):
        blocks = iter((2, 2))
        return C.makePackedDecomp(comm, 3, dm.ptrs([C.makeLocalDecomp(n) if d == ax else
                                                    C.makeBlockDecomp(n, next(blocks)) for d in range(3)]))
    px, py, pz = pencil(0), pencil(1), pencil(2)
    per = max(C.extent(px), C.extent(py), C.extent(pz))
    rng = np.random.default_rng(5)
    glob = rng.uniform(-1, 1, (n, n, n)) + 1j * rng.uniform(-1, 1, (n, n, n))
    X = np.fft.fftn(glob)
    host = np.zeros(P * per, np.complex128)
    host[flat_index(px, per).ravel()] = glob.ravel()
    b0, b1 = dm.DistBuffer(comm, per * 16), dm.DistBuffer(comm, per * 16)
    fx = C.dm_plan_axis_fft(px, py, 0, 1, dm.FFTW_FORWARD)
    fy = C.dm_plan_axis_fft(py, pz, 1, 1, dm.FFTW_FORWARD)
    fz = C.dm_plan_axis_fft(pz, pz, 2, 1, dm.FFTW_FORWARD)

    def forward():
        C.dm_run_axis_fft(fx, b0.ptr, b1.ptr)
        C.dm_run_axis_fft(fy, b1.ptr, b0.ptr)
        C.dm_run_axis_fft(fz, b0.ptr, b1.ptr)
    b0.put_all(host.reshape(P, per))
    forward()
    got = b1.get_all(per * 16, np.complex128).ravel()[flat_index(pz, per)]
    err_p = float(np.linalg.norm(got - X) / np.linalg.norm(X))
    for _ in range(warmup):
        forward()
    C.dm_synchronize(comm)
    t0 = time.perf_counter()
    for _ in range(steps):
        forward()
    C.dm_synchronize(comm)
    ms_p = (time.perf_counter() - t0) / steps * 1e3
    # the split FFT of the same x-pencil start (forward only)
    length = 2 * per + 1024
    dat, work = dm.DistBuffer(comm, per * 16), dm.DistBuffer(comm, length * 16)
    dat.put_all(host.reshape(P, per))
    plan = C.makeSplitPlan(0, px, dat.ptr, length, work.ptr)
    C.runSplitPlan(plan, dm.FFTW_FORWARD)
    ext = dm.split_plan(plan).extent
    got = work.get_all(ext * 16, np.complex128).ravel()[flat_index(C.getTransformed(plan), ext)]
    err_s = float(np.linalg.norm(got - X) / np.linalg.norm(X))
    for _ in range(warmup):
        C.runSplitPlan(plan, dm.FFTW_FORWARD)
    C.dm_synchronize(comm)
    t0 = time.perf_counter()
    for _ in range(steps):
        C.runSplitPlan(plan, dm.FFTW_FORWARD)
    C.dm_synchronize(comm)
    ms_s = (time.perf_counter() - t0) / steps * 1e3
    C.freeSplitPlan(plan)
    for f in (fx, fy, fz):
        C.dm_free_axis_fft(f)
    for b in (b0, b1, dat, work):
        b.free()
    for d in (px, py, pz):
        C.freeDecomp(d)
    C.dm_world_free(comm)
    return {"workload": f"{n}^3 complex double, x -> y -> z pencils on a 2x2 grid, forward",
            "pencil_ms": round(ms_p, 4), "pencil_rel_l2_vs_fp64_dft": err_p,
            "split_plan_ms": round(ms_s, 4), "split_plan_rel_l2_vs_fp64_dft": err_s,
            "tolerance": 1e-12, "ok": bool(err_p < 1e-12 and err_s < 1e-12)}


def batched_bench(C, dm, steps, warmup, schedule="auto"):
    """256 independent complex-float 3D FFTs of seeded uneven shapes (each
    axis from {32, 48, 64, 96, 128}) in place, the batched launches replayed
    from a HIP graph; grid, tile and radix order from dm_batch_autotune, or
    fixed (schedule "full" or "planes": one workgroup per tile, large
    radices first) so that a profile sees one schedule throughout."""
    import ctypes
    from decomp_mpi_amd import workloads
    shapes = workloads.batched_shapes(256, seed=4)
    comm = C.dm_world_create(1)
    stream = C.dm_get_stream(comm)
    offs, at = [], 0
    for sh in shapes:
        offs.append(at)
        at += (int(np.prod(sh)) * 8 + 255) // 256 * 256
    buf = dm.DistBuffer(comm, at)
    rng = np.random.default_rng(6)
    xs = [workloads.uniform_complex(rng, sh, 0) for sh in shapes]
    plan = C.dm_batch_plan(len(shapes), 3, dm.ints([v for sh in shapes for v in sh]), 0, dm.FFTW_FORWARD)
    ptrs = (ctypes.c_void_p * len(shapes))(*[buf.ptr + o for o in offs])
    if not plan or C.dm_batch_bind(plan, ptrs) != 0:
        raise RuntimeError("batched plan failed")
    if schedule == "auto":
        tune_ms = C.dm_batch_autotune(plan, stream)
    else:
        tile_id = {"full": dm.DM_BATCH_TILE_FULL, "planes": dm.DM_BATCH_TILE_PLANES}[schedule]
        if C.dm_batch_set_schedule(plan, tile_id, 0) != 0:
            raise RuntimeError("batched schedule failed")
        tune_ms = -1.0
    tile, order = ctypes.c_int(), ctypes.c_int()
    C.dm_batch_info(plan, None, None, ctypes.byref(tile), ctypes.byref(order))
    grid_mode = C.dm_batch_grid(plan)

    def upload():
        for x, o in zip(xs, offs):
            buf.put(0, x, o)
    upload()
    C.dm_batch_graph_launch(plan, stream)
    C.dm_synchronize(comm)
    max_rel, max_abs = 0.0, 0.0
    for x, o, sh in zip(xs, offs, shapes):
        got = buf.get(0, x.nbytes, np.complex64, o).reshape(sh).astype(np.complex128)
        want = np.fft.fftn(x.astype(np.complex128))
        max_rel = max(max_rel, float(np.linalg.norm(got - want) / np.linalg.norm(want)))
        max_abs = max(max_abs, float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
    for _ in range(warmup):
        C.dm_batch_graph_launch(plan, stream)
    C.dm_synchronize(comm)
    t0 = time.perf_counter()
    for _ in range(steps):
        C.dm_batch_graph_launch(plan, stream)
    C.dm_synchronize(comm)
    dt = (time.perf_counter() - t0) / steps
    flops = sum(5 * np.prod(sh) * np.log2(np.prod(sh)) for sh in shapes)
    alg = C.dm_batch_bytes(plan)
    moved = C.dm_batch_bytes_moved(plan)
    C.dm_batch_free(plan)
    buf.free()
    C.dm_world_free(comm)
    return {"workload": "256 complex-float 3D FFTs, axes from {32,48,64,96,128} (seed 4), in place, "
                        "one HIP graph of the autotuned schedule's batched launches",
            "ms_per_batch": round(dt * 1e3, 4), "transforms_per_s": round(len(shapes) / dt, 1),
            "GFLOP_per_s": round(flops / dt / 1e9, 1), "GB_per_s": round(alg / dt / 1e9, 1),
            "frac_of_peak": round(alg / dt / 1e9 / PEAK_HBM_GBS, 4),
            "bytes_algorithmic": alg, "bytes_moved": moved,
            "moved_GB_per_s": round(moved / dt / 1e9, 1),
            "max_rel_l2_vs_fp64": max_rel, "max_abs_err_over_max_vs_fp64": max_abs,
            "schedule": {"tile": ["full", "half", "planes (last two axes per plane tile where they fit)"][tile.value],
                         "radix_order": ["large_first", "small_first"][order.value],
                         "grid": ["one workgroup per tile", "persistent"][grid_mode],
                         "autotune_best_ms": round(tune_ms, 4) if tune_ms > 0 else None,
                         "chosen_by": "dm_batch_autotune" if schedule == "auto" else "--batch-schedule " + schedule}}


def serving_bench(C, dm, count=1000, streams=4, max_batch=8, budget=64 << 30):
    """Poisson arrivals of mixed 3D FFT requests (axes log-uniform in
    [32, 256] rounded to 16; 70% complex float, 30% complex double; random
    direction) through dm_serve with host (pinned) input and output, at 60%
    of the closed-loop capacity measured first.  Latency = finish - arrival."""
    import ctypes
    from decomp_mpi_amd import workloads
    probe = workloads.serving_requests(count, 1.0, seed=5)   # shapes; arrivals rescaled below
    ring = 32
    sizes = [0] * ring
    for i, (_, sh, prec, _) in enumerate(probe):
        sizes[i % ring] = max(sizes[i % ring], int(np.prod(sh)) * (16 if prec else 8))
    bufs = [C.dm_host_alloc(sz) for sz in sizes]
    if not all(bufs):
        raise MemoryError("pinned host buffers")
    rng = np.random.default_rng(8)
    for b, sz in zip(bufs, sizes):   # inputs: uniform[-1,1) values
        v = rng.uniform(-1, 1, sz // 4).astype(np.float32)
        ctypes.memmove(b, v.ctypes.data, sz)

    def run(reqs, arrivals):
        s = C.dm_server_create(budget, streams, max_batch)
        owner = [-1] * ring
        ids = []
        t_start = time.perf_counter()
        for i, ((_, sh, prec, d), t_arr) in enumerate(zip(reqs, arrivals)):
            while time.perf_counter() - t_start < t_arr:
                C.dm_server_poll(s)
            j = i % ring
            if owner[j] >= 0:
                C.dm_server_wait(s, owner[j])
            rid = C.dm_server_submit(s, 3, dm.ints(list(sh)), prec, d, bufs[j], bufs[j])
            owner[j] = rid
            ids.append(rid)
            C.dm_server_poll(s)
        C.dm_server_drain(s)
        sub, fin = ctypes.c_double(), ctypes.c_double()
        times, parts = [], []
        ms = (ctypes.c_double * 6)()
        for rid in ids:
            C.dm_server_times(s, rid, ctypes.byref(sub), None, ctypes.byref(fin))
            times.append((sub.value, fin.value))
            C.dm_server_breakdown(s, rid, ms)
            parts.append(list(ms))
        nb, mx = ctypes.c_longlong(), ctypes.c_longlong()
        peak = ctypes.c_size_t()
        C.dm_server_stats(s, ctypes.byref(nb), ctypes.byref(mx), ctypes.byref(peak), None)
        cached, builds = ctypes.c_longlong(), ctypes.c_longlong()
        C.dm_server_plan_stats(s, ctypes.byref(cached), ctypes.byref(builds), None)
        C.dm_server_destroy(s)
        return times, nb.value, mx.value, peak.value, np.array(parts), builds.value

    # closed loop: everything at time 0
    n0 = min(256, count)
    times = run(probe[:n0], [0.0] * n0)[0]
    capacity = n0 / (max(f for _, f in times) - min(sub for sub, _ in times))
    lam = 0.6 * capacity
    reqs = workloads.serving_requests(count, lam, seed=5)
    arrivals = [a for a, _, _, _ in reqs]
    times, batches, max_seen, peak, parts, builds = run(reqs, arrivals)
    first = min(sub for sub, _ in times)
    lat = np.array([fin - (first + arr) for (sub, fin), arr in zip(times, arrivals)]) * 1e3
    # where the time went: mean breakdown of the requests at or above p99,
    # and of those around the median; "arrival_lag" is the submit time after
    # the scheduled arrival (the driver loop's own delay)
    lag = np.array([(sub - first - arr) * 1e3 for (sub, _), arr in zip(times, arrivals)])
    names = ["queue", "host_enqueue_overlapped", "copy_in", "kernels", "copy_out", "completion_lag"]

    def mean_parts(sel):
        out = {n: round(float(parts[sel, i].mean()), 3) for i, n in enumerate(names)}
        out["arrival_lag"] = round(float(lag[sel].mean()), 3)
        out["bytes_mean"] = int(np.mean([np.prod(reqs[j][1]) * (16 if reqs[j][2] else 8)
                                         for j in np.nonzero(sel)[0]]))
        return out
    p50, p99 = np.percentile(lat, 50), np.percentile(lat, 99)
    breakdown = {"at_or_above_p99": mean_parts(lat >= p99),
                 "around_p50": mean_parts((lat >= np.percentile(lat, 40)) & (lat <= np.percentile(lat, 60)))}
    span = max(f for _, f in times) - first
    # accuracy per precision class on a separate mixed set
    acc = {0: 0.0, 1: 0.0}
    s = C.dm_server_create(budget, streams, max_batch)
    checks = []
    for (_, sh, prec, d) in workloads.serving_requests(24, 1.0, seed=9):
        sh = tuple(min(a, 128) for a in sh)
        x = workloads.uniform_complex(rng, sh, prec)
        pin = C.dm_host_alloc(x.nbytes)
        ctypes.memmove(pin, x.ctypes.data, x.nbytes)
        checks.append((x, pin, d, C.dm_server_submit(s, 3, dm.ints(list(sh)), prec, d, pin, pin)))
        C.dm_server_poll(s)
    C.dm_server_drain(s)
    for x, pin, d, _ in checks:
        got = np.frombuffer((ctypes.c_char * x.nbytes).from_address(pin), dtype=x.dtype).reshape(x.shape)
        x64 = x.astype(np.complex128)
        want = np.fft.fftn(x64) if d == -1 else np.fft.ifftn(x64) * x64.size
        prec = 1 if x.dtype == np.complex128 else 0
        acc[prec] = max(acc[prec], float(np.linalg.norm(got.astype(np.complex128) - want) / np.linalg.norm(want)))
        C.dm_host_free(pin)
    C.dm_server_destroy(s)
    for b in bufs:
        C.dm_host_free(b)
    return {"workload": f"{count} Poisson requests, axes log-uniform [32,256] step 16, 70% c64 / 30% c128, "
                        f"random direction, pinned host in/out, budget {budget >> 30} GiB, {streams} streams, "
                        f"batches of up to {max_batch}",
            "closed_loop_capacity_req_per_s": round(capacity, 1), "lambda_req_per_s": round(lam, 1),
            "throughput_req_per_s": round(count / span, 1),
            "latency_ms": {"p50": round(float(p50), 3), "p99": round(float(p99), 3),
                           "mean": round(float(lat.mean()), 3)},
            "latency_breakdown_ms": breakdown,
            "batches": batches, "largest_batch": max_seen, "peak_device_bytes": peak,
            "plans_built": builds,
            "max_rel_l2_vs_fp64": {"c64": acc[0], "c128": acc[1]}}


def measured_traffic(bytes_per_launch):
    """HBM bytes per FFT pass launch from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what it
    reads for a plain copy of known size), when that summary exists."""
    path = os.path.join(ROOT, "profiles", "fft_pass_512cube_c128_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return {"bytes": int(t["hbm_bytes_per_launch"]),
            "ratio_to_algorithmic": round(t["hbm_bytes_per_launch"] / bytes_per_launch, 4),
            "source": "profiles/fft_pass_512cube_c128_traffic.json"}


TIMER_NAMES = ["RESHAPE_1", "FFT_1", "TWIDDLE", "RESHAPE_2", "FFT_2", "ALL"]


def time_steps(C, dm, comm, plan, steps, warmup):
    """Warm-up, then exactly `steps` forward+inverse steps between two
    device synchronisations.  Per step: the HIP-event time of its passes
    (plan->times[ALL] accumulates the per-pass event intervals, recorded on
    the stream the passes run on).  Returns (wall seconds, per-step event
    seconds, the plan's Timers per step)."""
    sp = dm.split_plan(plan)
    for _ in range(warmup):
        C.runSplitPlan(plan, dm.FFTW_FORWARD)
        C.runSplitPlan(plan, dm.FFTW_BACKWARD)
    for i in range(dm.NUM_TIMERS):
        sp.times[i] = 0.0
    per_step = []
    C.dm_synchronize(comm)
    t0 = time.perf_counter()
    for _ in range(steps):
        before = sp.times[dm.ALL]
        C.runSplitPlan(plan, dm.FFTW_FORWARD)
        C.runSplitPlan(plan, dm.FFTW_BACKWARD)
        per_step.append(sp.times[dm.ALL] - before)
    C.dm_synchronize(comm)
    wall = time.perf_counter() - t0
    timers = {name: round(sp.times[i] / steps * 1e3, 4) for i, name in enumerate(TIMER_NAMES)}
    return wall, np.array(per_step), timers


def step_stats(per_step):
    ms = per_step * 1e3
    return {"median": round(float(np.median(ms)), 4), "min": round(float(ms.min()), 4),
            "mean": round(float(ms.mean()), 4), "n": int(ms.size)}


def cpu_reshape_baseline():
    """The oracle's reshape (the reference's message pairing restated in C,
    one host core) on the 512^3 int64 chain: GB/s per hop at 2 * N * 8
    bytes."""
    import oracle
    shifts = [int(v) for v in np.random.default_rng(11).integers(0, L, 3)]
    block = oracle.Decomp(P, [(oracle.Block(L, 4), 8), (oracle.Block(L, 8), 1), (oracle.Local(L), 1)],
                          strides=[(L // 8) * L, L, 1])
    rot = oracle.Decomp(P, [(oracle.Rotate(shifts[0], oracle.Block(L, 4)), 8),
                            (oracle.Rotate(shifts[1], oracle.Block(L, 8)), 1),
                            (oracle.Rotate(shifts[2], oracle.Local(L)), 1)],
                        strides=[L // 8, 1, (L // 4) * (L // 8)])
    inter = oracle.Decomp(P, [(oracle.Interleave(oracle.Block(4, 4), L // 4), 8), (oracle.Block(L, 8), 1),
                              (oracle.Local(L), 1)])
    ext = [int(max(d.extents)) for d in (block, rot, inter)]
    bufs = [np.arange(P * ext[0], dtype=np.int64).reshape(P, ext[0])]
    bufs += [np.zeros((P, e), np.int64) for e in ext[1:]]
    secs = []
    for k, (a, b) in enumerate(((block, rot), (rot, inter))):
        t0 = time.perf_counter()
        oracle.reshape(a, b, bufs[k].view(np.uint8), bufs[k + 1].view(np.uint8), 8)
        secs.append(time.perf_counter() - t0)
    alg = 2 * L ** 3 * 8
    return {"GB_per_s": [round(alg / t / 1e9, 4) for t in secs], "s_per_hop": [round(t, 3) for t in secs],
            "cores": 1, "kind": "port",
            "sample": "oracle reshape (the reference's per-message pairing restated in C), both hops of the "
                      "512^3 int64 chain, one host core"}


def cpu_split_port_baseline(n=512):
    """The oracle's split FFT (the reference's two stages restated in C,
    radix-2 stage DFTs, one core), one forward transform of the 512^3 array
    on the same 4x8 grid with the reference's default split."""
    import oracle
    start = oracle.Decomp(P, [(oracle.Local(n), 32), (oracle.Block(n, 4), 8), (oracle.Block(n, 8), 1)])
    a = max(q for q in range(1, n) if n % q == 0 and q * q <= n)
    A, B = [a] * 3, [n // a] * 3
    ga, gb = oracle.processor_grid(P, A), oracle.processor_grid(P, B)
    ext = oracle.split_extent(P, A, B, ga, gb)
    per = int(max(start.extents))
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (P, per)) + 1j * rng.uniform(-1, 1, (P, per))
    w = np.zeros((P, 2 * ext), np.complex128)
    t0 = time.perf_counter()
    oracle.split_fft(start, A, B, ga, gb, x, w, ext, -1)
    dt = time.perf_counter() - t0
    return {"GB_per_s": round(6 * n ** 3 * 16 / dt / 1e9, 4), "cores": 1, "kind": "port",
            "sample": f"oracle split FFT, one forward transform, {n}^3 complex double, 4x8 virtual grid, "
                      f"A={A}, {dt:.2f} s on one host core"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the FP64 reference spectrum and the error checks (profiling runs)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-phases", action="store_true", help="skip the phase-order timing")
    ap.add_argument("--no-reshape", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batch-schedule", choices=["auto", "full", "planes"], default="auto",
                    help="batched section: autotuned schedule (default) or a fixed one")
    ap.add_argument("--factors", type=str, default=None,
                    help="B factors b0,b1,b2 for makeParallelPlan instead of the autotuner (profiling runs)")
    ap.add_argument("--only", choices=["batched", "reshape", "serving"], default=None,
                    help="run one section alone and print its figures (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU; the virtual-rank grid stands in for "
                                   "the processor grid (multi-GPU is a design, not a run)"}))
        sys.exit(2)
    import decomp_mpi_amd as dm
    from decomp_mpi_amd.layout import flat_index
    C = dm.lib()
    if not C.dm_device_available():
        raise SystemExit("no HIP device")
    if args.only is not None:
        if args.only == "batched":
            res = batched_bench(C, dm, max(args.steps, 5), max(args.warmup, 1), args.batch_schedule)
            # a plain 1 GiB copy in the same run: the byte counters' calibration
            comm = C.dm_world_create(1)
            cb = dm.DistBuffer(comm, 2 << 30)
            res["copy_1GiB_ms"] = round(C.dm_copy_bench(comm, cb.ptr + (1 << 30), cb.ptr, 1 << 30, 3), 4)
            cb.free()
        elif args.only == "serving":
            res = serving_bench(C, dm)
        else:
            comm = C.dm_world_create(P)
            res = reshape_chain(C, dm, comm)
            cb = dm.DistBuffer(comm, (2 << 30) // P)
            res["copy_1GiB_ms"] = round(C.dm_copy_bench(comm, cb.ptr + (1 << 30), cb.ptr, 1 << 30, 3), 4)
            cb.free()
        print(json.dumps({"section": args.only, **res}), flush=True)
        return
    print(f"bench: generating {L}^3 input", file=sys.stderr, flush=True)
    host_x, modes = make_input()
    N = L ** 3
    alg_step = 2 * 6 * N * 16
    X = None
    cpu_fft = None
    if not args.no_check:
        print("bench: FP64 reference spectrum (scipy.fft on the host)", file=sys.stderr, flush=True)
        glob = assemble(host_x)
        X, cpu_step_s, threads = fp64_reference(glob)
        del glob
        cpu_fft = {"value": round(alg_step / cpu_step_s / 1e9, 4), "unit": "GB/s", "cores": threads,
                   "kind": "port",
                   "sample": f"scipy.fft (pocketfft) FP64 fftn + ifftn of the same 512^3 array on {threads} host "
                             f"threads, 2 steps, {cpu_step_s:.2f} s per step; a library stand-in for the "
                             "reference's MPI+FFTW path (FFTW is not installed)",
                   "backend": "scipy.fft / pocketfft (FP64)", "cpu_model": cpu_model(),
                   "ms_per_step": round(cpu_step_s * 1e3, 1)}
    comm = C.dm_world_create(P)
    factors = [int(v) for v in args.factors.split(",")] if args.factors else None
    checks_ok = True

    def verify(plan, dat, work, per, dbl):
        """Forward against the FP64 spectrum and the planted peaks, then
        the round trip over every rank; leaves the input in dat."""
        ct, es, tol = (np.complex128, 16, 1e-12) if dbl else (np.complex64, 8, 1e-6)
        sp_ = dm.split_plan(plan)
        C.runSplitPlan(plan, dm.FFTW_FORWARD)
        spec = work.get_all(sp_.extent * es, ct).ravel()[flat_index(C.getTransformed(plan), sp_.extent)]
        err, peak, ok = spectrum_checks(spec.reshape(L, L, L).astype(np.complex128), X, modes, tol)
        del spec
        C.runSplitPlan(plan, dm.FFTW_BACKWARD)
        back = dat.get_all(per * es, ct).astype(np.complex128) / N
        rt = float(np.linalg.norm(back - host_x) / np.linalg.norm(host_x))
        del back
        dat.put_all(host_x if dbl else host_x.astype(np.complex64))
        return {"rel_l2_vs_fp64_dft": err, "max_planted_peak_rel_dev": peak, "round_trip_rel_l2_all_ranks": rt,
                "tolerance": tol, "ok": bool(ok and rt < tol)}

    print("bench: planning FP64", file=sys.stderr, flush=True)
    start, dat, work, plan, per = setup(C, dm, comm, True, host_x, factors)
    sp = dm.split_plan(plan)
    A64 = [sp.A_factors[i] for i in range(3)]
    chk64 = None
    if X is not None:
        chk64 = verify(plan, dat, work, per, True)
        checks_ok = checks_ok and chk64["ok"]
    print("bench: timing FP64", file=sys.stderr, flush=True)
    wall, per_step, timers64 = time_steps(C, dm, comm, plan, args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    passes = 6 * args.steps
    pass_ms = float(per_step.sum()) / passes * 1e3
    pass_bytes = 2 * N * 16
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9
    C.freeSplitPlan(plan)
    dat.free()
    work.free()

    # the reference's default split (A = 16, B = 32 on every axis) in the
    # reference's phase order, each phase timed, beside the fused program
    # for the same split
    phases = None
    if not args.no_phases:
        print("bench: phase order, default split", file=sys.stderr, flush=True)
        _, datp, workp, planp, _ = setup(C, dm, comm, True, host_x, [32, 32, 32])
        wall_f, per_f, _ = time_steps(C, dm, comm, planp, 5, 2)
        C.dm_split_set_phase_order(planp, 1)
        datp.put_all(host_x)   # the timed round trips scaled it by N per step
        chkp = verify(planp, datp, workp, per, True) if X is not None else None
        checks_ok = checks_ok and (chkp is None or chkp["ok"])
        wall_p, per_p, timers_p = time_steps(C, dm, comm, planp, 5, 2)
        phases = {"workload": "the same 512^3 transform with the reference's default split (A = 16, B = 32 per "
                              "axis), forward+inverse",
                  "phase_order_ms_per_step": round(wall_p / 5 * 1e3, 4),
                  "timers_ms_per_step": timers_p,
                  "fused_ms_per_step_same_split": round(wall_f / 5 * 1e3, 4),
                  "checks": chkp}
        C.freeSplitPlan(planp)
        datp.free()
        workp.free()

    fp32 = None
    if not args.no_fp32:
        print("bench: planning + timing FP32", file=sys.stderr, flush=True)
        start32, dat32, work32, plan32, _ = setup(C, dm, comm, False, host_x, factors)
        sp32 = dm.split_plan(plan32)
        A32 = [sp32.A_factors[i] for i in range(3)]
        chk32 = None
        if X is not None:
            chk32 = verify(plan32, dat32, work32, per, False)
            checks_ok = checks_ok and chk32["ok"]
        wall32, per_step32, timers32 = time_steps(C, dm, comm, plan32, args.steps, args.warmup)
        # the same plan with plain float twiddles, beside the compensated
        # products of the timed form
        assert C.dm_split_set_twiddles(plan32, dm.DM_TWIDDLES_PLAIN) == 0
        dat32.put_all(host_x.astype(np.complex64))
        chk_plain = verify(plan32, dat32, work32, per, False) if X is not None else None
        wall_pl, _, _ = time_steps(C, dm, comm, plan32, args.steps, args.warmup)
        fp32 = {"ms_per_step": round(wall32 / args.steps * 1e3, 4),
                "GB_per_s": round(alg_step / (wall32 / args.steps) / 1e9, 2),
                "speedup_vs_fp64": round(wall / wall32, 3),
                "step_ms_events": step_stats(per_step32), "timers_ms_per_step": timers32,
                "checks": chk32, "A_factors": A32,
                "pass_ms": round(float(per_step32.sum()) / passes * 1e3, 4),
                "twiddles": "compensated (float + bfloat16 correction, FP64-generated)",
                "plain_float_twiddles": {"ms_per_step": round(wall_pl / args.steps * 1e3, 4),
                                         "rel_l2_vs_fp64_dft": chk_plain and chk_plain["rel_l2_vs_fp64_dft"],
                                         "round_trip_rel_l2_all_ranks":
                                             chk_plain and chk_plain["round_trip_rel_l2_all_ranks"]}}
        C.freeSplitPlan(plan32)
        dat32.free()
        work32.free()
    X = None

    # plain 2 GiB device copy: the achievable-bandwidth reference
    cbuf = dm.DistBuffer(comm, per * 16 * 2)
    half = P * per * 16                       # 2 GiB: the array size
    copy_ms = C.dm_copy_bench(comm, cbuf.ptr + half, cbuf.ptr, half, 5)
    copy_gbs = 2 * half / (copy_ms * 1e-3) / 1e9
    cbuf.free()

    chain = None
    if not args.no_reshape:
        print("bench: timing the 512^3 int64 reshape chain", file=sys.stderr, flush=True)
        chain = reshape_chain(C, dm, comm)

    print("bench: 64^3 pencil pipeline", file=sys.stderr, flush=True)
    pencil = pencil_bench(C, dm)
    checks_ok = checks_ok and pencil["ok"]

    batched = None
    if not args.no_batched:
        print("bench: batched 256 uneven FFTs", file=sys.stderr, flush=True)
        batched = batched_bench(C, dm, max(args.steps, 5), max(args.warmup, 1), args.batch_schedule)
    serving = None
    if not args.no_serving:
        print("bench: serving loop", file=sys.stderr, flush=True)
        serving = serving_bench(C, dm)

    cpu = cpu_fft
    if cpu is not None and not args.no_cpu:
        print("bench: CPU ports (oracle reshape chain, oracle split FFT sample)", file=sys.stderr, flush=True)
        cpu = dict(cpu)
        cpu["reshape_chain_port"] = cpu_reshape_baseline()
        cpu["split_fft_port"] = cpu_split_port_baseline()
    value = alg_step / (wall / args.steps) / 1e9
    out = {
        "metric": "512^3 c2c 3D FFT time & HBM GB/s, FP32 vs FP64 form; batched req/s + p99",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (f64)",
        "data": "synthetic: uniform[-1,1) re/im + 8 planted modes, seed 7",
        "config": {"workload": "split FFT forward+inverse, 512^3 complex double, pencil start "
                               "{Local(512), Block(512,4) rs8, Block(512,8) rs1} on 32 virtual ranks",
                   "A_factors": A64, "gflops": round(2 * 5 * N * np.log2(N) / (wall / args.steps) / 1e9, 1),
                   "step_ms_events": step_stats(per_step), "timers_ms_per_step": timers64,
                   "checks": chk64, "phase_order": phases, "fp32_form": fp32, "reshape_chain": chain,
                   "pencil_64cube": pencil, "batched": batched, "serving": serving},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": (measured_traffic(pass_bytes) or {}).get("bytes"),
                     "traffic_detail": measured_traffic(pass_bytes),
                     "kernel": "fft_pass_kernel<double> (one axis pass)",
                     "pass_ms": round(pass_ms, 4), "bytes_per_launch": pass_bytes,
                     "copy_2GiB_GBps": round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if not checks_ok:
        print("bench: a result check failed (see config.checks)", file=sys.stderr, flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
