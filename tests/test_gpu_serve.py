"""The request loop (dm_serve.h) on the device: every result against
numpy's FP64 DFT (complex float within 1e-6, complex double within 1e-12
relative L2) over the full axis range of the serving workload ([32, 256]
in steps of 16), each result the same (bit for bit) alone and in a mix of
other requests arriving while it runs, the HBM budget respected, plans
built once per shape and reused, and the latency breakdown accounting for
each request's time."""
import ctypes

import numpy as np
import pytest

import decomp_mpi_amd as dm
from decomp_mpi_amd import workloads

pytestmark = pytest.mark.gpu

TOL = {1: 1e-12, 0: 1e-6}


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


class Pinned:
    def __init__(self, C, arr):
        self.C = C
        self.shape, self.dtype = arr.shape, arr.dtype
        self.ptr = C.dm_host_alloc(arr.nbytes)
        assert self.ptr
        ctypes.memmove(self.ptr, arr.ctypes.data, arr.nbytes)

    def array(self):
        n = int(np.prod(self.shape)) * np.dtype(self.dtype).itemsize
        buf = (ctypes.c_char * n).from_address(self.ptr)
        return np.frombuffer(buf, dtype=self.dtype).reshape(self.shape).copy()

    def free(self):
        self.C.dm_host_free(self.ptr)


def make_requests(C, n, seed, max_axis=256):
    rng = np.random.default_rng(seed)
    reqs = []
    for (_, shape, prec, direction) in workloads.serving_requests(n, 1000.0, seed=seed):
        shape = tuple(min(a, max_axis) for a in shape)
        x = workloads.uniform_complex(rng, shape, prec)
        reqs.append((shape, prec, direction, x, Pinned(C, x), Pinned(C, np.zeros_like(x))))
    return reqs


def submit(C, s, req):
    shape, prec, direction, _, pin_in, pin_out = req
    rid = C.dm_server_submit(s, len(shape), dm.ints(list(shape)), prec, direction, pin_in.ptr, pin_out.ptr)
    assert rid >= 0
    return rid


def want(req):
    shape, prec, direction, x, _, _ = req
    x64 = x.astype(np.complex128)
    return np.fft.fftn(x64) if direction == -1 else np.fft.ifftn(x64) * x64.size


def release(reqs):
    for r in reqs:
        r[4].free()
        r[5].free()


def test_server_results_match_numpy(C):
    s = C.dm_server_create(8 << 30, 4, 8)
    reqs = make_requests(C, 32, seed=21)
    assert max(max(r[0]) for r in reqs) >= 192, "the draw should reach the top of the axis range"
    for r in reqs:
        submit(C, s, r)
        C.dm_server_poll(s)
    assert C.dm_server_drain(s) == 0
    for r in reqs:
        got = r[5].array().astype(np.complex128)
        w = want(r)
        err = np.linalg.norm(got - w) / np.linalg.norm(w)
        assert err < TOL[r[1]], (r[0], r[1], r[2], err)
    nb, mx = ctypes.c_longlong(), ctypes.c_longlong()
    C.dm_server_stats(s, ctypes.byref(nb), ctypes.byref(mx), None, None)
    assert nb.value >= 1 and mx.value >= 1
    C.dm_server_destroy(s)
    release(reqs)


@pytest.mark.parametrize("prec", [0, 1])
def test_server_lines_longer_than_lds_twiddles(C, prec):
    """Requests with lines of more than 256 points (their batches read the
    twiddles from device memory) mixed with short-line ones (twiddles in
    LDS when batched alone), both directions, against numpy."""
    rng = np.random.default_rng(33)
    s = C.dm_server_create(8 << 30, 2, 8)
    reqs = []
    for shape, direction in (((512, 16, 16), -1), ((32, 48, 64), -1), ((16, 1024, 8), 1), ((64, 32, 32), 1)):
        x = workloads.uniform_complex(rng, shape, prec)
        reqs.append((shape, prec, direction, x, Pinned(C, x), Pinned(C, np.zeros_like(x))))
    print(f"serve long lines: precision {prec}", flush=True)
    for r in reqs:
        submit(C, s, r)
    assert C.dm_server_drain(s) == 0
    for r in reqs:
        got = r[5].array().astype(np.complex128)
        w = want(r)
        err = np.linalg.norm(got - w) / np.linalg.norm(w)
        assert err < TOL[r[1]], (r[0], r[2], err)
    C.dm_server_destroy(s)
    release(reqs)


def test_server_alone_equals_mix(C):
    probe = make_requests(C, 5, seed=31)
    alone = []
    for r in probe:
        s = C.dm_server_create(8 << 30, 1, 8)
        rid = submit(C, s, r)
        assert C.dm_server_wait(s, rid) == 0
        alone.append(r[5].array())
        C.dm_server_destroy(s)
    for r in probe:  # clear the outputs
        ctypes.memset(r[5].ptr, 0, r[3].nbytes)
    others = make_requests(C, 24, seed=32)
    s = C.dm_server_create(8 << 30, 3, 6)
    ids = []
    k = 0
    for i, r in enumerate(others):  # probes arrive while others are in flight
        submit(C, s, r)
        C.dm_server_poll(s)
        if i % 5 == 3 and k < len(probe):
            ids.append(submit(C, s, probe[k]))
            k += 1
    assert k == len(probe)
    assert C.dm_server_drain(s) == 0
    for r, a in zip(probe, alone):
        assert np.array_equal(r[5].array(), a)
        err = np.linalg.norm(r[5].array().astype(np.complex128) - want(r)) / np.linalg.norm(want(r))
        assert err < TOL[r[1]]
    C.dm_server_destroy(s)
    release(probe + others)


def test_server_budget(C):
    budget = 16 << 20
    s = C.dm_server_create(budget, 4, 4)
    rng = np.random.default_rng(41)
    reqs = []
    for i in range(24):
        shape = (64, 64, 64)  # 2 MiB complex float, 4 MiB complex double
        prec = i % 2
        x = workloads.uniform_complex(rng, shape, prec)
        reqs.append((shape, prec, -1, x, Pinned(C, x), Pinned(C, np.zeros_like(x))))
        submit(C, s, reqs[-1])
    assert C.dm_server_drain(s) == 0
    peak, held = ctypes.c_size_t(), ctypes.c_size_t()
    C.dm_server_stats(s, None, None, ctypes.byref(peak), ctypes.byref(held))
    assert 0 < peak.value <= budget and held.value <= budget
    for r in reqs:
        got = r[5].array().astype(np.complex128)
        assert np.linalg.norm(got - want(r)) / np.linalg.norm(want(r)) < TOL[r[1]]
    C.dm_server_destroy(s)
    release(reqs)


def test_server_plans_stay_on_the_device(C):
    """Plans are device-resident: 12 requests over 3 shapes build the
    schedules their axes need (one per line length and tile height), and
    12 more requests of the same shapes build none; every finished request
    has a latency breakdown whose parts are non-negative and add up to its
    latency (queue, copies, kernels, completion lag; the host enqueue time
    overlaps the device parts)."""
    s = C.dm_server_create(1 << 30, 2, 4)
    rng = np.random.default_rng(51)
    shapes = [(48, 64, 32), (256, 32, 32), (32, 32, 160)]
    cached, builds = ctypes.c_longlong(), ctypes.c_longlong()
    nbytes = ctypes.c_size_t()
    ms = (ctypes.c_double * 6)()
    sub, fin = ctypes.c_double(), ctypes.c_double()
    seen = []
    for rnd in range(2):
        reqs, ids = [], []
        for i in range(12):
            shape = shapes[i % 3]
            x = workloads.uniform_complex(rng, shape, 0)
            reqs.append((shape, 0, -1, x, Pinned(C, x), Pinned(C, np.zeros_like(x))))
        for r in reqs:
            ids.append(submit(C, s, r))
            C.dm_server_poll(s)
        assert C.dm_server_drain(s) == 0
        assert C.dm_server_plan_stats(s, ctypes.byref(cached), ctypes.byref(builds), ctypes.byref(nbytes)) == 0
        seen.append((builds.value, cached.value))
        assert 3 <= builds.value <= 9 and cached.value == builds.value and nbytes.value > 0
        for rid, r in zip(ids, reqs):
            assert C.dm_server_breakdown(s, rid, ms) == 0
            parts = list(ms)
            assert min(parts[:5]) >= 0 and parts[5] > -0.02, parts
            C.dm_server_times(s, rid, ctypes.byref(sub), None, ctypes.byref(fin))
            # ms[1] (host enqueue) overlaps the device parts
            assert abs(sum(parts) - parts[1] - (fin.value - sub.value) * 1e3) < 1e-6
            assert parts[3] > 0
            got = r[5].array().astype(np.complex128)
            assert np.linalg.norm(got - want(r)) / np.linalg.norm(want(r)) < TOL[0]
        release(reqs)
    assert seen[1] == seen[0], seen
    C.dm_server_destroy(s)


def test_server_refuses_without_abort(C):
    """A request the server cannot take (an axis longer than one tile of
    the batched kernel) returns -1 with the reason in dm_last_error, in the
    default error mode, and the server keeps serving."""
    s = C.dm_server_create(1 << 30, 1, 2)
    x = np.zeros((8192, 2, 2), np.complex64)
    pin = Pinned(C, x)
    C.dm_clear_error()
    assert C.dm_server_submit(s, 3, dm.ints([8192, 2, 2]), 0, -1, pin.ptr, pin.ptr) == -1
    assert b"not supported" in C.dm_last_error()
    C.dm_clear_error()
    y = workloads.uniform_complex(np.random.default_rng(3), (32, 32, 32), 1)
    req = ((32, 32, 32), 1, -1, y, Pinned(C, y), Pinned(C, np.zeros_like(y)))
    assert C.dm_server_wait(s, submit(C, s, req)) == 0
    assert np.linalg.norm(req[5].array() - want(req)) / np.linalg.norm(want(req)) < TOL[1]
    C.dm_server_destroy(s)
    pin.free()
    release([req])


def test_server_prime_axes(C):
    """Axes with primes above 13 (direct-sum stages) in both precisions and
    directions, mixed with ordinary ones in the same batches."""
    s = C.dm_server_create(1 << 30, 2, 4)
    rng = np.random.default_rng(31)
    reqs = []
    for shape, prec, direction in (((34, 32, 17), 0, -1), ((19, 16, 32), 1, 1), ((32, 32, 32), 0, 1),
                                   ((23, 46, 16), 1, -1), ((48, 17, 64), 0, -1)):
        x = workloads.uniform_complex(rng, shape, prec)
        reqs.append((shape, prec, direction, x, Pinned(C, x), Pinned(C, np.zeros_like(x))))
    ids = [submit(C, s, r) for r in reqs]
    for rid in ids:
        assert C.dm_server_wait(s, rid) == 0
    for r in reqs:
        err = np.linalg.norm(r[5].array() - want(r)) / np.linalg.norm(want(r))
        assert err < TOL[r[1]], (r[0], r[1], err)
    C.dm_server_destroy(s)
    release(reqs)
