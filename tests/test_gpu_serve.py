"""The request loop (dm_serve.h) on the device: every result against
numpy's FP64 DFT (complex float within 1e-6, complex double within 1e-12
relative L2), each result the same (bit for bit) alone and in a mix of
other requests arriving while it runs, and the HBM budget respected."""
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


def make_requests(C, n, seed, max_axis=96):
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
    s = C.dm_server_create(4 << 30, 4, 8)
    reqs = make_requests(C, 40, seed=21)
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


def test_server_alone_equals_mix(C):
    probe = make_requests(C, 5, seed=31)
    alone = []
    for r in probe:
        s = C.dm_server_create(4 << 30, 1, 8)
        rid = submit(C, s, r)
        assert C.dm_server_wait(s, rid) == 0
        alone.append(r[5].array())
        C.dm_server_destroy(s)
    for r in probe:  # clear the outputs
        ctypes.memset(r[5].ptr, 0, r[3].nbytes)
    others = make_requests(C, 30, seed=32)
    s = C.dm_server_create(4 << 30, 3, 6)
    ids = []
    k = 0
    for i, r in enumerate(others):  # probes arrive while others are in flight
        submit(C, s, r)
        C.dm_server_poll(s)
        if i % 6 == 3 and k < len(probe):
            ids.append(submit(C, s, probe[k]))
            k += 1
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
