"""dm_batch_run_noalloc (dm_batch.h): the batched local FFTs in one call that
allocates no device memory, against the bound entry point dm_batch_run on
the same inputs (bit for bit: same kernels, same tables) and against
numpy's FP64 DFT (1e-12 / 1e-6 relative L2 per transform).  Two calls in a
row on one stream, the second with a smaller batch, so the second run
reuses the table area the first one filled."""
import ctypes

import numpy as np
import pytest

import decomp_mpi_amd as dm

pytestmark = pytest.mark.gpu

TOL = {1: 1e-12, 0: 1e-6}


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


def carve(shapes, es):
    offs, at = [], 0
    for s in shapes:
        offs.append(at)
        at += (int(np.prod(s)) * es + 255) // 256 * 256
    return offs, at


def run_pair(C, comm, stream, shapes, dbl, direction, seed):
    """The same inputs through dm_batch_run (buffer a) and
    dm_batch_run_noalloc (buffer b); returns both outputs and the inputs."""
    es = 16 if dbl else 8
    ct = np.complex128 if dbl else np.complex64
    offs, total = carve(shapes, es)
    a, b = dm.DistBuffer(comm, total), dm.DistBuffer(comm, total)
    rng = np.random.default_rng(seed)
    xs = [(rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s)).astype(ct) for s in shapes]
    for x, o in zip(xs, offs):
        a.put(0, x, o)
        b.put(0, x, o)
    flat = dm.ints([v for s in shapes for v in s])
    nd = len(shapes[0])
    plan = C.dm_batch_plan(len(shapes), nd, flat, dbl, direction)
    assert plan
    pa = (ctypes.c_void_p * len(shapes))(*[a.ptr + o for o in offs])
    assert C.dm_batch_bind(plan, pa) == 0
    assert C.dm_batch_run(plan, stream) == 0
    pb = (ctypes.c_void_p * len(shapes))(*[b.ptr + o for o in offs])
    print(f"noalloc: {len(shapes)} transforms, precision {dbl}, direction {direction}", flush=True)
    assert C.dm_batch_run_noalloc(len(shapes), nd, flat, dbl, direction, pb, stream) == 0
    C.dm_synchronize(comm)
    ga = [a.get(0, int(np.prod(s)) * es, ct, o).reshape(s) for s, o in zip(shapes, offs)]
    gb = [b.get(0, int(np.prod(s)) * es, ct, o).reshape(s) for s, o in zip(shapes, offs)]
    C.dm_batch_free(plan)
    a.free()
    b.free()
    return ga, gb, xs


@pytest.mark.parametrize("dbl", [0, 1])
def test_noalloc_matches_bound_run_twice(C, dbl):
    comm = C.dm_world_create(1)
    stream = C.dm_get_stream(comm)
    first = [(64, 96, 128), (32, 48, 64), (96, 96, 32), (80, 112, 48), (176, 32, 16), (48, 64, 128)] * 3
    second = [(32, 32, 32), (16, 48, 40), (64, 16, 16)]
    for k, (shapes, direction) in enumerate(((first, -1), (second, 1))):
        ga, gb, xs = run_pair(C, comm, stream, shapes, dbl, direction, seed=10 + k)
        for x, u, v in zip(xs, ga, gb):
            assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), "noalloc differs from dm_batch_run"
            x64 = x.astype(np.complex128)
            want = np.fft.fftn(x64) if direction == -1 else np.fft.ifftn(x64) * x64.size
            err = np.linalg.norm(v.astype(np.complex128) - want) / np.linalg.norm(want)
            assert err < TOL[dbl], err
    C.dm_world_free(comm)


def test_noalloc_refuses_batch_beyond_scratch(C):
    """A batch whose tables exceed the fixed area is refused with an error,
    before anything is copied or launched."""
    comm = C.dm_world_create(1)
    buf = dm.DistBuffer(comm, 256)
    n = 600000   # 600k one-point transforms: 4.8 MB of pointers + 19 MB of tiles
    ptrs = (ctypes.c_void_p * n)(*([buf.ptr] * n))
    C.dm_set_error_mode(1)
    try:
        C.dm_clear_error()
        rc = C.dm_batch_run_noalloc(n, 1, dm.ints([1] * n), 0, -1, ptrs, C.dm_get_stream(comm))
        assert rc == 1
        assert b"scratch array holds" in C.dm_last_error()
    finally:
        C.dm_clear_error()
        C.dm_set_error_mode(0)
    buf.free()
    C.dm_world_free(comm)
