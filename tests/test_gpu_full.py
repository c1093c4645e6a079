"""Full-size cases on the device, through the C ABI.

* The 512^3 int64 reshape chain: block decomposition on a 4x8 grid ->
  rotated decomposition with axis permutation (z slowest, x, y fastest in
  memory) -> interleave decomposition with stride 4 on axis 0.  Every
  element holds its own global linear index, so the check is exact and
  size-independent: each destination slot holds the index its descriptor
  prescribes, and every slot no coordinate maps to keeps its fill value.
* The 512^3 complex-double split FFT (pencil start on 4x8) with the
  default split (A=16, B=32), the mixed split B=(1,1,256) and the
  autotuned plan, forward against scipy's FP64 DFT (relative L2 < 1e-12),
  the planted modes, and the unnormalised round trip on every rank; the
  FP32 form within 1e-6 of the same FP64 spectrum.
"""
import numpy as np
import pytest

import decomp_mpi_amd as dm
from helpers import flat_index, rel_l2

pytestmark = pytest.mark.gpu

L = 512
P = 32
N = L ** 3


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


def reshape_chain_layouts(C, comm, shifts):
    """(block, rotated + permuted, interleaved) decompositions of 512^3."""
    block = C.makeDecomp(3, dm.ptrs([C.makeAxisDecomp(comm, C.makeBlockDecomp(L, 4), 64 * L, 8),
                                     C.makeAxisDecomp(comm, C.makeBlockDecomp(L, 8), L, 1),
                                     C.makeAxisDecomp(comm, C.makeLocalDecomp(L), 1, 1)]))
    rot = []
    for s, nested in zip(shifts, [C.makeBlockDecomp(L, 4), C.makeBlockDecomp(L, 8), C.makeLocalDecomp(L)]):
        rot.append(C.makeRotatedDecomp(int(s), nested))
        C.freeDimDecomp(nested)
    # memory order: axis 2 (z) slowest, axis 0 (x), axis 1 (y) fastest
    rotated = C.makeDecomp(3, dm.ptrs([C.makeAxisDecomp(comm, rot[0], 64, 8),
                                       C.makeAxisDecomp(comm, rot[1], 1, 1),
                                       C.makeAxisDecomp(comm, rot[2], 128 * 64, 1)]))
    inner = C.makeBlockDecomp(4, 4)
    il = C.makeInterleaveDecomp(inner, 128)
    C.freeDimDecomp(inner)
    inter = C.makePackedDecomp(comm, 3, dm.ptrs([il, C.makeBlockDecomp(L, 8), C.makeLocalDecomp(L)]))
    return block, rotated, inter


def test_reshape_chain_512_int64(C):
    shifts = np.random.default_rng(11).integers(0, L, 3)
    comm = C.dm_world_create(P)
    block, rotated, inter = reshape_chain_layouts(C, comm, shifts)
    slots = [C.extent(d) for d in (block, rotated, inter)]
    bufs = [dm.DistBuffer(comm, s * 8) for s in slots]
    idx0 = flat_index(block, slots[0])
    host = np.full(P * slots[0], -1, np.int64)
    host[idx0.ravel()] = np.arange(N, dtype=np.int64)
    bufs[0].put_all(host.reshape(P, slots[0]))
    del host, idx0
    for b, s in zip(bufs[1:], slots[1:]):
        b.fill(0xFF, s * 8)   # every byte 0xFF: int64 -1
    t = dm.mpi_type("long_long")
    plans = [C.makeReshapePlan(comm, t, block, rotated), C.makeReshapePlan(comm, t, rotated, inter)]
    assert all(plans)
    C.runReshapePlan(comm, plans[0], bufs[0].ptr, bufs[1].ptr)
    C.runReshapePlan(comm, plans[1], bufs[1].ptr, bufs[2].ptr)
    for d, b, s in ((rotated, bufs[1], slots[1]), (inter, bufs[2], slots[2])):
        got = b.get_all(s * 8, np.int64).ravel()
        idx = flat_index(d, s).ravel()
        assert np.array_equal(got[idx], np.arange(N, dtype=np.int64))
        mask = np.ones(got.size, bool)
        mask[idx] = False
        assert np.all(got[mask] == -1)
        del got, idx, mask
    for p in plans:
        C.freeReshapePlan(p)
    for b in bufs:
        b.free()
    for d in (block, rotated, inter):
        C.freeDecomp(d)
    C.dm_world_free(comm)


def planted_input(seed=7):
    rng = np.random.default_rng(seed)
    modes = rng.integers(0, L, (8, 3))
    x = rng.uniform(-1, 1, (L, L, L)) + 1j * rng.uniform(-1, 1, (L, L, L))
    n = np.arange(L)
    for k in modes:
        e = [np.exp(2j * np.pi * k[d] * n / L) for d in range(3)]
        x += 10.0 * e[0][:, None, None] * e[1][None, :, None] * e[2][None, None, :]
    return x, modes


def dft3(x):
    try:
        import scipy.fft
        return scipy.fft.fftn(x, workers=16)
    except ImportError:
        return np.fft.fftn(x)


@pytest.fixture(scope="module")
def c3_data():
    x, modes = planted_input()
    return x, dft3(x), modes


PLANS = [("default_B32", [32, 32, 32]), ("B_1_1_256", [1, 1, 256]), ("autotuned", None),
         ("phase_order_B32", [32, 32, 32])]


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("name,factors", PLANS, ids=[p[0] for p in PLANS])
def test_split_fft_512_vs_fp64_dft(C, c3_data, dbl, name, factors):
    """The reference's default split (A = 16, B = 32 on every axis), the
    mixed family B = (1, 1, 256), whatever autotuneParallelPlan[f] returns
    (the plan bench.py times) and the default split run in the reference's
    phase order, each against the FP64 DFT, the planted modes and the
    round trip on every rank."""
    x, X, modes = c3_data
    comm = C.dm_world_create(P)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([C.makeLocalDecomp(L), C.makeBlockDecomp(L, 4),
                                                 C.makeBlockDecomp(L, 8)]))
    per = C.extent(start)
    es = 16 if dbl else 8
    ct = np.complex128 if dbl else np.complex64
    length = 2 * per + 1024
    dat = dm.DistBuffer(comm, per * es)
    work = dm.DistBuffer(comm, length * es)
    sidx = flat_index(start, per)
    host = np.zeros(P * per, ct)
    host[sidx.ravel()] = x.ravel()
    dat.put_all(host.reshape(P, per))
    del host
    if factors is None:
        tune = C.autotuneParallelPlan if dbl else C.autotuneParallelPlanf
        plan = tune(start, dat.ptr, length, work.ptr)
        # the autotuner timed candidates on dat: put the input back
        host = np.zeros(P * per, ct)
        host[sidx.ravel()] = x.ravel()
        dat.put_all(host.reshape(P, per))
        del host
    else:
        make = C.makeParallelPlan if dbl else C.makeParallelPlanf
        plan = make(start, dm.ints(factors), dat.ptr, length, work.ptr)
    assert plan
    if name.startswith("phase_order"):
        assert C.dm_split_set_phase_order(plan, 1) == 0
    sp = dm.split_plan(plan)
    A = [sp.A_factors[i] for i in range(3)]
    if factors is not None:
        assert A == [L // b for b in factors]
    print(f"plan {name}: A = {A}", flush=True)
    ext = sp.extent
    C.runSplitPlan(plan, dm.FFTW_FORWARD)
    got = work.get_all(ext * es, ct).ravel()
    spec = got[flat_index(C.getTransformed(plan), ext)].astype(np.complex128)
    del got
    tol = 1e-12 if dbl else 1e-6
    err = rel_l2(spec, X)
    assert err < tol, f"forward relative L2 {err:.3e} (A = {A})"
    for k in modes:
        assert abs(abs(spec[tuple(k)]) / (10.0 * N) - 1) < 0.01
    del spec
    C.runSplitPlan(plan, dm.FFTW_BACKWARD)
    back = dat.get_all(per * es, ct).ravel()[sidx].astype(np.complex128) / N
    rt = rel_l2(back, x)
    assert rt < tol, f"round trip relative L2 {rt:.3e}"
    C.freeSplitPlan(plan)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)
