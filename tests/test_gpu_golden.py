"""Device results against the committed golden fixtures (tests/golden):
test.c's reshape cases and the descriptor zoo bit for bit, and the split
FFT of small cubes against numpy's FP64 DFT (1e-12 relative L2; the FP32
form within 1e-6)."""
import os

import numpy as np
import pytest

import decomp_mpi_amd as dm
import make_golden
import oracle
import refmodel as rm
from helpers import elem_type, flat_index, prod_dim, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


def prod_layout(C, comm, descs, rs=None, strides=None):
    if rs is None and strides is None:
        return C.makePackedDecomp(comm, len(descs), dm.ptrs([prod_dim(C, d) for d in descs]))
    if strides is None:
        raise ValueError("explicit rank strides need explicit memory strides here")
    rs = rs or rm.Layout(1, descs).rs
    axes = [C.makeAxisDecomp(comm, prod_dim(C, d), strides[i], rs[i]) for i, d in enumerate(descs)]
    return C.makeDecomp(len(descs), dm.ptrs(axes))


def run_reshape(C, comm, sd, dd, src, init, elem):
    P = src.shape[0]
    plan = C.makeReshapePlan(comm, elem_type(C, elem), sd, dd)
    assert plan
    sb = dm.DistBuffer(comm, src.shape[1] * elem)
    db = dm.DistBuffer(comm, init.shape[1] * elem)
    sb.put_all(np.ascontiguousarray(src))
    db.put_all(np.ascontiguousarray(init))
    C.runReshapePlan(comm, plan, sb.ptr, db.ptr)
    got = db.get_all(init.shape[1] * elem, init.dtype)
    C.freeReshapePlan(plan)
    sb.free()
    db.free()
    assert got.shape == (P, init.shape[1])
    return got


@pytest.mark.parametrize("npe", [1, 2, 3, 4, 6])
def test_golden_test_c(C, npe):
    f = np.load(os.path.join(GOLD, "reshape_testc.npz"))
    S = 10
    comm = C.dm_world_create(npe)
    lay = {
        "orig": prod_layout(C, comm, [("single", S, 0, 1), ("single", S, 0, npe)], [1, 1], [1, S]),
        "new": prod_layout(C, comm, [("rotate", 2, ("local", S)), ("rotate", 1, ("block", S, npe))], [1, 1], [1, S]),
    }
    pairs = [("orig", "new")]
    if npe > 1 and npe % 2 == 0:
        lay["d2"] = prod_layout(C, comm, [("rotate", 3, ("block", S, 2)), ("rotate", 3, ("block", S, npe // 2))],
                                [1, 2], [1, S // 2])
        pairs += [("orig", "d2"), ("new", "d2")]
    for s, d in pairs:
        key = f"npe{npe}_{s}_{d}"
        got = run_reshape(C, comm, lay[s], lay[d], f[key + "_src"], f[key + "_init"], 4)
        assert np.array_equal(got, f[key + "_want"]), key
    for v in lay.values():
        C.freeDecomp(v)
    C.dm_world_free(comm)


@pytest.mark.parametrize("case", range(len(make_golden.ZOO)))
def test_golden_zoo(C, case):
    f = np.load(os.path.join(GOLD, "reshape_zoo.npz"))
    P, sd, dd, drs, dst = make_golden.ZOO[case]
    comm = C.dm_world_create(P)
    src = prod_layout(C, comm, sd)
    if drs is not None and dst is None:
        dst = None
        pk = rm.Layout(P, dd, drs)
        assert all((pk.m[r] == pk.m[0]).all() for r in range(P))
        dst = [int(v) for v in pk.m[0]]
    dstd = prod_layout(C, comm, dd, drs, dst)
    want = f[f"zoo{case}_want"]
    init = np.full_like(want, -2)
    got = run_reshape(C, comm, src, dstd, f[f"zoo{case}_src"], init, 8)
    assert np.array_equal(got, want)
    C.freeDecomp(src)
    C.freeDecomp(dstd)
    C.dm_world_free(comm)


@pytest.mark.parametrize("case", range(len(make_golden.SPLIT)))
@pytest.mark.parametrize("dbl", [1, 0])
def test_golden_split_fft(C, case, dbl):
    f = np.load(os.path.join(GOLD, "split_fft.npz"))
    x, X = f[f"split{case}_x"], f[f"split{case}_X"]
    A = [int(a) for a in f[f"split{case}_A"]]
    B = [n // a for n, a in zip(x.shape, A)]
    P = 4
    grid = oracle.processor_grid(P, list(x.shape))
    comm = C.dm_world_create(P)
    start = prod_layout(C, comm, [("block", n, g) for n, g in zip(x.shape, grid)])
    per = C.extent(start)
    es = 16 if dbl else 8
    ct = np.complex128 if dbl else np.complex64
    length = 2 * per + 256
    dat, work = dm.DistBuffer(comm, per * es), dm.DistBuffer(comm, length * es)
    sidx = flat_index(start, per)
    host = np.zeros(P * per, ct)
    host[sidx.ravel()] = x.ravel()
    dat.put_all(host.reshape(P, per))
    plan = (C.makeParallelPlan if dbl else C.makeParallelPlanf)(start, dm.ints(B), dat.ptr, length, work.ptr)
    assert plan
    ext = dm.split_plan(plan).extent
    C.runSplitPlan(plan, dm.FFTW_FORWARD)
    got = work.get_all(ext * es, ct).ravel()
    spec = got[flat_index(C.getTransformed(plan), ext)].astype(np.complex128)
    tol = 1e-12 if dbl else 1e-6
    assert rel_l2(spec, X) < tol
    C.runSplitPlan(plan, dm.FFTW_BACKWARD)
    back = dat.get_all(per * es, ct).ravel()[sidx].astype(np.complex128) / x.size
    assert rel_l2(back, x) < tol
    C.freeSplitPlan(plan)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)
