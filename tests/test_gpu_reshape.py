"""Device reshape (runReshapePlan) against the oracle's restatement of the
reference's message pairing (reshape.c:351-492, typelist.c:74-119):
bit-exact on every destination byte, including slots the reshape must leave
untouched."""
import numpy as np
import pytest

import decomp_mpi_amd as dm
import oracle
from helpers import elem_type, make_pair, oracle_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


def run_case(C, P, src_descs, dst_descs, elem, src_strides=None, dst_strides=None, src_rs=None, dst_rs=None,
             seed=0):
    """Random source and destination bytes; compare every destination byte
    with the oracle applied to the same bytes."""
    comm = C.dm_world_create(P)
    ps, os_ = make_pair(C, comm, P, src_descs, src_rs, src_strides)
    pd, od = make_pair(C, comm, P, dst_descs, dst_rs, dst_strides)
    plan = C.makeReshapePlan(comm, elem_type(C, elem), ps, pd)
    assert plan
    sb = dm.DistBuffer(comm, int(max(os_.extents)) * elem)
    db = dm.DistBuffer(comm, int(max(od.extents)) * elem)
    rng = np.random.default_rng(seed)
    hs = rng.integers(0, 256, (P, sb.slot), dtype=np.uint8)
    hd = rng.integers(0, 256, (P, db.slot), dtype=np.uint8)
    sb.put_all(hs)
    db.put_all(hd)
    C.runReshapePlan(comm, plan, sb.ptr, db.ptr)
    got = db.get_all(db.slot, np.uint8)
    want = hd.copy()
    oracle.reshape(os_, od, hs, want, elem)
    C.freeReshapePlan(plan)
    sb.free()
    db.free()
    C.dm_world_free(comm)
    return got, want


# test.c:118-231: a 10x10 int array on rank 0 (Single decomps), reshaped to
# {Rotated(2,Local), Rotated(1,Block(npe))}; with npe even also to
# {Rotated(3,Block(2)), Rotated(3,Block(npe/2)) stride 5 rank stride 2}.
SIZE = 10


def tc_orig(npe):
    return [("single", SIZE, 0, 1), ("single", SIZE, 0, npe)], [1, 1], [1, SIZE]


def tc_new(npe):
    return [("rotate", 2, ("local", SIZE)), ("rotate", 1, ("block", SIZE, npe))], [1, 1], [1, SIZE]


def tc_d2(npe):
    return [("rotate", 3, ("block", SIZE, 2)), ("rotate", 3, ("block", SIZE, npe // 2))], [1, 2], [1, SIZE // 2]


@pytest.mark.parametrize("npe", [1, 2, 3, 4, 6])
def test_reference_test_c_bit_exact(C, npe):
    o, n = tc_orig(npe), tc_new(npe)
    pairs = [(o, n)]
    if npe > 1 and npe % 2 == 0:
        d2 = tc_d2(npe)
        pairs += [(o, d2), (n, d2)]
    for (sd, srs, sst), (dd, drs, dst) in pairs:
        got, want = run_case(C, npe, sd, dd, 4, sst, dst, srs, drs, seed=npe)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("npe", [1, 2, 3, 4, 6])
def test_reference_test_c_self_index(C, npe):
    """printSlab's check (test.c:16-43): after the reshape every local
    element equals j + SIZE*i, the linear index of its coordinate."""
    cases = [(tc_orig(npe), tc_new(npe))]
    if npe > 1 and npe % 2 == 0:
        cases.append((tc_orig(npe), tc_d2(npe)))
    for (sd, srs, sst), (dd, drs, dst) in cases:
        comm = C.dm_world_create(npe)
        ps, os_ = make_pair(C, comm, npe, sd, srs, sst)
        pd, od = make_pair(C, comm, npe, dd, drs, dst)
        plan = C.makeReshapePlan(comm, dm.mpi_type("int"), ps, pd)
        sb = dm.DistBuffer(comm, SIZE * SIZE * 4)
        db = dm.DistBuffer(comm, SIZE * SIZE * 4)
        hs = np.zeros((npe, sb.slot // 4), np.int32)
        hs[0, :SIZE * SIZE] = np.arange(SIZE * SIZE)
        sb.put_all(hs)
        db.put_all(np.tile(-np.arange(npe, dtype=np.int32)[:, None], (1, db.slot // 4)))
        C.runReshapePlan(comm, plan, sb.ptr, db.ptr)
        got = db.get_all(db.slot, np.int32)
        errors = 0
        for i in range(SIZE):
            for j in range(SIZE):
                r, off = od.owner((j, i))
                if r is not None and got[r, off] != j + SIZE * i:
                    errors += 1
        assert errors == 0
        C.freeReshapePlan(plan)
        sb.free()
        db.free()


ZOO = [
    ("block", 12, 3), ("block", 12, 4, 2), ("local", 12), ("rotate", 5, ("block", 12, 3)),
    ("interleave", ("block", 4, 2), 3), ("frev", 3, 4, ("block", 12, 2)), ("single", 12, 1, 2),
    ("rotate", -7, ("interleave", ("block", 3, 3), 4)), ("block", 12, 5),
]


@pytest.mark.parametrize("seed", range(10))
def test_descriptor_zoo_bit_exact(C, seed):
    """Random pairs of 3-axis decompositions built from every descriptor
    type, uneven blocks included, packed or with explicit strides."""
    rng = np.random.default_rng(100 + seed)
    P = int(rng.choice([4, 6, 8, 12, 18, 30]))
    elem = int(rng.choice([1, 4, 8, 16, 24]))

    def pick():
        while True:
            descs = [ZOO[int(i)] for i in rng.integers(0, len(ZOO), 2)] + [("block", 8, 2)]
            prod = 1
            for d in descs:
                prod *= oracle.proc_size(oracle_spec(d))
            if prod <= P:
                return descs
    sd, dd = pick(), pick()
    if rng.integers(0, 2):
        # an axis permutation expressed by memory strides (fftw_test.c:277-281)
        got, want = run_case(C, P, sd, dd, elem, src_strides=[1, 64, 16], seed=seed)
    else:
        got, want = run_case(C, P, sd, dd, elem, seed=seed)
    assert np.array_equal(got, want)


# Both sides share their fastest axis but not its order along it: block
# against interleave (the split FFT's start -> A and A -> B reshapes), with
# rotation, uneven blocks and a factor-reversed order.
SAME_FASTEST = [
    (4, [("block", 8, 2), ("local", 6), ("block", 32, 2)],
     [("block", 8, 2), ("local", 6), ("interleave", ("block", 4, 2), 8)]),
    (8, [("local", 4), ("block", 12, 2), ("block", 48, 4)],
     [("local", 4), ("block", 12, 2), ("interleave", ("block", 8, 4), 6)]),
    (6, [("block", 6, 3), ("local", 5), ("rotate", 3, ("block", 20, 2))],
     [("block", 6, 3), ("local", 5), ("interleave", ("block", 4, 2), 5)]),
    (4, [("local", 3), ("interleave", ("block", 4, 2), 6), ("interleave", ("block", 8, 2), 4)],
     [("local", 3), ("block", 24, 2), ("block", 32, 2)]),
    (6, [("block", 5, 3), ("local", 7), ("frev", 4, 9, ("block", 36, 2))],
     [("block", 5, 3), ("local", 7), ("block", 36, 2)]),
    (4, [("local", 2), ("block", 10, 2), ("block", 50, 2)],
     [("local", 2), ("block", 10, 2), ("interleave", ("block", 10, 2), 5)]),
    (4, [("block", 8, 2), ("local", 6), ("interleave", ("block", 4, 2), 8)],
     [("block", 8, 2), ("local", 6), ("block", 32, 2)]),
    (8, [("local", 4), ("block", 12, 2), ("interleave", ("block", 8, 4), 6)],
     [("local", 4), ("block", 12, 2), ("block", 48, 4)]),
]


@pytest.mark.parametrize("elem", [4, 8, 16])
@pytest.mark.parametrize("case", range(len(SAME_FASTEST)))
def test_same_fastest_axis_bit_exact(C, case, elem):
    P, sd, dd = SAME_FASTEST[case]
    print(f"same fastest axis case {case}, {elem}-byte elements", flush=True)
    got, want = run_case(C, P, sd, dd, elem, seed=case)
    assert np.array_equal(got, want)


def pencil_descs(n, local_axis, grid=(2, 2)):
    it = iter(grid)
    return [("local", n) if d == local_axis else ("block", n, next(it)) for d in range(3)]


def test_pencil_64_c128_bit_exact(C):
    """The 64^3 complex-double pencil reshapes on a 2x2 grid."""
    n, P = 64, 4
    x, y, z = pencil_descs(n, 0), pencil_descs(n, 1), pencil_descs(n, 2)
    for a, b in [(x, y), (y, z), (z, x)]:
        got, want = run_case(C, P, a, b, 16, seed=7)
        assert np.array_equal(got, want)
