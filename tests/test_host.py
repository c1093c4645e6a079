"""Host-side logic of the product library, no GPU needed: descriptor
functions, per-rank decomposition queries, processor grids, split-plan
factors and output placement, reshape-plan message lists, error paths.
Compared with the pinned oracle and tests/refmodel.py."""
import ctypes
import itertools

import numpy as np
import pytest

import decomp_mpi_amd as dm
import oracle
import refmodel as rm
from helpers import prod_dim
from test_oracle import DESCS

MAPPER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int)


@pytest.fixture(scope="module")
def C():
    return dm.lib()


def dim_fields(ptr):
    return ctypes.cast(ptr, ctypes.POINTER(dm.DimDecompS)).contents


@pytest.mark.parametrize("d", DESCS, ids=str)
def test_descriptor_functions(C, d):
    """rank_offset / type_offset through the struct's function table, sizes,
    and clone (reshape.h:45-67)."""
    p = prod_dim(C, d)
    f = dim_fields(p)
    assert (f.coord_size, f.proc_size) == (rm.coord_size(d), rm.proc_size(d))
    ro, to = MAPPER(f.rank_offset), MAPPER(f.type_offset)
    for g in range(f.coord_size):
        assert (ro(p, g), to(p, g)) == rm.rank_off(d, g), (d, g)
    c = C.cloneDimDecomp(p)
    assert c and c != p
    fc = dim_fields(c)
    assert [MAPPER(fc.type_offset)(c, g) for g in range(f.coord_size)] == [to(p, g) for g in range(f.coord_size)]
    assert C.sizeDimDecomp(p) == f.coord_size
    C.freeDimDecomp(c)
    C.freeDimDecomp(p)


def test_exported_offset_functions(C):
    """The per-kind *_rank_offset / *_type_offset exports (block_decomp.h etc.)."""
    cases = [("block", ("block", 10, 3)), ("local", ("local", 5)), ("single", ("single", 6, 2, 3)),
             ("rotate", ("rotate", 4, ("block", 9, 2))), ("interleave", ("interleave", ("block", 4, 2), 3)),
             ("factor_reverse", ("frev", 3, 4, ("block", 12, 2)))]
    for kind, d in cases:
        p = prod_dim(C, d)
        rfn, tfn = getattr(C, kind + "_rank_offset"), getattr(C, kind + "_type_offset")
        for g in range(rm.coord_size(d)):
            assert (rfn(p, g), tfn(p, g)) == rm.rank_off(d, g)
        C.freeDimDecomp(p)


def test_constructor_null_returns(C):
    """block_decomp.c:37-39 and factor_reverse_decomp.c:49-52."""
    assert not C.makeBlockDecompWithFactor(10, 3, 2)
    n = C.makeLocalDecomp(12)
    assert not C.makeFactorReverseDecomp(5, 3, n)
    assert C.makeFactorReverseDecomp(4, 3, n)
    C.freeDimDecomp(n)


LAYOUTS = [
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)], None, None),
    (6, [("block", 12, 5), ("rotate", 7, ("local", 8)), ("interleave", ("block", 4, 1), 4)], None, None),
    (6, [("block", 12, 3), ("local", 8), ("interleave", ("block", 4, 2), 4)], [2, 1, 1], None),
    (6, [("block", 12, 2), ("local", 8), ("block", 16, 3)], [1, 1, 2], [1, 6, 48]),
    (3, [("block", 10, 3), ("frev", 2, 5, ("local", 10))], None, None),
    (8, [("interleave", ("block", 4, 4), 8), ("block", 16, 2, 4)], None, None),
]


def prod_view(C, comm, r, descs, rs, strides):
    """The decomposition as virtual rank r builds it (SPMD view)."""
    v = C.dm_comm_view(comm, r)
    if strides is None and rs is None:
        return C.makePackedDecomp(v, len(descs), dm.ptrs([prod_dim(C, d) for d in descs]))
    st = strides if strides is not None else [0] * len(descs)
    axes = [C.makeAxisDecomp(v, prod_dim(C, d), st[i], rs[i]) for i, d in enumerate(descs)]
    return C.makeDecomp(len(descs), dm.ptrs(axes))


@pytest.mark.parametrize("case", range(len(LAYOUTS)))
def test_decomp_queries_per_rank(C, case):
    """makeAxisDecomp / makeDecomp / makePackedDecomp (reshape.c:47-248) and
    the rank-local queries (reshape.c:281-324) on every virtual rank."""
    P, descs, rs, strides = LAYOUTS[case]
    if rs is not None and strides is None:
        # packed memory strides, explicit rank strides: compute the packing
        pk = rm.Layout(P, descs, rs)
        strides = None
    want = rm.Layout(P, descs, rs, strides)
    comm = C.dm_world_create(P)
    for r in range(P):
        if rs is not None and strides is None:
            st = [int(x) for x in pk.m[r]]
            pd = prod_view(C, comm, r, descs, rs, st)
        else:
            pd = prod_view(C, comm, r, descs, rs, strides)
        assert pd
        d = dm.decomp(pd)
        assert d.proc_used == want.proc_used
        if r < want.proc_used:
            assert C.myRank(pd) == r
        for i in range(len(descs)):
            ax = d.dims[i].contents
            assert ax.rank_stride == want.rs[i]
            assert ax.my_rank_in_dimension == (r // want.rs[i]) % rm.proc_size(descs[i])
            assert ax.max_local == want.maxl[r, i]
        if rs is None or strides is not None:
            assert C.extent(pd) == want.extent[r]
        for g in itertools.product(*[range(n) for n in want.L]):
            owner, off = want.owner(g)
            # isLocal compares positions only (reshape.c:101-103), so a rank
            # past proc_used still sees the coordinates of its positions
            local = all(want.tab[i][g[i]][0] == ax_pos(want, r, i) for i in range(len(g)))
            assert local == (owner == r) or r >= want.proc_used
            assert C.areCoordsLocal(pd, dm.ints(list(g))) == int(local)
            for i in range(len(g)):
                assert C.isCoordLocal(pd, i, g[i]) == int(want.tab[i][g[i]][0] == ax_pos(want, r, i))
            if owner == r and (rs is None or strides is not None):
                assert C.getOffset(pd, dm.ints(list(g))) == off
        C.freeDecomp(pd)
    C.dm_world_free(comm)


def ax_pos(lay, r, i):
    return (r // lay.rs[i]) % rm.proc_size(lay.descs[i])


@pytest.mark.parametrize("shape,grid", [((8, 6, 10), (2, 3, 1)), ((12, 4, 8), (1, 2, 4)), ((6, 6), (3, 2))])
def test_even_packed_block_is_c_order(C, shape, grid):
    """Closed form for makePackedDecomp (reshape.c:203-248) on evenly
    divided blocks: each rank owns one block of L_d / p_d points per axis,
    stores it as the C-order array of those local sizes, and the blocks of
    the P ranks are distinct."""
    P = int(np.prod(grid))
    loc = [n // p for n, p in zip(shape, grid)]
    comm = C.dm_world_create(P)
    blocks = set()
    for r in range(P):
        pd = C.makePackedDecomp(C.dm_comm_view(comm, r), len(shape),
                                dm.ptrs([prod_dim(C, ("block", n, p)) for n, p in zip(shape, grid)]))
        assert C.extent(pd) == int(np.prod(loc))
        mine = []
        for g in itertools.product(*[range(n) for n in shape]):
            if not C.areCoordsLocal(pd, dm.ints(list(g))):
                continue
            rel = [x % l for x, l in zip(g, loc)]
            assert C.getOffset(pd, dm.ints(list(g))) == int(np.ravel_multi_index(rel, loc))
            mine.append(tuple(x // l for x, l in zip(g, loc)))
        assert len(mine) == int(np.prod(loc)) and len(set(mine)) == 1
        blocks.add(mine[0])
        C.freeDecomp(pd)
    assert len(blocks) == P
    C.dm_world_free(comm)


def test_compatible(C):
    comm = C.dm_world_create(4)
    a = C.makePackedDecomp(comm, 2, dm.ptrs([prod_dim(C, ("block", 8, 2)), prod_dim(C, ("block", 6, 2))]))
    b = C.makePackedDecomp(comm, 2, dm.ptrs([prod_dim(C, ("local", 8)), prod_dim(C, ("block", 6, 4))]))
    c = C.makePackedDecomp(comm, 2, dm.ptrs([prod_dim(C, ("local", 8)), prod_dim(C, ("block", 7, 4))]))
    assert C.compatible(a, b) == 1 and C.compatible(a, c) == 0
    for x in (a, b, c):
        C.freeDecomp(x)
    C.dm_world_free(comm)


def test_processor_grid_matches_oracle(C):
    for P in (1, 2, 4, 6, 8, 32):
        comm = C.dm_world_create(P)
        for tasks in [(64, 64, 64), (512, 512, 512), (16, 16, 16), (32, 32, 32), (3, 8, 5), (12, 7), (1, 9, 4)]:
            g = C.makeProcessorGrid(comm, len(tasks), dm.ints(list(tasks)))
            got = [g[i] for i in range(len(tasks))]
            assert got == oracle.processor_grid(P, list(tasks)), (P, tasks)
        C.dm_world_free(comm)


class HostBuffers:
    """Registered (never dereferenced) memory for plan construction."""

    def __init__(self, C, comm, P, slot_elems):
        self.C = C
        self.arr = np.zeros(P * slot_elems * 2 + 64, np.float64)
        self.ptr = self.arr.ctypes.data
        assert C.dm_register(comm, self.ptr, slot_elems * 16) == 0

    def free(self):
        self.C.dm_unregister(self.ptr)


@pytest.mark.parametrize("shape,grid,P", [((64, 64, 64), (1, 2, 2), 4), ((512, 512, 512), (1, 4, 8), 32),
                                          ((16, 8, 32), (2, 2, 2), 8)])
def test_split_plan_factors_and_placement(C, shape, grid, P):
    """makeSplitPlan factors (split_fft.c:590-601: A = largest divisor
    <= sqrt(L)), extent (split_fft.c:299-308) and getTransformed's placement:
    frequency k along each axis at B-position k/B + A (k%B)."""
    comm = C.dm_world_create(P)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("block", n, g)) for n, g in zip(shape, grid)]))
    per = C.extent(start)
    length = 2 * per + 1024
    dat = HostBuffers(C, comm, P, per)
    work = HostBuffers(C, comm, P, length)
    plan = C.makeSplitPlan(0, start, dat.ptr, length, work.ptr)
    assert plan
    sp = dm.split_plan(plan)
    A = [sp.A_factors[i] for i in range(3)]
    assert A == [max(a for a in range(1, n + 1) if n % a == 0 and a * a <= n) for n in shape]
    B = [n // a for n, a in zip(shape, A)]
    ga, gb = oracle.processor_grid(P, A), oracle.processor_grid(P, B)
    assert sp.extent == oracle.split_extent(P, A, B, ga, gb)
    tr = C.getTransformed(plan)
    assert tr
    td, bd = dm.decomp(tr), dm.decomp(sp.B)
    for i in range(3):
        ft, fb = td.dims[i].contents.decomp.contents, bd.dims[i].contents.decomp.contents
        assert td.dims[i].contents.memory_stride == bd.dims[i].contents.memory_stride
        tdp = ctypes.cast(td.dims[i].contents.decomp, ctypes.c_void_p).value
        bdp = ctypes.cast(bd.dims[i].contents.decomp, ctypes.c_void_p).value
        for k in range(shape[i]):
            p = rm.split_position(k, A[i], B[i])
            assert MAPPER(ft.rank_offset)(tdp, k) == MAPPER(fb.rank_offset)(bdp, p)
            assert MAPPER(ft.type_offset)(tdp, k) == MAPPER(fb.type_offset)(bdp, p)
    C.freeSplitPlan(plan)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)


def test_slab_plan_factors(C):
    """makeSplitPlan(slab=1): B = L on all axes but the last, B_last = 1
    (split_fft.c:584-589)."""
    comm = C.dm_world_create(4)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("block", 16, 1)), prod_dim(C, ("block", 16, 2)),
                                                 prod_dim(C, ("block", 16, 2))]))
    per = C.extent(start)
    dat, work = HostBuffers(C, comm, 4, per), HostBuffers(C, comm, 4, 4 * per + 64)
    plan = C.makeSplitPlan(1, start, dat.ptr, 4 * per + 64, work.ptr)
    assert plan
    sp = dm.split_plan(plan)
    assert [sp.A_factors[i] for i in range(3)] == [1, 1, 16]
    C.freeSplitPlan(plan)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)


def test_build_plan_rejects_short_workspace(C):
    """buildParallelPlan returns NULL when length < 2 * extent
    (split_fft.c:304-308)."""
    comm = C.dm_world_create(4)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("block", 16, 1)), prod_dim(C, ("block", 16, 2)),
                                                 prod_dim(C, ("block", 16, 2))]))
    per = C.extent(start)
    dat, work = HostBuffers(C, comm, 4, per), HostBuffers(C, comm, 4, 4 * per)
    A, B, L = [4, 4, 4], [4, 4, 4], [16, 16, 16]
    ga, gb = oracle.processor_grid(4, A), oracle.processor_grid(4, B)
    ext = oracle.split_extent(4, A, B, ga, gb)
    args = lambda n: (0, 0, 0, start, 3, dm.ints(A), dm.ints(B), dm.ints(L), dm.ints(ga), dm.ints(gb),
                      dat.ptr, n, work.ptr)
    assert not C.buildParallelPlan(*args(2 * ext - 1))
    plan = C.buildParallelPlan(*args(2 * ext))
    assert plan
    C.freeSplitPlan(plan)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)


def test_reshape_plan_bounds(C):
    """checkReshapePlanBounds (reshape.c:523-567): 0 when the arrays hold
    every message, 1 when one is short."""
    comm = C.dm_world_create(4)
    a = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("local", 16)), prod_dim(C, ("block", 16, 2)),
                                             prod_dim(C, ("block", 16, 2))]))
    b = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("block", 16, 2)), prod_dim(C, ("local", 16)),
                                             prod_dim(C, ("block", 16, 2))]))
    plan = C.makeReshapePlan(comm, dm.mpi_type("c_double_complex"), a, b)
    assert plan
    n = 16 * 8 * 8
    assert C.checkReshapePlanBounds(plan, 16, n, n) == 0
    assert C.checkReshapePlanBounds(plan, 16, n, n - 1) == 1
    assert C.checkReshapePlanBounds(plan, 16, n - 1, n) == 1
    C.freeReshapePlan(plan)
    C.freeDecomp(a)
    C.freeDecomp(b)
    C.dm_world_free(comm)


def test_error_paths_record_mode(C):
    """raise_error's message text (reshape.c:12-21) and the failing calls:
    illegal rank stride (reshape.c:59-61), rank out of range
    (reshape.c:196-200), duplicate rank strides (reshape.c:205-215)."""
    C.dm_set_error_mode(1)
    try:
        comm = C.dm_world_create(4)
        C.dm_clear_error()
        C.makeAxisDecomp(comm, C.makeBlockDecomp(8, 4), 1, 2)
        assert b"Error in reshape:" in C.dm_last_error()
        assert b"rank_stride" in C.dm_last_error()
        C.dm_clear_error()
        big = C.dm_world_create(8)
        ax = [C.makeAxisDecomp(big, C.makeBlockDecomp(8, 2), 1, 2), C.makeAxisDecomp(big, C.makeBlockDecomp(8, 2), 8, 2)]
        assert not C.makeDecomp(2, dm.ptrs(ax))
        assert b"duplicate rank strides" in C.dm_last_error()
        C.dm_world_free(big)
        C.dm_clear_error()
        ax = [C.makeAxisDecomp(comm, C.makeBlockDecomp(8, 2), 1, 2), C.makeAxisDecomp(comm, C.makeBlockDecomp(8, 4), 8, 1)]
        assert not C.makeDecomp(2, dm.ptrs(ax))
        assert b"out of range" in C.dm_last_error()
        C.dm_world_free(comm)
    finally:
        C.dm_clear_error()
        C.dm_set_error_mode(0)


def test_batch_plan_byte_counts(C):
    """dm_batch_bytes counts one read and one write per axis of every
    transform; with DM_BATCH_TILE_PLANES, dm_batch_bytes_moved counts one
    pass fewer for each transform whose last two axes make a plane tile:
    complex float up to 8192 points per plane, complex double 4096, every
    side at most 256 (planning only, no device)."""
    shapes = [(32, 48, 64), (64, 96, 128), (40, 128, 64), (8, 512, 8), (2, 16, 16)]
    for dbl, plane_ok in ((0, [True, False, True, False, True]), (1, [True, False, False, False, True])):
        es = 16 if dbl else 8
        plan = C.dm_batch_plan(len(shapes), 3, dm.ints([v for s in shapes for v in s]), dbl, -1)
        assert plan
        alg = sum(2 * 3 * int(np.prod(s)) * es for s in shapes)
        assert C.dm_batch_bytes(plan) == alg
        assert C.dm_batch_bytes_moved(plan) == alg          # line tiles: three passes each
        assert C.dm_batch_set_schedule(plan, dm.DM_BATCH_TILE_PLANES, 0) == 0
        moved = sum(2 * (2 if ok else 3) * int(np.prod(s)) * es for s, ok in zip(shapes, plane_ok))
        assert C.dm_batch_bytes_moved(plan) == moved
        C.dm_batch_free(plan)
