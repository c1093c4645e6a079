"""Host-side pieces of the split-FFT plan and the index lists, no GPU needed:
makeIndexList (reshape.c:104-136) against a brute-force run list, the
plan's twiddle tables (split_fft.c:348-420) and decomp_scale
(dimscale.c:7-67) against the oracle for every view rank, visitData
(split_fft.c:862-889) against the decomposition's own placement, and the
plan-time rejection of axes the device passes cannot transform."""
import ctypes

import numpy as np
import pytest

import decomp_mpi_amd as dm
import oracle
import refmodel as rm
from helpers import prod_dim
from test_host import HostBuffers


@pytest.fixture(scope="module")
def C():
    return dm.lib()


def index_runs(ptr):
    ix = ctypes.cast(ptr, ctypes.POINTER(dm.IndexS)).contents
    return [(ix.offsets[i], ix.counts[i]) for i in range(ix.count)]


def brute_runs(mine, my_pos, remote, procs):
    """For ascending g held at my_pos in `mine`: append mine's local offset
    to the list of remote's owner of g, merging an offset that extends the
    previous run by one."""
    lists = [[] for _ in range(procs)]
    for g in range(rm.coord_size(mine)):
        pos, off = rm.rank_off(mine, g)
        if pos != my_pos:
            continue
        rr, _ = rm.rank_off(remote, g)
        runs = lists[rr]
        if runs and runs[-1][0] + runs[-1][1] == off:
            runs[-1] = (runs[-1][0], runs[-1][1] + 1)
        else:
            runs.append((off, 1))
    return lists


INDEX_CASES = [
    (("block", 12, 3), ("block", 12, 4)),
    (("block", 16, 2), ("interleave", ("block", 4, 2), 4)),
    (("rotate", 5, ("block", 12, 3)), ("block", 12, 2)),
    (("interleave", ("block", 4, 2), 3), ("frev", 3, 4, ("block", 12, 3))),
    (("local", 10), ("block", 10, 3)),
]


@pytest.mark.parametrize("mine,remote", INDEX_CASES, ids=str)
def test_index_list_matches_brute_force(C, mine, remote):
    """makeIndexList: per remote rank, the runs of my local offsets in
    ascending global order, NULL-terminated."""
    procs_m = rm.proc_size(mine)
    comm = C.dm_world_create(procs_m)
    for r in range(procs_m):
        view = C.dm_comm_view(comm, r)
        ax = C.makeAxisDecomp(view, prod_dim(C, mine), 1, 1)
        rd = prod_dim(C, remote)
        lst = C.makeIndexList(ax, rd)
        assert lst
        arr = ctypes.cast(lst, ctypes.POINTER(ctypes.c_void_p))
        want = brute_runs(mine, r, remote, rm.proc_size(remote))
        got = [index_runs(arr[i]) for i in range(len(want))]
        assert arr[len(want)] is None
        assert got == want, (r, got, want)
        C.freeIndexList(lst)
        C.freeDimDecomp(rd)
        C.freeAxisDecomp(ax)
    C.dm_world_free(comm)


SPLITS = [((16, 8, 12), [4, 2, 3], 4), ((16, 16, 16), [2, 4, 8], 8), ((12, 10, 8), [3, 5, 2], 6)]


def plan_for_rank(C, comm, r, shape, B, P):
    """A plan built on rank r's view of a block start decomposition."""
    grid = oracle.processor_grid(P, list(shape))
    view = C.dm_comm_view(comm, r)
    start = C.makePackedDecomp(view, 3, dm.ptrs([prod_dim(C, ("block", n, g)) for n, g in zip(shape, grid)]))
    per = max(oracle.Decomp(P, [(oracle.Block(n, g), 1) for n, g in zip(shape, grid)]).extents)
    length = 4 * int(np.prod(shape)) + 64
    dat, work = HostBuffers(C, comm, P, per), HostBuffers(C, comm, P, length)
    plan = C.makeParallelPlan(start, dm.ints(B), dat.ptr, length, work.ptr)
    assert plan
    return plan, start, dat, work


@pytest.mark.parametrize("shape,B,P", SPLITS)
def test_plan_twiddles_and_decomp_scale_match_oracle(C, shape, B, P):
    """plan->twiddle[i] is the oracle's table for the view rank (exact
    quarter turns, the rest within 2 ulp), and decomp_scale with those
    tables scales a rank's A-layout block as the oracle does, sign -1
    (forward) and +1 (conjugate)."""
    A = [n // b for n, b in zip(shape, B)]
    ga, gb = oracle.processor_grid(P, A), oracle.processor_grid(P, B)
    Ad, _ = oracle.split_decomps(P, A, B, ga, gb)
    comm = C.dm_world_create(P)
    for r in range(P):
        plan, start, dat, work = plan_for_rank(C, comm, r, shape, B, P)
        sp = dm.split_plan(plan)
        assert [sp.A_factors[i] for i in range(3)] == A
        tw = ctypes.cast(sp.twiddle, ctypes.POINTER(ctypes.c_void_p))
        tabs = []
        for i in range(3):
            want = oracle.twiddle_table(P, Ad, r, A, list(shape), i)
            got = np.ctypeslib.as_array(ctypes.cast(tw[i], ctypes.POINTER(ctypes.c_double)),
                                        (2 * len(want),)).view(np.complex128).copy()
            exact = (np.abs(want.real) == 1) | (np.abs(want.imag) == 1) | (want == 0)
            assert np.array_equal(got[exact], want[exact])
            assert np.array_equal(np.signbit(got[exact].real), np.signbit(want[exact].real))
            assert np.allclose(got, want, rtol=0, atol=4.5e-16)
            tabs.append(got)
        ext = int(Ad.extents[r])
        rng = np.random.default_rng(r)
        x = rng.uniform(-1, 1, ext) + 1j * rng.uniform(-1, 1, ext)
        for sign in (-1, 1):
            mine, ref = x.copy(), x.copy()
            C.decomp_scale(sp.A, sp.twiddle, mine.ctypes.data, sign)
            oracle.decomp_scale(P, Ad, r, A, list(shape), ref.view(np.float64), sign)
            assert np.allclose(mine, ref, rtol=0, atol=1e-15)
            assert not np.array_equal(mine, x)
        C.freeSplitPlan(plan)
        dat.free()
        work.free()
        C.freeDecomp(start)
    C.dm_world_free(comm)


VISITOR = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p)


def test_visit_data_reports_every_slot(C):
    """visitData calls the visitor once per slot of the rank's local
    bounding box; a slot that holds a coordinate gets that coordinate, the
    others -1 on the axes where they hold none."""
    P, shape, grid = 6, (10, 7, 9), (2, 3, 1)
    comm = C.dm_world_create(P)
    od = oracle.Decomp(P, [(oracle.Block(n, g), rs) for n, g, rs in zip(shape, grid, (3, 1, 1))])
    for r in range(P):
        view = C.dm_comm_view(comm, r)
        d = C.makePackedDecomp(view, 3, dm.ptrs([prod_dim(C, ("block", n, g)) for n, g in zip(shape, grid)]))
        ext = C.extent(d)
        buf = np.zeros(ext, np.complex128)
        base = buf.ctypes.data
        seen = {}

        def visit(nd, gpos, ptr):
            seen[(ptr - base) // 16] = tuple(gpos[i] for i in range(nd))

        C.visitData(d, VISITOR(visit), base)
        assert len(seen) == ext
        owned = {}
        for g in np.ndindex(*shape):
            rank, off = od.owner(g)
            if rank == r:
                owned[off] = g
        full = {o: g for o, g in seen.items() if min(g) >= 0}
        assert full == owned
        C.freeDecomp(d)
    C.dm_world_free(comm)


@pytest.mark.parametrize("L,dbl,A,ok,phased", [
    (272, 1, 16, True, False),       # 17 * 16: a prime above 13 runs as direct sums
    (17 * 8, 0, 8, True, False),
    (4099, 1, 1, False, None),       # a prime longer than a complex-double tile
    (8209, 0, 1, False, None),       # a prime longer than a complex-float tile
    (8192, 1, 1, False, None),       # longer than a pass tile, and so is B = 8192
    (8192, 1, 64, True, True),       # 64 x 128: phase order only
    (65536, 1, 256, True, True),
    (16384, 0, 128, True, True),
    (8192, 0, 64, True, False),      # one complex-float pass tile
    (4096, 1, 64, True, False),
    (16 * 27 * 5, 1, 36, True, False),
])
def test_split_plan_axis_support_at_plan_time(C, L, dbl, A, ok, phased):
    """An axis that neither one fused pass nor the plan's A x B factors can
    transform (a factor longer than one tile, a prime one included)
    makes buildParallelPlan return NULL, as the reference's other plan
    rejections do (split_fft.c:252-308), and makeSplitPlan raise 'Cannot
    generate plan' with the reason.  An axis longer than one pass tile
    whose factors each fit one gives a plan that runs in the reference's
    phase order (and refuses to leave it)."""
    comm = C.dm_world_create(2)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("local", L)), prod_dim(C, ("block", 2, 2)),
                                                 prod_dim(C, ("local", 1))]))
    per = C.extent(start)
    length = 4 * 2 * L + 64
    dat, work = HostBuffers(C, comm, 2, per), HostBuffers(C, comm, 2, length)
    As, B, Ls = [A, 1, 1], [L // A, 2, 1], [L, 2, 1]
    ga, gb = oracle.processor_grid(2, As), oracle.processor_grid(2, B)
    build = C.buildParallelPlan if dbl else C.buildParallelPlanf
    make = C.makeSplitPlan if dbl else C.makeSplitPlanf
    C.dm_set_error_mode(1)
    try:
        C.dm_clear_error()
        p = build(0, 0, 0, start, 3, dm.ints(As), dm.ints(B), dm.ints(Ls), dm.ints(ga), dm.ints(gb),
                  dat.ptr, length, work.ptr)
        assert bool(p) == ok
        assert C.dm_last_error() == b""   # a rejection, not an error
        if p:
            assert C.dm_split_phase_order(p) == int(phased)
            if phased:
                assert C.dm_split_set_phase_order(p, 0) != 0
                assert b"phase order only" in C.dm_last_error()
                C.dm_clear_error()
            else:
                assert C.dm_split_set_phase_order(p, 1) == 0 and C.dm_split_phase_order(p) == 1
                assert C.dm_split_set_phase_order(p, 0) == 0 and C.dm_split_phase_order(p) == 0
            C.freeSplitPlan(p)
        if A > 1 and A * A <= L and not any(L % q == 0 and A < q and q * q <= L for q in range(2, L)):
            # A is makeSplitPlan's own choice (largest divisor not above sqrt(L))
            p = make(0, start, dat.ptr, length, work.ptr)
            assert bool(p) == ok
            if ok:
                C.freeSplitPlan(p)
            else:
                msg = C.dm_last_error()
                assert b"Cannot generate plan" in msg and b"axis 0" in msg, msg
    finally:
        C.dm_clear_error()
        C.dm_set_error_mode(0)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)


def test_phase_order_flag(C):
    """RUN_PHASE_ORDER among buildParallelPlan's flags makes a plan that
    runs in phase order; without it the fused program is the default."""
    comm = C.dm_world_create(4)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("block", 16, 1)), prod_dim(C, ("block", 16, 2)),
                                                 prod_dim(C, ("block", 16, 2))]))
    per = C.extent(start)
    dat, work = HostBuffers(C, comm, 4, per), HostBuffers(C, comm, 4, 4 * per + 64)
    A, B, L = [4, 4, 4], [4, 4, 4], [16, 16, 16]
    ga, gb = oracle.processor_grid(4, A), oracle.processor_grid(4, B)
    for flags, want in ((0, 0), (dm.RUN_PHASE_ORDER, 1)):
        p = C.buildParallelPlan(flags, 0, 0, start, 3, dm.ints(A), dm.ints(B), dm.ints(L), dm.ints(ga),
                                dm.ints(gb), dat.ptr, 4 * per + 64, work.ptr)
        assert p and C.dm_split_phase_order(p) == want
        C.freeSplitPlan(p)
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)


def test_twiddle_mode_setting(C):
    """dm_split_set_twiddles: compensated products by default, plain on
    request, unknown modes and NULL plans refused."""
    comm = C.dm_world_create(4)
    start = C.makePackedDecomp(comm, 3, dm.ptrs([prod_dim(C, ("block", 16, 1)), prod_dim(C, ("block", 16, 2)),
                                                 prod_dim(C, ("block", 16, 2))]))
    per = C.extent(start)
    dat, work = HostBuffers(C, comm, 4, per), HostBuffers(C, comm, 4, 4 * per + 64)
    for make in (C.makeSplitPlanf, C.makeSplitPlan):
        p = make(0, start, dat.ptr, 4 * per + 64, work.ptr)
        assert p and C.dm_split_twiddles(p) == dm.DM_TWIDDLES_COMPENSATED
        assert C.dm_split_set_twiddles(p, dm.DM_TWIDDLES_PLAIN) == 0
        assert C.dm_split_twiddles(p) == dm.DM_TWIDDLES_PLAIN
        C.dm_set_error_mode(1)
        assert C.dm_split_set_twiddles(p, 7) != 0
        C.dm_clear_error()
        C.dm_set_error_mode(0)
        assert C.dm_split_twiddles(p) == dm.DM_TWIDDLES_PLAIN
        assert C.dm_split_set_twiddles(p, dm.DM_TWIDDLES_COMPENSATED) == 0
        assert C.dm_split_twiddles(p) == dm.DM_TWIDDLES_COMPENSATED
        C.freeSplitPlan(p)
    assert C.dm_split_set_twiddles(None, 0) != 0 and C.dm_split_twiddles(None) == -1
    dat.free()
    work.free()
    C.freeDecomp(start)
    C.dm_world_free(comm)


IO_CASES = [
    (6, [("block", 10, 2), ("rotate", 3, ("block", 7, 3)), ("local", 5)], None, None),
    (4, [("interleave", ("block", 4, 2), 3), ("block", 9, 2)], None, None),
    (8, [("block", 6, 2), ("local", 4), ("frev", 2, 4, ("block", 8, 2))], None, None),
    (6, [("block", 8, 2), ("local", 6), ("block", 10, 3)], [1, 1, 2], [1, 8, 48]),
    (5, [("block", 6, 2), ("local", 4)], None, None),     # ranks 2..4 hold nothing
]


@pytest.mark.parametrize("P,descs,rs,strides", IO_CASES)
def test_io_type_view_is_the_row_major_file_view(C, P, descs, rs, strides):
    """makeIOType (reshape.c:494-521): rank r's view lists, in the order of
    its loops over its coordinates (first axis outermost, ascending), the
    byte offsets of its elements in the global row-major file; the type
    spans the whole file, and is empty on ranks the decomposition leaves
    out.  expandIOType returns the next axis' stride (reshape.c:158-172)."""
    from helpers import make_pair
    shape = [rm.coord_size(d) for d in descs]
    es = 16
    world = C.dm_world_create(P)
    _, od = make_pair(C, world, P, descs, rank_strides=rs, strides=strides)
    for r in range(P):
        view = C.dm_comm_view(world, r)
        pd, _ = make_pair(C, view, P, descs, rank_strides=rs, strides=strides)
        t = ctypes.c_void_p()
        C.makeIOType(ctypes.byref(t), dm.mpi_type("c_double_complex"), pd)
        n = C.dm_type_view(t, None, 0)
        offs = (ctypes.c_longlong * max(n, 1))()
        assert C.dm_type_view(t, offs, n) == n
        want = []
        for g in np.ndindex(*shape):
            rank, _ = od.owner(g)
            if rank == r:
                want.append(int(np.ravel_multi_index(g, shape)) * es)
        assert list(offs)[:n] == want
        ext = ctypes.c_long()
        C.dm_type_get_extent(t, None, ctypes.byref(ext))
        assert ext.value == (int(np.prod(shape)) * es if r < od.proc_used else 0)
        C.dm_type_free(ctypes.byref(t))
        # the last axis alone: stride 1 -> returns its length
        ax = dm.decomp(pd).dims[len(shape) - 1]
        t2 = ctypes.c_void_p()
        assert C.expandIOType(ctypes.byref(t2), es, dm.mpi_type("c_double_complex"),
                              ctypes.cast(ax, ctypes.c_void_p).value, 1) == shape[-1]
        C.dm_type_free(ctypes.byref(t2))
        C.freeDecomp(pd)
    C.dm_world_free(world)


class DatatypeHead(ctypes.Structure):
    """The leading fields of this library's datatype (dm_mpi.h's
    dm_datatype): bytes, extent, predefined, count, true lower bound and
    true extent."""
    _fields_ = [("size", ctypes.c_size_t), ("extent", ctypes.c_long), ("predefined", ctypes.c_int),
                ("count", ctypes.c_int), ("true_lb", ctypes.c_long), ("true_extent", ctypes.c_long)]


MSG_CASES = [
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)], [("block", 8, 2), ("local", 8), ("block", 8, 2)]),
    (6, [("block", 12, 3), ("rotate", 2, ("block", 10, 2)), ("local", 4)],
     [("interleave", ("block", 4, 2), 3), ("local", 10), ("block", 4, 3)]),
    (6, [("block", 10, 2), ("local", 6), ("block", 9, 3)], [("single", 10, 0, 1), ("block", 6, 3), ("local", 9)]),
]


@pytest.mark.parametrize("case", range(len(MSG_CASES)))
def test_plan_message_lists_match_oracle(C, case):
    """makeReshapePlan's send and receive lists (reshape.c:394-420 built by
    expandList, typelist.c:74-119): for every rank, one entry per peer the
    oracle's literal message pairing has a message with, carrying that
    message's element count, true lower bound and true extent in bytes."""
    from helpers import make_pair
    P, sd, dd = MSG_CASES[case]
    es = 8
    world = C.dm_world_create(P)
    _, os_ = make_pair(C, world, P, sd)
    _, od = make_pair(C, world, P, dd)
    for r in range(P):
        view = C.dm_comm_view(world, r)
        ps, _ = make_pair(C, view, P, sd)
        pd, _ = make_pair(C, view, P, dd)
        plan = C.makeReshapePlan(view, dm.mpi_type("long"), ps, pd)
        assert plan
        rp = dm.reshape_plan(plan)
        for send, lst in ((True, rp.sends), (False, rp.recvs)):
            got = {}
            if lst:
                tl = lst.contents
                for k in range(tl.count):
                    e = tl.list[k]
                    h = ctypes.cast(e.type, ctypes.POINTER(DatatypeHead)).contents
                    got[e.rank] = (h.count, h.true_lb, h.true_extent)
            want = {}
            for peer in range(P):
                m = oracle.message(os_, od, r, peer, send)
                if m is not None:
                    want[peer] = (len(m), int(m.min()) * es, int(m.max() - m.min() + 1) * es)
            assert got == want, (r, send, got, want)
        C.freeReshapePlan(plan)
        C.freeDecomp(ps)
        C.freeDecomp(pd)
    C.dm_world_free(world)
