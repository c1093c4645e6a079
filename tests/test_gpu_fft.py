"""Device split FFT (runSplitPlan) against the oracle's restatement of the
reference's two-stage transform (split_fft.c:249-682, dimscale.c), and the
per-axis FFT stage against the oracle's reshape plus numpy's DFT.

Tolerances: complex double within 1e-12 relative L2 (the stated bound
against an FP64 FFT); the FP32 form within 1e-6 relative L2 vs the FP64
oracle."""
import numpy as np
import pytest

import decomp_mpi_amd as dm
import oracle
from helpers import make_pair, rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


def block_start(shape, grid):
    return [("block", n, g) for n, g in zip(shape, grid)]


def run_split(C, P, shape, grid, B=None, slab=None, dbl=True, seed=0, start_descs=None, autotune=False,
              phased=False, length=None, plain_twiddles=False):
    """Forward then backward on the device; forward checked against the
    oracle on every slot of the output layout and against numpy's FP64 DFT
    of the global array, backward as N * x.  phased: the plan runs in the
    reference's phase order, and every phase's timer must have run.
    plain_twiddles: complex-float products without the compensation."""
    comm = C.dm_world_create(P)
    descs = start_descs or block_start(shape, grid)
    ps, os_ = make_pair(C, comm, P, descs)
    es = 16 if dbl else 8
    per = int(max(os_.extents))
    N = int(np.prod(shape))
    length = length or 2 * max(per, N // P + 1) * 4 + 64
    dat = dm.DistBuffer(comm, per * es)
    work = dm.DistBuffer(comm, length * es)
    if autotune:
        plan = (C.autotuneParallelPlan if dbl else C.autotuneParallelPlanf)(ps, dat.ptr, length, work.ptr)
    elif slab is not None:
        plan = (C.makeSplitPlan if dbl else C.makeSplitPlanf)(slab, ps, dat.ptr, length, work.ptr)
    else:
        plan = (C.makeParallelPlan if dbl else C.makeParallelPlanf)(ps, dm.ints(B), dat.ptr, length, work.ptr)
    assert plan, "planning failed"
    if phased:
        assert C.dm_split_set_phase_order(plan, 1) == 0
    if plain_twiddles:
        assert C.dm_split_set_twiddles(plan, dm.DM_TWIDDLES_PLAIN) == 0
    sp = dm.split_plan(plan)
    nd = len(shape)
    A = [sp.A_factors[i] for i in range(nd)]
    Bf = [shape[i] // A[i] for i in range(nd)]
    ga = oracle.processor_grid(P, A)
    gb = oracle.processor_grid(P, Bf)
    ext = sp.extent
    assert ext == oracle.split_extent(P, A, Bf, ga, gb)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (P, per)) + 1j * rng.uniform(-1, 1, (P, per))
    dat.put_all(x.astype(np.complex128 if dbl else np.complex64))
    C.runSplitPlan(plan, dm.FFTW_FORWARD)
    if phased:
        t = [sp.times[i] for i in range(dm.NUM_TIMERS)]
        two = any(a > 1 for a in A)
        for name, i in (("RESHAPE_1", dm.RESHAPE_1), ("FFT_1", dm.FFT_1)):
            assert t[i] > 0 or (name == "FFT_1" and all(b == 1 for b in Bf)), (name, t)
        if two:
            assert t[dm.RESHAPE_2] > 0 and t[dm.FFT_2] > 0, t
        if two and any(a > 1 and b > 1 for a, b in zip(A, Bf)):
            assert t[dm.TWIDDLE] > 0, t
        assert abs(t[dm.ALL] - sum(t[:dm.ALL])) < 1e-9, t
    dt = np.complex128 if dbl else np.complex64
    got = work.get_all(ext * es, dt).astype(np.complex128)
    # oracle forward on the same input (the oracle does not restate plans
    # whose every B_d is 1: those are checked against numpy alone)
    out_d = oracle.split_decomps(P, A, Bf, ga, gb)[1 if any(a > 1 for a in A) else 0]
    fwd_err = 0.0
    if not all(b == 1 for b in Bf):
        od = x.astype(np.complex128).copy()
        ow = np.zeros((P, 2 * ext + 16), np.complex128)
        oracle.split_fft(os_, A, Bf, ga, gb, od, ow, ext, -1)
        mask = owned_mask(out_d, ext)
        fwd_err = rel_l2(got[mask], ow[:, :ext][mask])
    # independently of the oracle: numpy's DFT of the global array, with
    # frequency k of axis d at output position k / B_d + A_d * (k % B_d)
    # (A = 1 on every axis: the single-stage output in the A layout)
    sr, so = placement(os_)
    glob = x[sr, so]
    grids = np.meshgrid(*[np.arange(n) // b + a * (np.arange(n) % b) for n, a, b in zip(shape, A, Bf)],
                        indexing="ij")
    orank, ooff = placement(out_d)
    pos = tuple(grids)
    spec = got[orank[pos], ooff[pos]]
    fwd_err = max(fwd_err, rel_l2(spec, np.fft.fftn(glob)))
    # backward: unnormalised inverse restores N * x on the start layout
    C.runSplitPlan(plan, dm.FFTW_BACKWARD)
    back = dat.get_all(per * es, dt).astype(np.complex128)
    smask = owned_mask(os_, per)
    rt_err = rel_l2(back[smask] / N, x[smask])
    C.freeSplitPlan(plan)
    dat.free()
    work.free()
    C.dm_world_free(comm)
    return fwd_err, rt_err, A


def placement(d):
    """(rank, local offset) arrays over every global coordinate of the
    oracle decomposition d (per-rank strides), from per-axis tables."""
    n = d.ndim
    rank = np.zeros([1] * n, np.int64)
    parts = []
    for i, (spec, rs) in enumerate(d.axes):
        tab = np.array([oracle.rank_offset(spec, g) for g in range(d.L[i])], np.int64)
        shp = [1] * n
        shp[i] = d.L[i]
        rank = rank + (tab[:, 0] * rs).reshape(shp)
        parts.append(tab[:, 1].reshape(shp))
    rank = np.broadcast_to(rank, d.L)
    assert np.all(rank < d.proc_used)
    off = np.zeros(d.L, np.int64)
    for i in range(n):
        off = off + d.strides[rank, i] * parts[i]
    return rank, off


def owned_mask(d, width):
    """Slots of each rank that hold a coordinate of decomposition d."""
    mask = np.zeros((d.P, width), bool)
    r, off = placement(d)
    mask[r, off] = True
    return mask


@pytest.mark.parametrize("P,shape,grid,B", [
    (1, (8, 8, 8), (1, 1, 1), None),
    (4, (16, 16, 16), (1, 2, 2), None),
    (4, (16, 16, 16), (1, 2, 2), [4, 2, 8]),      # A != B on every axis
    (8, (16, 8, 32), (2, 2, 2), [2, 8, 4]),
    (4, (12, 12, 12), (1, 2, 2), None),           # radix 3
    (6, (12, 20, 14), (1, 2, 3), [3, 4, 7]),      # radices 3, 5, 7
    (4, (22, 26, 16), (1, 2, 2), [2, 2, 4]),      # radices 11, 13
    (4, (16, 16), (2, 2), None),                  # 2D
    (4, (16, 16, 16), (1, 2, 2), [16, 16, 16]),   # single stage (A = 1)
    (4, (16, 16, 16), (1, 2, 2), [1, 1, 8]),      # B_d = 1 beside B_d > 1
    (8, (16, 16, 16), (2, 2, 2), [1, 4, 1]),
    (8, (16, 16, 16), (1, 2, 4), [1, 1, 16]),     # the 512^3 bench family, scaled down
    (32, (64, 64, 64), (1, 4, 8), [1, 1, 32]),
    (4, (34, 38, 16), (1, 2, 2), [2, 2, 4]),      # primes 17, 19: direct-sum stages after a radix-2 one
    (4, (17, 23, 16), (1, 2, 2), None),           # prime axes, single direct-sum stage
    (2, (289, 6, 4), (2, 1, 1), [17, 2, 4]),      # 17 x 17: a direct-sum stage with stage twiddles
    (2, (254, 4, 4), (2, 1, 1), [2, 4, 4]),       # 2 x 127
    (2, (4093, 2, 2), (2, 1, 1), None),           # the largest prime within one complex-double tile
])
def test_split_fft_fp64_vs_oracle(C, P, shape, grid, B):
    if B is None:
        fwd, rt, A = run_split(C, P, shape, grid, slab=0)
    else:
        fwd, rt, A = run_split(C, P, shape, grid, B=B)
    assert fwd < TOL64, f"forward rel L2 {fwd:.3e} (A={A})"
    assert rt < TOL64, f"round trip {rt:.3e}"


def test_split_fft_slab(C):
    fwd, rt, _ = run_split(C, 4, (16, 16, 16), (1, 2, 2), slab=1)
    assert fwd < TOL64 and rt < TOL64


@pytest.mark.parametrize("P,shape,grid,B", [
    (4, (16, 16, 16), (1, 2, 2), None),
    (8, (32, 16, 32), (2, 2, 2), [4, 4, 8]),
    (8, (16, 16, 16), (2, 2, 2), [1, 4, 1]),
    (8, (16, 16, 16), (1, 2, 4), [1, 1, 16]),
    (4, (22, 26, 16), (1, 2, 2), [2, 2, 4]),      # radices 11, 13
    (4, (34, 38, 16), (1, 2, 2), [2, 2, 4]),      # primes 17, 19 (direct sums)
    (2, (289, 6, 4), (2, 1, 1), [17, 2, 4]),
    (2, (8191, 2, 2), (2, 1, 1), None),           # the largest prime within one complex-float tile
])
def test_split_fft_fp32_vs_fp64_oracle(C, P, shape, grid, B):
    if B is None:
        fwd, rt, _ = run_split(C, P, shape, grid, slab=0, dbl=False)
    else:
        fwd, rt, _ = run_split(C, P, shape, grid, B=B, dbl=False)
    assert fwd < TOL32, f"FP32 forward rel L2 {fwd:.3e}"
    assert rt < TOL32


@pytest.mark.parametrize("P,shape,grid,B", [
    (8, (128, 128, 128), (1, 2, 4), [1, 1, 64]),    # uniform strides: fused compensated store
    (6, (96, 80, 112), (1, 2, 3), [8, 8, 8]),       # uneven blocks: per-rank strides
])
def test_fp32_compensated_twiddles_beat_plain(C, P, shape, grid, B):
    """The compensated twiddle products (float part plus a bfloat16
    correction of each FP64 root) against the float part alone, on the same
    plan and input: both within the FP32 tolerance of the FP64 DFT, the
    compensated forward and round trip strictly closer."""
    comp = run_split(C, P, shape, grid, B=B, dbl=False, seed=3)
    plain = run_split(C, P, shape, grid, B=B, dbl=False, seed=3, plain_twiddles=True)
    print(f"FP32 {shape}: compensated fwd {comp[0]:.3e} rt {comp[1]:.3e}; plain fwd {plain[0]:.3e} rt {plain[1]:.3e}")
    assert comp[0] < TOL32 and plain[0] < TOL32
    assert comp[0] < plain[0] and comp[1] < plain[1]


PHASED = [
    (4, (16, 16, 16), (1, 2, 2), [4, 4, 4], None),
    (4, (16, 16, 16), (1, 2, 2), [4, 2, 8], None),     # A != B
    (6, (12, 20, 14), (1, 2, 3), [3, 4, 7], None),     # radices 3, 5, 7, uneven blocks
    (4, (22, 26, 16), (1, 2, 2), [2, 2, 4], None),     # 11- and 13-point second stages
    (4, (16, 16, 16), (1, 2, 2), [16, 16, 16], None),  # single stage (A = 1)
    (8, (16, 16, 16), (2, 2, 2), [1, 4, 1], None),     # B_d = 1 beside B_d > 1
    (2, (8, 8, 8), (1, 1, 2), [1, 1, 1], 4096),        # single second stage (B = 1)
    (4, (16, 16), (2, 2), [4, 4], None),               # 2D
    (4, (34, 38, 16), (1, 2, 2), [2, 2, 4], None),     # 17- and 19-point direct-sum second stages
]


@pytest.mark.parametrize("dbl", [True, False])
@pytest.mark.parametrize("P,shape,grid,B,length", PHASED)
def test_split_fft_phase_order_vs_oracle(C, P, shape, grid, B, length, dbl):
    """The reference's phase order on the device -- reshape, B_d-point
    DFTs, twiddle multiply with the plan's tables, reshape, A_d-point DFTs,
    each timed into plan->times -- against the oracle and numpy, and back."""
    fwd, rt, A = run_split(C, P, shape, grid, B=B, dbl=dbl, phased=True, length=length)
    tol = TOL64 if dbl else TOL32
    assert fwd < tol, f"forward rel L2 {fwd:.3e} (A={A})"
    assert rt < tol, f"round trip {rt:.3e}"


@pytest.mark.parametrize("P,shape,grid,B,dbl", [
    (4, (8192, 8), (2, 2), [128, 8], True),      # 64 x 128 along axis 0: longer than a pass tile
    (2, (16384, 4), (2, 1), [128, 4], False),    # complex float: 128 x 128
])
def test_split_fft_axis_longer_than_a_tile(C, P, shape, grid, B, dbl):
    """An axis longer than one pass tile plans (in phase order only: the
    plan's own A x B split is what makes it reachable) and matches the
    oracle and numpy's DFT."""
    fwd, rt, A = run_split(C, P, shape, grid, B=B, dbl=dbl)
    tol = TOL64 if dbl else TOL32
    assert fwd < tol, f"forward rel L2 {fwd:.3e} (A={A})"
    assert rt < tol, f"round trip {rt:.3e}"


@pytest.mark.parametrize("dbl", [True, False])
@pytest.mark.parametrize("P,shape,grid", [(4, (16, 16, 16), (1, 2, 2)), (8, (32, 16, 32), (2, 2, 2)),
                                          (32, (64, 64, 64), (1, 4, 8)),
                                          (4, (34, 38, 16), (1, 2, 2))])   # primes 17, 19
def test_autotuned_plan_vs_oracle_and_dft(C, P, shape, grid, dbl):
    """Whatever autotuneParallelPlan[f] returns (split_fft.c:78-247 picks by
    timing; this library by a traffic model plus timing), its forward
    matches the oracle run with the plan's own factors and numpy's DFT, and
    its backward restores N * x."""
    fwd, rt, A = run_split(C, P, shape, grid, dbl=dbl, autotune=True)
    tol = TOL64 if dbl else TOL32
    assert fwd < tol, f"forward rel L2 {fwd:.3e} (A={A})"
    assert rt < tol, f"round trip {rt:.3e} (A={A})"


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("shape", [(64, 64, 64), (34, 19, 17)])
def test_pencil_axis_ffts_64(C, shape, dbl):
    """x-pencil -> (FFT x) -> y-pencil -> (FFT y) -> z-pencil -> (FFT z):
    three fused passes, result equals numpy's 3D DFT placed by the z-pencil
    decomposition; reshape placement checked through the oracle layout.
    (34, 19, 17): uneven blocks and axes of primes above 13 (direct-sum
    stages).  Complex float within the FP32 tolerance of the FP64 DFT."""
    P = 4
    es, ct = (16, np.complex128) if dbl else (8, np.complex64)
    comm = C.dm_world_create(P)

    def pencil(ax):
        it = iter((2, 2))
        return [("local", shape[d]) if d == ax else ("block", shape[d], next(it)) for d in range(3)]
    px, ox = make_pair(C, comm, P, pencil(0))
    py, oy = make_pair(C, comm, P, pencil(1))
    pz, oz = make_pair(C, comm, P, pencil(2))
    per = int(max(max(o.extents) for o in (ox, oy, oz)))
    b0 = dm.DistBuffer(comm, per * es)
    b1 = dm.DistBuffer(comm, per * es)
    rng = np.random.default_rng(3)
    glob = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    if not dbl:
        glob = glob.astype(np.complex64).astype(np.complex128)
    host = np.zeros((P, per), ct)
    for c in np.ndindex(*shape):
        r, off = ox.owner(c)
        host[r, off] = glob[c]
    b0.put_all(host)
    fx = C.dm_plan_axis_fft(px, py, 0, dbl, dm.FFTW_FORWARD)
    fy = C.dm_plan_axis_fft(py, pz, 1, dbl, dm.FFTW_FORWARD)
    fz = C.dm_plan_axis_fft(pz, pz, 2, dbl, dm.FFTW_FORWARD)
    assert fx and fy and fz
    assert C.dm_run_axis_fft(fx, b0.ptr, b1.ptr) == 0
    assert C.dm_run_axis_fft(fy, b1.ptr, b0.ptr) == 0
    assert C.dm_run_axis_fft(fz, b0.ptr, b1.ptr) == 0
    got = b1.get_all(per * es, ct).astype(np.complex128)
    ref = np.fft.fftn(glob)
    want = np.zeros_like(got)
    mask = np.zeros(got.shape, bool)
    for c in np.ndindex(*shape):
        r, off = oz.owner(c)
        want[r, off] = ref[c]
        mask[r, off] = True
    assert rel_l2(got[mask], want[mask]) < (TOL64 if dbl else TOL32)
    for f in (fx, fy, fz):
        C.dm_free_axis_fft(f)
    b0.free()
    b1.free()
