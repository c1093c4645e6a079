"""Batched local FFTs (dm_batch.h) on the device against numpy's FP64 DFT:
uneven 3D shapes (including radix 3, 5, 7, 11 and 13 lengths), 1D and 2D
batches, both directions, both precisions; HIP-graph replays checked one by
one against the eager run; the autotuner's schedule stays correct.

Tolerances: complex double 1e-12 relative L2 per transform, complex float
1e-6 per transform against the FP64 DFT."""
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


class Batch:
    """count transforms carved out of one device buffer (256-byte aligned)."""

    def __init__(self, C, shapes, dbl, direction, seed=0):
        self.C, self.dbl = C, dbl
        self.shapes = [tuple(s) for s in shapes]
        self.ndim = len(self.shapes[0])
        self.ct = np.complex128 if dbl else np.complex64
        es = 16 if dbl else 8
        self.offs, at = [], 0
        for s in self.shapes:
            self.offs.append(at)
            at += (int(np.prod(s)) * es + 255) // 256 * 256
        self.comm = C.dm_world_create(1)
        self.buf = dm.DistBuffer(self.comm, at)
        rng = np.random.default_rng(seed)
        self.x = [(rng.uniform(-1, 1, s) + 1j * rng.uniform(-1, 1, s)).astype(self.ct) for s in self.shapes]
        flat = [v for s in self.shapes for v in s]
        self.plan = C.dm_batch_plan(len(shapes), self.ndim, dm.ints(flat), dbl, direction)
        assert self.plan
        ptrs = (ctypes.c_void_p * len(shapes))(*[self.buf.ptr + o for o in self.offs])
        assert C.dm_batch_bind(self.plan, ptrs) == 0
        self.direction = direction

    def upload(self):
        for x, o in zip(self.x, self.offs):
            self.buf.put(0, np.ascontiguousarray(x), o)

    def download(self):
        es = 16 if self.dbl else 8
        return [self.buf.get(0, int(np.prod(s)) * es, self.ct, o).reshape(s) for s, o in zip(self.shapes, self.offs)]

    def want(self):
        out = []
        for x in self.x:
            x64 = x.astype(np.complex128)
            out.append(np.fft.fftn(x64) if self.direction == -1 else np.fft.ifftn(x64) * x64.size)
        return out

    def free(self):
        self.C.dm_batch_free(self.plan)
        self.buf.free()
        self.C.dm_world_free(self.comm)


def rel(a, b):
    return float(np.linalg.norm(a.astype(np.complex128) - b) / np.linalg.norm(b))


SHAPES3 = [(32, 48, 64), (64, 96, 128), (128, 32, 48), (96, 96, 32), (48, 64, 128), (32, 32, 32),
           (80, 112, 48), (176, 32, 16), (208, 16, 32), (16, 240, 32)]
# axes with primes above 13 (direct-sum stages), beside ordinary ones and
# axes of length 1
SHAPES_PRIME = [(17, 34, 16), (19, 32, 23), (64, 17, 48), (32, 32, 32), (127, 2, 8), (16, 16, 289),
                (1, 17, 32), (8, 1, 1), (4, 1, 64)]


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("direction", [-1, 1])
def test_batch_3d_matches_numpy(C, dbl, direction):
    b = Batch(C, SHAPES3, dbl, direction, seed=3)
    b.upload()
    assert C.dm_batch_run(b.plan, None) == 0
    got, want = b.download(), b.want()
    errs = [rel(g, w) for g, w in zip(got, want)]
    b.free()
    assert max(errs) < TOL[dbl], errs


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("direction", [-1, 1])
@pytest.mark.parametrize("tile", [0, 2])
def test_batch_prime_axes(C, dbl, direction, tile):
    """Batches whose axes include primes above 13 (the direct-sum kernel
    for those launches), in the full and plane schedules (planes are not
    formed for such axes), against numpy; and the no-alloc entry point on
    the same inputs bit for bit equal to the full schedule."""
    b = Batch(C, SHAPES_PRIME, dbl, direction, seed=17)
    assert C.dm_batch_set_schedule(b.plan, tile, 0) == 0
    b.upload()
    assert C.dm_batch_run(b.plan, None) == 0
    got = b.download()
    errs = [rel(g, w) for g, w in zip(got, b.want())]
    if tile == 0:
        b.upload()
        flat = dm.ints([v for s in b.shapes for v in s])
        ptrs = (ctypes.c_void_p * len(b.shapes))(*[b.buf.ptr + o for o in b.offs])
        assert C.dm_batch_run_noalloc(len(b.shapes), b.ndim, flat, dbl, direction, ptrs, None) == 0
        C.dm_synchronize(b.comm)
        again = b.download()
        assert all(np.array_equal(g, a) for g, a in zip(got, again))
    b.free()
    assert max(errs) < TOL[dbl], errs


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("shapes", [[(512,), (96,), (1000,), (4096,)], [(64, 48), (128, 128), (44, 52)],
                                    [(256, 128), (48, 256)]])
def test_batch_1d_2d(C, shapes, dbl):
    """1D and 2D batches; lines longer than 256 points (twiddles read from
    device memory for the whole launch) and launches whose lines all fit
    the twiddles in LDS."""
    b = Batch(C, shapes, dbl, -1, seed=5)
    b.upload()
    assert C.dm_batch_run(b.plan, None) == 0
    errs = [rel(g, w) for g, w in zip(b.download(), b.want())]
    b.free()
    assert max(errs) < TOL[dbl], errs


def test_batch_graph_replays_match_eager(C):
    """Each of several graph replays (with eager uploads and syncs between
    them) equals the eager run bit for bit."""
    shapes = [tuple(int(v) for v in s) for s in np.random.default_rng(9).choice([32, 48, 64, 96], (12, 3))]
    b = Batch(C, shapes, 0, -1, seed=7)
    b.upload()
    assert C.dm_batch_run(b.plan, None) == 0
    eager = b.download()
    for _ in range(3):
        b.upload()
        assert C.dm_batch_graph_launch(b.plan, None) == 0
        C.dm_synchronize(b.comm)
        got = b.download()
        assert all(np.array_equal(g, e) for g, e in zip(got, eager))
    errs = [rel(g, w) for g, w in zip(eager, b.want())]
    b.free()
    assert max(errs) < 1e-6


PLANE_CASES = [
    # 3D: planes of the last two axes where they fit (complex float at most
    # 8192 points, complex double 4096; the rest on line tiles), both plane
    # sizes classes
    SHAPES3,
    # 2D: the plane is the whole transform; (512, 16) has a side above 256,
    # (128, 128) is above either plane size
    [(64, 48), (128, 128), (44, 52), (512, 16), (16, 208), (96, 80)],
]


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("direction", [-1, 1])
@pytest.mark.parametrize("case", range(len(PLANE_CASES)))
def test_batch_planes_match_numpy(C, case, dbl, direction):
    """DM_BATCH_TILE_PLANES: the last two axes of each transform in one
    plane tile (rows' stages, then the columns' in the same LDS tile), line
    tiles for the rest; against numpy's FP64 DFT, and graph replays equal to
    the eager run."""
    b = Batch(C, PLANE_CASES[case], dbl, direction, seed=13)
    assert C.dm_batch_set_schedule(b.plan, dm.DM_BATCH_TILE_PLANES, 0) == 0
    tiles = (ctypes.c_longlong * b.ndim)()
    tile = ctypes.c_int()
    assert C.dm_batch_info(b.plan, tiles, None, ctypes.byref(tile), None) == 0
    assert tile.value == dm.DM_BATCH_TILE_PLANES
    print(f"planes: case {case} precision {dbl} direction {direction} tiles {list(tiles)}", flush=True)
    b.upload()
    assert C.dm_batch_run(b.plan, None) == 0
    eager = b.download()
    errs = [rel(g, w) for g, w in zip(eager, b.want())]
    b.upload()
    assert C.dm_batch_graph_launch(b.plan, None) == 0
    C.dm_synchronize(b.comm)
    replay = b.download()
    b.free()
    assert max(errs) < TOL[dbl], errs
    assert all(np.array_equal(g, e) for g, e in zip(replay, eager))


@pytest.mark.parametrize("dbl", [1, 0])
@pytest.mark.parametrize("tile", [0, 1, 2])
def test_batch_schedules_and_grids(C, tile, dbl):
    """Every schedule (full, half and plane tiles) in both grid modes (one
    workgroup per tile, persistent workgroups looping over the tile list)
    against numpy's FP64 DFT. Every other launch walks its tiles from the
    last transform to the first; the two grid modes run the same tiles
    through the same kernels, so their results are equal bit for bit."""
    shapes = SHAPES3 + [(17, 34, 16), (64, 17, 48)]
    b = Batch(C, shapes, dbl, -1, seed=21)
    out = []
    for grid in (0, 1):  # DM_BATCH_GRID_TILES, DM_BATCH_GRID_PERSISTENT
        assert C.dm_batch_set_schedule(b.plan, tile, 0) == 0
        assert C.dm_batch_set_grid(b.plan, grid) == 0
        assert C.dm_batch_grid(b.plan) == grid
        print(f"schedules: tile {tile} grid {grid} precision {dbl}", flush=True)
        b.upload()
        assert C.dm_batch_run(b.plan, None) == 0
        out.append(b.download())
    errs = [rel(g, w) for g, w in zip(out[0], b.want())]
    b.free()
    assert max(errs) < TOL[dbl], errs
    assert all(np.array_equal(x, y) for x, y in zip(out[0], out[1]))


def test_batch_autotune_keeps_results(C):
    shapes = [(64, 96, 128), (32, 48, 64), (96, 32, 128)]
    b = Batch(C, shapes, 0, -1, seed=11)
    ms = C.dm_batch_autotune(b.plan, None)
    assert ms > 0
    tile, order = ctypes.c_int(), ctypes.c_int()
    assert C.dm_batch_info(b.plan, None, None, ctypes.byref(tile), ctypes.byref(order)) == 0
    b.upload()
    assert C.dm_batch_run(b.plan, None) == 0
    errs = [rel(g, w) for g, w in zip(b.download(), b.want())]
    b.free()
    assert max(errs) < 1e-6
