"""The on-disk format and the FFTW-MPI transposed-out comparator, on the
device through the C ABI.

* dm_file_write / dm_file_read: a distributed array to and from a file of
  raw native-endian elements in global row-major order, the file the
  reference's drivers write through makeIOType's MPI-IO views
  (reshape.c:494-521, split_plan_test.c:160-200).  Self-index arrays make
  the check exact: the file must be 0, 1, ..., N-1, and reading it into
  another decomposition must put every index at the slot its descriptor
  prescribes, padding untouched.
* fftw_test.c:260-336: the slab pipeline (2D FFT on axis-0 slabs, reshape
  to the FFTW_MPI_TRANSPOSED_OUT layout, FFT along axis 0) against a split
  plan whose output is set to that layout (dm_split_set_output); both
  written with dm_file_write, the two files agree with each other and with
  numpy's FP64 DFT in row-major order (relative L2 < 1e-12).
"""
import ctypes
import os

import numpy as np
import pytest

import decomp_mpi_amd as dm
from helpers import make_pair, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    C = dm.lib()
    if not C.dm_device_available():
        pytest.skip("no HIP device")
    return C


def placement(od):
    """(rank, offset) of every global coordinate of oracle decomposition od."""
    out = {}
    for g in np.ndindex(*od.L):
        out[g] = od.owner(g)
    return out


IO_LAYOUTS = [
    (6, [("block", 12, 2), ("rotate", 5, ("block", 10, 3)), ("local", 8)], None, None),
    (4, [("interleave", ("block", 4, 2), 4), ("block", 12, 2), ("local", 6)], None, None),
    (8, [("block", 8, 2), ("local", 6), ("frev", 2, 5, ("block", 10, 4))], None, None),
    (6, [("block", 8, 2), ("local", 6), ("block", 10, 3)], [1, 1, 2], [1, 8, 48]),
]


@pytest.mark.parametrize("case", range(len(IO_LAYOUTS)))
def test_file_write_then_read_into_another_layout(C, case, tmp_path):
    P, descs, rs, strides = IO_LAYOUTS[case]
    comm = C.dm_world_create(P)
    src_p, src_o = make_pair(C, comm, P, descs, rank_strides=rs, strides=strides)
    shape = src_o.L
    N = int(np.prod(shape))
    # another layout of the same shape: different owners and memory order
    dst_descs = [("block", shape[0], 3 if P % 3 == 0 else 2), ("local", shape[1]), ("block", shape[2], 2)]
    dst_p, dst_o = make_pair(C, comm, P, dst_descs)
    es = 8
    sslot, dslot = int(max(src_o.extents)), int(max(dst_o.extents))
    a = dm.DistBuffer(comm, sslot * es)
    b = dm.DistBuffer(comm, dslot * es)
    host = np.full((P, sslot), -1, np.int64)
    for g, (r, off) in placement(src_o).items():
        host[r, off] = np.ravel_multi_index(g, shape)
    a.put_all(host)
    b.fill(0xFF, dslot * es)
    path = str(tmp_path / f"io_{case}.dat")
    t = dm.mpi_type("long_long")
    print(f"file I/O case {case}: {shape} on {P} ranks", flush=True)
    assert C.dm_file_write(path.encode(), t, src_p, a.ptr) == 0
    assert os.path.getsize(path) == N * es
    assert np.array_equal(np.fromfile(path, dtype=np.int64), np.arange(N, dtype=np.int64))
    assert C.dm_file_read(path.encode(), t, dst_p, b.ptr) == 0
    got = b.get_all(dslot * es, np.int64)
    seen = np.zeros_like(got, bool)
    for g, (r, off) in placement(dst_o).items():
        assert got[r, off] == np.ravel_multi_index(g, shape)
        seen[r, off] = True
    assert np.all(got[~seen] == -1)
    a.free()
    b.free()
    C.freeDecomp(src_p)
    C.freeDecomp(dst_p)
    C.dm_world_free(comm)


def slab_layouts(C, comm, n, npe):
    """fftw_test.c:260-281: start = axis 0 in blocks (FFTW-MPI's input slab),
    end = axis 1 in blocks with local order [k1][k0][k2]
    (FFTW_MPI_TRANSPOSED_OUT)."""
    start = C.makeDecomp(3, dm.ptrs([C.makeAxisDecomp(comm, C.makeBlockDecomp(n, npe), n * n, 1),
                                     C.makeAxisDecomp(comm, C.makeLocalDecomp(n), n, 1),
                                     C.makeAxisDecomp(comm, C.makeLocalDecomp(n), 1, 1)]))
    end = C.makeDecomp(3, dm.ptrs([C.makeAxisDecomp(comm, C.makeLocalDecomp(n), n, 1),
                                   C.makeAxisDecomp(comm, C.makeBlockDecomp(n, npe), n * n, 1),
                                   C.makeAxisDecomp(comm, C.makeLocalDecomp(n), 1, 1)]))
    return start, end


@pytest.mark.parametrize("n,npe", [(32, 4), (48, 2)])
def test_transposed_out_comparator(C, n, npe, tmp_path):
    comm = C.dm_world_create(npe)
    start, end = slab_layouts(C, comm, n, npe)
    per = n * n * ((n + npe - 1) // npe)
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (n, n, n)) + 1j * rng.uniform(-1, 1, (n, n, n))
    blk = (n + npe - 1) // npe
    host = np.zeros((npe, per), np.complex128)
    for r in range(npe):
        sl = x[r * blk:(r + 1) * blk].ravel()
        host[r, :sl.size] = sl
    t = dm.mpi_type("c_double_complex")
    # (1) the reference's USE_RESHAPE pipeline: FFT along axes 2 and 1 of the
    # slab, reshape start -> end, FFT along axis 0
    b_in, b_tmp, b_out = (dm.DistBuffer(comm, per * 16) for _ in range(3))
    b_in.put_all(host)
    f2 = C.dm_plan_axis_fft(start, start, 2, 1, dm.FFTW_FORWARD)
    f1 = C.dm_plan_axis_fft(start, start, 1, 1, dm.FFTW_FORWARD)
    f0 = C.dm_plan_axis_fft(end, end, 0, 1, dm.FFTW_FORWARD)
    rp = C.makeReshapePlan(comm, t, start, end)
    assert f2 and f1 and f0 and rp
    assert C.dm_run_axis_fft(f2, b_in.ptr, b_tmp.ptr) == 0
    assert C.dm_run_axis_fft(f1, b_tmp.ptr, b_out.ptr) == 0
    C.runReshapePlan(comm, rp, b_out.ptr, b_in.ptr)
    assert C.dm_run_axis_fft(f0, b_in.ptr, b_out.ptr) == 0
    path1 = str(tmp_path / "reshape.dat")
    assert C.dm_file_write(path1.encode(), t, end, b_out.ptr) == 0
    # (2) a split plan writing its forward result in the transposed-out layout
    length = 4 * per + 64
    dat = dm.DistBuffer(comm, per * 16)
    work = dm.DistBuffer(comm, length * 16)
    dat.put_all(host)
    plan = C.makeSplitPlan(0, start, dat.ptr, length, work.ptr)
    assert plan
    assert C.dm_split_set_output(plan, end) == 0
    assert C.getTransformed(plan) == end
    C.runSplitPlan(plan, dm.FFTW_FORWARD)
    path2 = str(tmp_path / "split.dat")
    assert C.dm_file_write(path2.encode(), t, end, work.ptr) == 0
    X = np.fft.fftn(x)
    got1 = np.fromfile(path1, dtype=np.complex128).reshape(n, n, n)
    got2 = np.fromfile(path2, dtype=np.complex128).reshape(n, n, n)
    assert rel_l2(got1, X) < 1e-12
    assert rel_l2(got2, X) < 1e-12
    assert rel_l2(got2, got1) < 1e-12
    # backward from the transposed-out layout: N * x back in the slab
    C.runSplitPlan(plan, dm.FFTW_BACKWARD)
    back = dat.get_all(per * 16, np.complex128)
    assert rel_l2(back / n ** 3, host) < 1e-12
    for f in (f0, f1, f2):
        C.dm_free_axis_fft(f)
    C.freeReshapePlan(rp)
    C.freeSplitPlan(plan)
    for b in (b_in, b_tmp, b_out, dat, work):
        b.free()
    C.freeDecomp(start)
    C.freeDecomp(end)
    C.dm_world_free(comm)
