"""Oracle: CPU restatement of the reference's hot path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker (or as the timed CPU baseline).  The
product (decomp_mpi_amd) never imports it.

Descriptors are node lists mirroring the reference's DimDecomp strategies:
    Block(len, procs, factor=1)   block_decomp.c:6-51
    Local(len)                    local_decomp.c:6-28
    Single(len, rank, procs)      single_decomp.c:6-30
    Rotate(shift, nested)         rotate_decomp.c:7-50
    Interleave(nested, B)         interleave_decomp.c:6-56
    FactorReverse(A, B, nested)   factor_reverse_decomp.c:8-54
An axis is (descriptor, rank_stride); a decomposition is a list of axes plus
either packed strides (reshape.c:218-241) or explicit per-axis strides
(makeAxisDecomp's `stride`, the same on every rank).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BLOCK, LOCAL, SINGLE, ROTATE, INTERLEAVE, FREV = 1, 2, 3, 4, 5, 6


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.or_split_extent.restype = ctypes.c_long
        _lib.or_message.restype = ctypes.c_long
    return _lib


class Spec:
    """A DimDecomp node list (root first)."""

    def __init__(self, nodes):
        self.nodes = nodes

    def array(self):
        return np.array([v for n in self.nodes for v in n], dtype=np.int32)


def _nest(head, nested):
    # renumber nested nodes behind the head node
    shift = 1
    rows = [list(head[:4]) + [shift]]
    for n in nested.nodes:
        n = list(n)
        if n[4] >= 0:
            n[4] += shift
        rows.append(n)
    return Spec(rows)


def Block(length, procs, factor=1):
    return Spec([[BLOCK, length, factor, procs, -1]])


def Local(length):
    return Spec([[LOCAL, length, 0, 0, -1]])


def Single(length, rank, procs):
    return Spec([[SINGLE, length, rank, procs, -1]])


def Rotate(shift, nested):
    return _nest([ROTATE, shift, 0, 0], nested)


def Interleave(nested, B):
    return _nest([INTERLEAVE, B, 0, 0], nested)


def FactorReverse(A, B, nested):
    return _nest([FREV, A, B, 0], nested)


def _ip(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rank_offset(spec, g):
    arr = spec.array()
    return lib().or_rank(_ip(arr), ctypes.c_int(g)), lib().or_offset(_ip(arr), ctypes.c_int(g))


def coord_size(spec):
    return lib().or_coord_size(_ip(spec.array()), 0)


def proc_size(spec):
    return lib().or_proc_size(_ip(spec.array()), 0)


class Decomp:
    """axes: list of (Spec, rank_stride); strides None -> packed layout."""

    def __init__(self, P, axes, strides=None):
        self.P = P
        self.axes = axes
        self.ndim = len(axes)
        arrs = [s.array() for s, _ in axes]
        offs, at = [], 0
        for a in arrs:
            offs.append(at)
            at += len(a)
        self.specs = np.concatenate(arrs).astype(np.int32)
        self.offs = np.array(offs, dtype=np.int32)
        self.rs = np.array([r for _, r in axes], dtype=np.int32)
        self.packed = 1 if strides is None else 0
        self.expl = np.array(strides if strides is not None else [0] * self.ndim, dtype=np.int64)
        st = np.zeros(P * self.ndim, dtype=np.int64)
        ex = np.zeros(P, dtype=np.int64)
        self.proc_used = lib().or_decomp_info(P, self.ndim, _ip(self.specs), _ip(self.offs), _ip(self.rs),
                                              self.packed, _ip(self.expl), _ip(st), _ip(ex))
        self.strides = st.reshape(P, self.ndim)
        self.extents = ex
        self.L = [coord_size(s) for s, _ in axes]

    def args(self):
        return [_ip(self.specs), _ip(self.offs), _ip(self.rs), ctypes.c_int(self.packed), _ip(self.expl)]

    def owner(self, coord):
        """(rank, local offset) of a global coordinate tuple (or (None, None))."""
        r, off = 0, []
        for i, (spec, rs) in enumerate(self.axes):
            q, o = rank_offset(spec, coord[i])
            r += q * rs
            off.append(o)
        if r >= self.proc_used:
            return None, None
        return r, int(sum(self.strides[r, i] * off[i] for i in range(self.ndim)))


def reshape(src, dst, src_buf, dst_buf, elem):
    """runReshapePlan over all ranks: src_buf / dst_buf are uint8 arrays of
    P slots (shape [P, slot_bytes]); dst_buf is updated in place."""
    assert src_buf.flags.c_contiguous and dst_buf.flags.c_contiguous
    rc = lib().or_reshape(src.P, src.ndim, *src.args(), *dst.args(), ctypes.c_int(elem), _ip(src_buf),
                          ctypes.c_long(src_buf.shape[1]), _ip(dst_buf), ctypes.c_long(dst_buf.shape[1]))
    if rc != 0:
        raise RuntimeError(f"oracle reshape: unmatched messages (code {rc})")


def message(src, dst, me, peer, send):
    """Element offsets of rank me's message to (send) or from rank peer,
    in the order the reference's nested datatype lists them, or None."""
    cap = int(max(src.extents.max(), dst.extents.max())) + 1
    out = np.zeros(cap, dtype=np.int64)
    n = lib().or_message(src.P, src.ndim, *src.args(), *dst.args(), ctypes.c_int(me), ctypes.c_int(peer),
                         ctypes.c_int(1 if send else 0), _ip(out), ctypes.c_long(cap))
    return None if n < 0 else out[:n].copy()


def processor_grid(P, tasks):
    t = np.array(tasks, dtype=np.int32)
    out = np.zeros(len(tasks), dtype=np.int32)
    lib().or_processor_grid(P, len(tasks), _ip(t), _ip(out))
    return [int(v) for v in out]


def split_extent(P, A, B, ga, gb):
    n = len(A)
    a = [np.array(v, dtype=np.int32) for v in (A, B, ga, gb)]
    return int(lib().or_split_extent(P, n, *[_ip(x) for x in a]))


def split_fft(start, A, B, ga, gb, dat, work, extent, direction):
    """runSplitPlan restated: dat [P, dat_slot] complex128, work [P, work_slot]
    complex128 (updated in place).  direction -1 forward, +1 backward."""
    n = start.ndim
    a = [np.array(v, dtype=np.int32) for v in (A, B, ga, gb)]
    rc = lib().or_split_fft(start.P, n, *start.args(), *[_ip(x) for x in a], ctypes.c_long(extent), _ip(dat),
                            ctypes.c_long(dat.shape[1]), _ip(work), ctypes.c_long(work.shape[1]),
                            ctypes.c_int(direction))
    if rc != 0:
        raise RuntimeError(f"oracle split FFT failed (code {rc})")


def twiddle_table(P, A_axes_decomp, rank, A, L, axis):
    out = np.zeros(2 * (max(L) + 1), dtype=np.float64)
    d = A_axes_decomp
    n = lib().or_twiddle_table(P, d.ndim, _ip(d.specs), _ip(d.offs), _ip(d.rs), rank,
                               _ip(np.array(A, dtype=np.int32)), _ip(np.array(L, dtype=np.int32)), axis, _ip(out))
    return out[: 2 * n].view(np.complex128)


def decomp_scale(P, A_decomp, rank, A, L, data, sign):
    d = A_decomp
    lib().or_decomp_scale(P, d.ndim, _ip(d.specs), _ip(d.offs), _ip(d.rs), rank,
                          _ip(np.array(A, dtype=np.int32)), _ip(np.array(L, dtype=np.int32)), _ip(data), sign)


def split_decomps(P, A, B, ga, gb):
    """The A (interleave) and B (blocked) decompositions of a split plan
    (split_fft.c:36-77), packed, rank strides in C order."""
    n = len(A)
    rs_a, rs_b, ra, rb = [0] * n, [0] * n, 1, 1
    for i in reversed(range(n)):
        rs_a[i], rs_b[i] = ra, rb
        ra *= ga[i]
        rb *= gb[i]
    Ad = Decomp(P, [(Interleave(Block(A[i], ga[i]), B[i]), rs_a[i]) for i in range(n)])
    Bd = Decomp(P, [(Block(A[i] * B[i], gb[i], A[i]), rs_b[i]) for i in range(n)])
    return Ad, Bd
