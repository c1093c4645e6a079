"""Host-side view of a decomposition's placement, for reading results.

flat_index(d, slot) gives, for every global coordinate, the element index
rank * slot + local offset at which decomposition d (a Decomp pointer of
this library, equal strides on every rank) stores it, or -1 where no rank
owns it.  Applied to getTransformed(plan) it orders a split FFT's output by
frequency."""
import ctypes

import numpy as np

from . import DimDecompS, decomp

MAPPER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int)


def layout_tables(dptr):
    """Per axis: (owner position table, local index table, rank stride,
    memory stride), and the decomposition's proc_used."""
    d = decomp(dptr)
    out = []
    for i in range(d.ndim):
        ax = d.dims[i].contents
        dd = ctypes.cast(ax.decomp, ctypes.POINTER(DimDecompS)).contents
        p = ctypes.cast(ax.decomp, ctypes.c_void_p).value
        ro, to = MAPPER(dd.rank_offset), MAPPER(dd.type_offset)
        rk = np.array([ro(p, g) for g in range(dd.coord_size)], np.int64)
        of = np.array([to(p, g) for g in range(dd.coord_size)], np.int64)
        out.append((rk, of, ax.rank_stride, ax.memory_stride))
    return out, d.proc_used


def flat_index(dptr, slot):
    tabs, proc_used = layout_tables(dptr)
    n = len(tabs)
    rank = np.zeros([1] * n, np.int64)
    off = np.zeros([1] * n, np.int64)
    for i, (rk, of, rs, ms) in enumerate(tabs):
        shape = [1] * n
        shape[i] = len(rk)
        rank = rank + (rk * rs).reshape(shape)
        off = off + (of * ms).reshape(shape)
    idx = rank * slot + off
    return np.where(rank < proc_used, idx, -1)
