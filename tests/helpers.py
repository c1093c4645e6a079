"""Shared test helpers: one description of a decomposition, built both as a
product object (through the C ABI) and as an oracle object."""
import ctypes

import numpy as np

import decomp_mpi_amd as dm
import oracle

ELEM_TYPE = {1: "byte", 4: "int", 8: "long", 16: "c_double_complex"}


def prod_dim(C, d):
    kind = d[0]
    if kind == "block":
        return C.makeBlockDecompWithFactor(d[1], d[3] if len(d) > 3 else 1, d[2])
    if kind == "local":
        return C.makeLocalDecomp(d[1])
    if kind == "single":
        return C.makeSingleDecomp(d[1], d[2], d[3])
    if kind == "rotate":
        nested = prod_dim(C, d[2])
        r = C.makeRotatedDecomp(d[1], nested)
        C.freeDimDecomp(nested)
        return r
    if kind == "interleave":
        nested = prod_dim(C, d[1])
        r = C.makeInterleaveDecomp(nested, d[2])
        C.freeDimDecomp(nested)
        return r
    if kind == "frev":
        nested = prod_dim(C, d[3])
        r = C.makeFactorReverseDecomp(d[1], d[2], nested)
        C.freeDimDecomp(nested)
        return r
    raise ValueError(kind)


def oracle_spec(d):
    kind = d[0]
    if kind == "block":
        return oracle.Block(d[1], d[2], d[3] if len(d) > 3 else 1)
    if kind == "local":
        return oracle.Local(d[1])
    if kind == "single":
        return oracle.Single(d[1], d[2], d[3])
    if kind == "rotate":
        return oracle.Rotate(d[1], oracle_spec(d[2]))
    if kind == "interleave":
        return oracle.Interleave(oracle_spec(d[1]), d[2])
    if kind == "frev":
        return oracle.FactorReverse(d[1], d[2], oracle_spec(d[3]))
    raise ValueError(kind)


def proc_size(d):
    return oracle.proc_size(oracle_spec(d))


def packed_rank_strides(descs):
    rs, r = [0] * len(descs), 1
    for i in reversed(range(len(descs))):
        rs[i] = r
        r *= proc_size(descs[i])
    return rs


def make_pair(C, comm, P, descs, rank_strides=None, strides=None):
    """strides None: makePackedDecomp (rank strides packed too); else
    makeAxisDecomp with the given memory strides and rank strides."""
    if strides is None:
        dims = [prod_dim(C, d) for d in descs]
        pd = C.makePackedDecomp(comm, len(descs), dm.ptrs(dims))
        rs = packed_rank_strides(descs)
        od = oracle.Decomp(P, [(oracle_spec(d), rs[i]) for i, d in enumerate(descs)])
    else:
        rs = rank_strides or packed_rank_strides(descs)
        axes = [C.makeAxisDecomp(comm, prod_dim(C, d), strides[i], rs[i]) for i, d in enumerate(descs)]
        pd = C.makeDecomp(len(descs), dm.ptrs(axes))
        od = oracle.Decomp(P, [(oracle_spec(d), rs[i]) for i, d in enumerate(descs)], strides=strides)
    assert pd, "product decomposition failed"
    return pd, od


def elem_type(C, elem):
    if elem in ELEM_TYPE:
        return dm.mpi_type(ELEM_TYPE[elem])
    assert elem % 8 == 0
    t = ctypes.c_void_p()
    C.dm_type_contiguous(elem // 8, dm.mpi_type("double"), ctypes.byref(t))
    return t.value


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))



# result placement helpers live in the package (bench.py uses them too)
from decomp_mpi_amd.layout import MAPPER, flat_index, layout_tables  # noqa: E402,F401
