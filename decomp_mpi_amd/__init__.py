"""decomp_mpi_amd: MI355X-native decomposition-driven reshape and split 3D FFT.

The product is the C-ABI library libdecomp_mpi_amd.so (headers in include/),
whose entry points are the reference's reshape.h / split_fft.h / *_decomp.h
API plus dm_* extensions.  This module loads it with ctypes and declares
every signature, so Python callers (tests, bench) use the same names with
the same argument meaning as C callers:

    import decomp_mpi_amd as dm
    C = dm.lib()
    comm = C.dm_world_create(4)
    a = C.makeAxisDecomp(comm, C.makeBlockDecomp(10, 4), 1, 1)

There is no fallback: when the shared library is missing, lib() raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdecomp_mpi_amd.so")

P = ctypes.c_void_p
I = ctypes.c_int
IP = ctypes.POINTER(ctypes.c_int)
D = ctypes.c_double
SZ = ctypes.c_size_t

FFTW_FORWARD = -1
FFTW_BACKWARD = 1
NO_FLAGS, PREFER_LARGE_FFT, PREFER_STAGE1, FORCE_USE_FIRST, RUN_PHASE_ORDER = 0, 1, 2, 4, 0x100
DM_TWIDDLES_COMPENSATED, DM_TWIDDLES_PLAIN = 0, 1
DM_BATCH_TILE_FULL, DM_BATCH_TILE_HALF, DM_BATCH_TILE_PLANES = 0, 1, 2
RESHAPE_1, FFT_1, TWIDDLE, RESHAPE_2, FFT_2, ALL, NUM_TIMERS = range(7)


# ---- struct mirrors (field names as in the reference headers) -------------

class DimDecompS(ctypes.Structure):
    _fields_ = [("rank_offset", P), ("type_offset", P), ("destroy", P), ("printer", P),
                ("cloner", P), ("type", ctypes.c_char_p), ("coord_size", I), ("proc_size", I)]


class IndexS(ctypes.Structure):
    _fields_ = [("size", I), ("count", I), ("offsets", IP), ("counts", IP)]


class AxisDecompS(ctypes.Structure):
    _fields_ = [("decomp", ctypes.POINTER(DimDecompS)), ("comm", P), ("memory_stride", I),
                ("rank_stride", I), ("my_rank_in_dimension", I), ("n_local", I), ("max_local", I),
                ("index", ctypes.POINTER(IndexS)), ("ext", P)]


class DecompS(ctypes.Structure):
    _fields_ = [("ndim", I), ("comm", P), ("proc_used", I),
                ("dims", ctypes.POINTER(ctypes.POINTER(AxisDecompS))), ("ext", P)]


class TypeEntryS(ctypes.Structure):
    _fields_ = [("rank", I), ("type", P)]


class TypeEntryListS(ctypes.Structure):
    _fields_ = [("size", I), ("count", I), ("list", ctypes.POINTER(TypeEntryS))]


class ReshapePlanS(ctypes.Structure):
    _fields_ = [("comm", P), ("src", P), ("dest", P), ("order", IP),
                ("sends", ctypes.POINTER(TypeEntryListS)), ("recvs", ctypes.POINTER(TypeEntryListS)),
                ("prog", P)]


class SplitPlanS(ctypes.Structure):
    _fields_ = [("comm", P), ("A_factors", IP), ("A", P), ("B", P), ("B_flipped", P), ("src", P),
                ("work", P), ("extent", I), ("plan_a", P), ("plan_b", P), ("plan_a_back", P),
                ("plan_b_back", P), ("start_to_A", P), ("A_to_start", P), ("twiddle", P), ("A_to_B", P),
                ("B_to_A", P), ("times", ctypes.POINTER(D)), ("ext", P)]


_SIGS = {
    # virtual world / types (dm_mpi.h)
    "dm_world_create": (P, [I]), "dm_comm_view": (P, [P, I]), "dm_world_free": (None, [P]),
    "dm_comm_rank": (I, [P, IP]), "dm_comm_size": (I, [P, IP]),
    "dm_type_contiguous": (I, [I, P, ctypes.POINTER(P)]), "dm_type_free": (I, [ctypes.POINTER(P)]),
    "dm_type_get_extent": (I, [P, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
    "dm_type_size": (I, [P, IP]),
    # runtime extensions (dm_runtime.h)
    "dm_malloc": (P, [P, SZ]), "dm_free": (I, [P]), "dm_register": (I, [P, P, SZ]),
    "dm_unregister": (I, [P]), "dm_slot_bytes": (SZ, [P]), "dm_put": (I, [P, I, P, SZ]),
    "dm_get": (I, [P, P, I, SZ]), "dm_fill": (I, [P, I, SZ]), "dm_device_available": (I, []),
    "dm_set_async": (I, [P, I]), "dm_synchronize": (I, [P]), "dm_get_stream": (P, [P]),
    "dm_set_stream": (I, [P, P]), "dm_set_error_mode": (None, [I]),
    "dm_last_error": (ctypes.c_char_p, []), "dm_clear_error": (None, []),
    "dm_plan_axis_fft": (P, [P, P, I, I, I]), "dm_run_axis_fft": (I, [P, P, P]),
    "dm_free_axis_fft": (None, [P]), "dm_copy_bench": (D, [P, P, P, SZ, I]),
    # index.h
    "makeIndex": (P, []), "freeIndex": (None, [P]), "addIndex": (None, [P, I]),
    # descriptors
    "makeBlockDecomp": (P, [I, I]), "makeBlockDecompWithFactor": (P, [I, I, I]),
    "makeLocalDecomp": (P, [I]), "makeSingleDecomp": (P, [I, I, I]),
    "makeRotatedDecomp": (P, [I, P]), "makeInterleaveDecomp": (P, [P, I]),
    "makeFactorReverseDecomp": (P, [I, I, P]), "makeFactorReverseAxisDecomp": (P, [I, I, P]),
    "block_rank_offset": (I, [P, I]), "block_type_offset": (I, [P, I]),
    "local_rank_offset": (I, [P, I]), "local_type_offset": (I, [P, I]),
    "single_rank_offset": (I, [P, I]), "single_type_offset": (I, [P, I]),
    "rotate_rank_offset": (I, [P, I]), "rotate_type_offset": (I, [P, I]),
    "interleave_rank_offset": (I, [P, I]), "interleave_type_offset": (I, [P, I]),
    "factor_reverse_rank_offset": (I, [P, I]), "factor_reverse_type_offset": (I, [P, I]),
    "freeDimDecomp": (None, [P]), "cloneDimDecomp": (P, [P]), "sizeDimDecomp": (I, [P]),
    # axis / decomp
    "makeAxisDecomp": (P, [P, P, I, I]), "freeAxisDecomp": (None, [P]), "isLocal": (I, [P, I]),
    "makeIndexList": (P, [P, P]), "makeGlobalIndex": (P, [P]), "freeIndexList": (None, [P]),
    "makeIOType": (None, [ctypes.POINTER(P), P, P]), "expandIOType": (I, [ctypes.POINTER(P), ctypes.c_long, P, P, I]),
    "dm_type_view": (ctypes.c_longlong, [P, ctypes.POINTER(ctypes.c_longlong), ctypes.c_longlong]),
    "dm_file_write": (I, [ctypes.c_char_p, P, P, P]), "dm_file_read": (I, [ctypes.c_char_p, P, P, P]),
    "freeIndex": (None, [P]),
    "makeDecomp": (P, [I, P]), "makePackedDecomp": (P, [P, I, P]), "freeDecomp": (None, [P]),
    "extent": (I, [P]), "compatible": (I, [P, P]), "equivalent": (I, [P, P]), "myRank": (I, [P]),
    "isCoordLocal": (I, [P, I, I]), "areCoordsLocal": (I, [P, IP]), "getOffset": (I, [P, IP]),
    # reshape
    "makeReshapePlan": (P, [P, P, P, P]), "freeReshapePlan": (None, [P]),
    "runReshapePlan": (None, [P, P, P, P]), "checkReshapePlanBounds": (I, [P, I, I, I]),
    # split FFT
    "buildParallelPlan": (P, [I, I, I, P, I, IP, IP, IP, IP, IP, P, I, P]),
    "proposeParallelPlan": (P, [I, P, IP, P, I, P]),
    "autotuneParallelPlan": (P, [P, P, I, P]), "makeSplitPlan": (P, [I, P, P, I, P]),
    "makeParallelPlan": (P, [P, IP, P, I, P]),
    "buildParallelPlanf": (P, [I, I, I, P, I, IP, IP, IP, IP, IP, P, I, P]),
    "proposeParallelPlanf": (P, [I, P, IP, P, I, P]),
    "autotuneParallelPlanf": (P, [P, P, I, P]), "makeSplitPlanf": (P, [I, P, P, I, P]),
    "makeParallelPlanf": (P, [P, IP, P, I, P]),
    "getTransformed": (P, [P]), "runSplitPlan": (None, [P, I]), "freeSplitPlan": (None, [P]),
    "dm_split_set_phase_order": (I, [P, I]), "dm_split_phase_order": (I, [P]),
    "dm_split_set_twiddles": (I, [P, I]), "dm_split_twiddles": (I, [P]),
    "dm_split_set_output": (I, [P, P]),
    "makeInterleavedDecomp": (P, [I, P, IP, IP, IP]), "makeBlockedDecomp": (P, [I, P, IP, IP, IP]),
    "gcd": (ctypes.c_uint, [ctypes.c_uint, ctypes.c_uint]), "decomp_scale": (None, [P, P, P, I]),
    "makeProcessorGrid": (IP, [P, I, IP]), "visitData": (None, [P, P, P]),
    "timer": (D, []),
    # batched local transforms (dm_batch.h)
    "dm_batch_plan": (P, [I, I, IP, I, I]), "dm_batch_bind": (I, [P, ctypes.POINTER(P)]),
    "dm_batch_set_schedule": (I, [P, I, I]), "dm_batch_run": (I, [P, P]),
    "dm_batch_graph_launch": (I, [P, P]), "dm_batch_autotune": (D, [P, P]),
    "dm_batch_run_noalloc": (I, [I, I, IP, I, I, ctypes.POINTER(P), P]),
    "dm_batch_set_grid": (I, [P, I]), "dm_batch_grid": (I, [P]),
    "dm_batch_info": (I, [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), IP, IP]),
    "dm_batch_bytes_moved": (D, [P]),
    "dm_batch_bytes": (D, [P]), "dm_batch_free": (None, [P]),
    # request loop (dm_serve.h)
    "dm_server_create": (P, [SZ, I, I]), "dm_server_submit": (ctypes.c_longlong, [P, I, IP, I, I, P, P]),
    "dm_server_breakdown": (I, [P, ctypes.c_longlong, ctypes.POINTER(D)]),
    "dm_server_plan_stats": (I, [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(SZ)]),
    "dm_server_poll": (I, [P]), "dm_server_wait": (I, [P, ctypes.c_longlong]), "dm_server_drain": (I, [P]),
    "dm_server_times": (I, [P, ctypes.c_longlong, ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(D)]),
    "dm_server_stats": (I, [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                            ctypes.POINTER(SZ), ctypes.POINTER(SZ)]),
    "dm_server_destroy": (None, [P]), "dm_host_alloc": (P, [SZ]), "dm_host_free": (None, [P]),
}

# predefined element types (addresses of exported structs)
TYPE_NAMES = ["byte", "char", "int", "long", "long_long", "float", "double",
              "c_float_complex", "c_double_complex"]

_lib = None


def lib():
    """Load libdecomp_mpi_amd.so and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m decomp_mpi_amd.build` "
                           "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def mpi_type(name):
    """Address of a predefined element type (MPI_INT etc.)."""
    return ctypes.addressof(ctypes.c_char.in_dll(lib(), "dm_type_" + name))


def ints(vals):
    return (ctypes.c_int * len(vals))(*vals)


def ptrs(vals):
    return (ctypes.c_void_p * len(vals))(*vals)


def decomp(ptr):
    return ctypes.cast(ptr, ctypes.POINTER(DecompS)).contents


def axis(ptr):
    return ctypes.cast(ptr, ctypes.POINTER(AxisDecompS)).contents


def split_plan(ptr):
    return ctypes.cast(ptr, ctypes.POINTER(SplitPlanS)).contents


def reshape_plan(ptr):
    return ctypes.cast(ptr, ctypes.POINTER(ReshapePlanS)).contents


class DistBuffer:
    """A distributed device buffer: nranks slots of slot_bytes (dm_malloc)."""

    def __init__(self, comm, bytes_per_rank):
        C = lib()
        self.comm = comm
        self.ptr = C.dm_malloc(comm, bytes_per_rank)
        if not self.ptr:
            raise MemoryError(f"dm_malloc({bytes_per_rank}) failed")
        self.slot = C.dm_slot_bytes(self.ptr)
        n = ctypes.c_int()
        C.dm_comm_size(comm, ctypes.byref(n))
        self.nranks = n.value

    def put(self, rank, arr, offset_bytes=0):
        a = np.ascontiguousarray(arr)
        rc = lib().dm_put(self.ptr + offset_bytes, rank, a.ctypes.data, a.nbytes)
        if rc != 0:
            raise ValueError("dm_put failed")

    def get(self, rank, nbytes, dtype, offset_bytes=0):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        rc = lib().dm_get(out.ctypes.data, self.ptr + offset_bytes, rank, out.nbytes)
        if rc != 0:
            raise ValueError("dm_get failed")
        return out

    def put_all(self, arr2d, offset_bytes=0):
        for r in range(arr2d.shape[0]):
            self.put(r, arr2d[r], offset_bytes)

    def get_all(self, nbytes, dtype, offset_bytes=0):
        return np.stack([self.get(r, nbytes, dtype, offset_bytes) for r in range(self.nranks)])

    def fill(self, value, nbytes):
        lib().dm_fill(self.ptr, value, nbytes)

    def free(self):
        if self.ptr:
            lib().dm_free(self.ptr)
            self.ptr = None
