"""Writes the golden fixtures in this folder (TEST INFRASTRUCTURE).

The reference ships no golden vectors for this path, so the fixtures are
computed here from definitions that do not go through the oracle:
  * reshape_*.npz: test.c's own cases (test.c:118-231) and a descriptor zoo,
    placed by brute-force global-coordinate pairing (tests/refmodel.py);
  * split_fft.npz: seeded inputs and numpy's FP64 DFT of each (pocketfft),
    the quantity runSplitPlan computes (split_fft.c:622-682);
  * twiddle.npz: exp(-2 pi i y l / L) per split_fft.c:373-393.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refmodel as rm  # noqa: E402

SIZE = 10


def testc_layouts(npe):
    """test.c:118-180: the 10x10 array on rank 0, the rotated local/block
    layout, and (even npe) the rotated 2 x npe/2 block layout."""
    orig = rm.Layout(npe, [("single", SIZE, 0, 1), ("single", SIZE, 0, npe)], [1, 1], [1, SIZE])
    new = rm.Layout(npe, [("rotate", 2, ("local", SIZE)), ("rotate", 1, ("block", SIZE, npe))], [1, 1], [1, SIZE])
    d2 = None
    if npe > 1 and npe % 2 == 0:
        d2 = rm.Layout(npe, [("rotate", 3, ("block", SIZE, 2)), ("rotate", 3, ("block", SIZE, npe // 2))],
                       [1, 2], [1, SIZE // 2])
    return orig, new, d2


def testc_cases():
    out = {}
    for npe in (1, 2, 3, 4, 6):
        orig, new, d2 = testc_layouts(npe)
        slot = SIZE * ((SIZE + npe - 1) // npe)
        src = np.zeros((npe, max(SIZE * SIZE, slot)), np.int32)
        src[0, :SIZE * SIZE] = np.arange(SIZE * SIZE)
        dest = np.repeat(-np.arange(npe, dtype=np.int32)[:, None], slot, axis=1)
        want1 = dest.copy()
        rm.reshape(orig, new, src, want1)
        out[f"npe{npe}_orig_new_src"] = src
        out[f"npe{npe}_orig_new_init"] = dest
        out[f"npe{npe}_orig_new_want"] = want1
        if d2 is not None:
            d2b = np.repeat(-np.arange(npe, dtype=np.int32)[:, None], slot, axis=1)
            want2 = d2b.copy()
            rm.reshape(orig, d2, src, want2)
            out[f"npe{npe}_orig_d2_src"] = src
            out[f"npe{npe}_orig_d2_init"] = d2b
            out[f"npe{npe}_orig_d2_want"] = want2
            d2c = np.repeat(-1000 * np.arange(npe, dtype=np.int32)[:, None], slot, axis=1)
            want3 = d2c.copy()
            rm.reshape(new, d2, want1, want3)
            out[f"npe{npe}_new_d2_src"] = want1
            out[f"npe{npe}_new_d2_init"] = d2c
            out[f"npe{npe}_new_d2_want"] = want3
    return out


# Descriptor zoo: (P, src descs, dst descs, dst rank strides, dst strides).
# Every descriptor kind, uneven blocks, idle ranks and an explicit-stride
# (axis-permuted) destination.
ZOO = [
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)], [("block", 8, 2), ("local", 8), ("block", 8, 2)],
     None, None),
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)],
     [("rotate", 3, ("block", 8, 2)), ("rotate", 5, ("local", 8)), ("block", 8, 2)], None, None),
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)],
     [("interleave", ("block", 2, 2), 4), ("block", 8, 2), ("local", 8)], None, None),
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)],
     [("frev", 2, 4, ("block", 8, 2)), ("local", 8), ("block", 8, 2)], None, None),
    (4, [("local", 8), ("block", 8, 2), ("block", 8, 2)],
     [("single", 8, 1, 2), ("block", 8, 2), ("local", 8)], None, None),
    (3, [("block", 8, 3), ("local", 8), ("local", 8)], [("block", 8, 3, 2), ("local", 8), ("local", 8)],
     None, None),
    (6, [("local", 12), ("block", 8, 2), ("block", 16, 3)],
     [("block", 12, 5), ("rotate", 7, ("local", 8)), ("interleave", ("block", 4, 1), 4)], None, None),
    (6, [("local", 12), ("block", 8, 2), ("block", 16, 3)],
     [("block", 12, 3), ("local", 8), ("interleave", ("block", 4, 2), 4)], [2, 1, 1], None),
    (6, [("local", 12), ("block", 8, 2), ("block", 16, 3)],
     [("block", 12, 2), ("local", 8), ("block", 16, 3)], [1, 1, 2], [1, 6, 48]),
]


def zoo_cases():
    out = {}
    for i, (P, sd, dd, drs, dst) in enumerate(ZOO):
        src = rm.Layout(P, sd)
        dstl = rm.Layout(P, dd, drs, dst)
        sb = np.full((P, max(src.extent)), -1, np.int64)
        shape = src.L
        for g in np.ndindex(*shape):
            r, o = src.owner(g)
            if r is not None:
                sb[r, o] = np.ravel_multi_index(g, shape)
        db = np.full((P, max(dstl.extent)), -2, np.int64)
        rm.reshape(src, dstl, sb, db)
        out[f"zoo{i}_src"] = sb
        out[f"zoo{i}_want"] = db
    return out


# Split FFT cases: (shape, per-axis A factor); B = L / A.
SPLIT = [((8, 8, 8), (2, 4, 2)), ((16, 16, 16), (4, 4, 4)), ((16, 8, 12), (4, 2, 3)), ((16, 16), (2, 8))]


def split_cases():
    out = {}
    for i, (shape, A) in enumerate(SPLIT):
        rng = np.random.default_rng(100 + i)
        x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
        out[f"split{i}_x"] = x
        out[f"split{i}_X"] = np.fft.fftn(x)
        out[f"split{i}_A"] = np.array(A, np.int32)
    return out


def twiddle_cases():
    out = {}
    for L in (8, 12, 16, 64, 512):
        A = max(a for a in range(1, L + 1) if L % a == 0 and a * a <= L)
        out[f"tw{L}_A"] = np.array(A)
        out[f"tw{L}"] = np.array([rm.twiddle(g % A, g // A, L) for g in range(L)], np.complex128)
    return out


def main():
    np.savez_compressed(os.path.join(HERE, "reshape_testc.npz"), **testc_cases())
    np.savez_compressed(os.path.join(HERE, "reshape_zoo.npz"), **zoo_cases())
    np.savez_compressed(os.path.join(HERE, "split_fft.npz"), **split_cases())
    np.savez_compressed(os.path.join(HERE, "twiddle.npz"), **twiddle_cases())


if __name__ == "__main__":
    main()
