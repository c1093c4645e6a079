"""Pure-Python model of the reference's data placement (TEST INFRASTRUCTURE).

A second, independent statement of the rules the oracle (oracle/oracle.c)
and the product both follow, written as brute force over global coordinates
for small shapes.  It pins the oracle: the oracle restates the reference's
per-message pairing (reshape.c:326-449, typelist.c:74-119), this module the
global-coordinate pairing that pairing is meant to produce, so an index or
ordering mistake in either shows up as a mismatch.

Descriptors are the tuples used by the tests:
    ("block", L, procs[, factor])      block_decomp.c:6-51
    ("local", L)                       local_decomp.c:6-28
    ("single", L, rank, procs)         single_decomp.c:6-30
    ("rotate", shift, nested)          rotate_decomp.c:7-50
    ("interleave", nested, B)          interleave_decomp.c:6-56
    ("frev", A, B, nested)             factor_reverse_decomp.c:8-58
"""
import itertools

import numpy as np


def coord_size(d):
    k = d[0]
    if k in ("block", "local", "single"):
        return d[1]
    if k == "rotate":
        return coord_size(d[2])
    if k == "interleave":
        return coord_size(d[1]) * d[2]
    if k == "frev":
        return coord_size(d[3])
    raise ValueError(k)


def proc_size(d):
    k = d[0]
    if k == "block":
        return d[2]
    if k == "local":
        return 1
    if k == "single":
        return d[3]
    if k == "rotate":
        return proc_size(d[2])
    if k == "interleave":
        return proc_size(d[1])
    if k == "frev":
        return proc_size(d[3])
    raise ValueError(k)


def block_size(L, procs, factor=1):
    # block_decomp.c:36-45: blocks are whole multiples of `factor`
    per = L // factor
    return factor * ((per + procs - 1) // procs)


def rank_off(d, g):
    """(rank position along the processor axis, local index) of coordinate g."""
    k = d[0]
    if k == "block":
        bs = block_size(d[1], d[2], d[3] if len(d) > 3 else 1)
        return g // bs, g % bs
    if k == "local":
        return 0, g
    if k == "single":
        return d[2], g
    if k == "rotate":
        L = coord_size(d[2])
        return rank_off(d[2], (g + L + d[1]) % L)
    if k == "interleave":
        A = coord_size(d[1])
        y, z = g % A, g // A
        r, o = rank_off(d[1], y)
        return r, z + d[2] * o
    if k == "frev":
        A, B = d[1], d[2]
        return rank_off(d[3], g // A + B * (g % A))
    raise ValueError(k)


class Layout:
    """A decomposition over P ranks: per axis a descriptor and a rank stride,
    memory strides packed (reshape.c:218-241) or given (same on every rank)."""

    def __init__(self, P, descs, rank_strides=None, strides=None):
        self.P = P
        self.descs = descs
        self.n = len(descs)
        if rank_strides is None:
            rank_strides, r = [0] * self.n, 1
            for i in reversed(range(self.n)):
                rank_strides[i] = r
                r *= proc_size(descs[i])
        self.rs = rank_strides
        self.L = [coord_size(d) for d in descs]
        self.proc_used = 1 + sum(rs * (proc_size(d) - 1) for d, rs in zip(descs, self.rs))
        self.tab = [[rank_off(d, g) for g in range(coord_size(d))] for d in descs]
        # per rank, per axis: max local index (0 when the rank holds nothing,
        # reshape.c:72-86)
        self.maxl = np.zeros((P, self.n), dtype=np.int64)
        for r in range(P):
            for i, d in enumerate(descs):
                pos = (r // self.rs[i]) % proc_size(d)
                offs = [o for q, o in self.tab[i] if q == pos]
                self.maxl[r, i] = max(offs) if offs else 0
        self.m = np.zeros((P, self.n), dtype=np.int64)
        for r in range(P):
            if strides is None:
                s = 1
                for i in reversed(range(self.n)):
                    self.m[r, i] = s
                    s *= int(self.maxl[r, i]) + 1
            else:
                self.m[r, :] = strides
        self.extent = [1 + int(sum(self.m[r] * self.maxl[r])) for r in range(P)]

    def owner(self, g):
        """(rank, element offset) of global coordinate g, or (None, None)."""
        R = sum(self.tab[i][g[i]][0] * self.rs[i] for i in range(self.n))
        if R >= self.proc_used:
            return None, None
        return R, int(sum(self.m[R, i] * self.tab[i][g[i]][1] for i in range(self.n)))


def reshape(src, dst, src_buf, dst_buf):
    """Global-coordinate pairing: every coordinate owned in both layouts is
    copied from its source slot to its destination slot (element arrays
    [P, slot]); everything else in dst_buf is left as it was."""
    for g in itertools.product(*[range(n) for n in src.L]):
        S, so = src.owner(g)
        D, do = dst.owner(g)
        if S is not None and D is not None:
            dst_buf[D, do] = src_buf[S, so]


def processor_grid(P, tasks):
    """makeProcessorGrid (processor_grid.c:24-97) by flat enumeration in the
    same order: smallest max-over-axes load, ties to the smaller
    (max ceil) / (min floor) asymmetry, first found wins."""
    best_load, best_asym, best = int(np.prod(tasks)), 10000.0, [1] * len(tasks)
    ranges = [range(1, min(P, t) + 1) for t in tasks]
    for grid in itertools.product(*ranges):
        if int(np.prod(grid)) > P:
            continue
        load = 1
        for t, g in zip(tasks, grid):
            load *= -(-t // g)
        hi = max(max(-(-t // g) for t, g in zip(tasks, grid)), 0)
        lo = min(min(t // g for t, g in zip(tasks, grid)), max(tasks))
        asym = hi / lo
        if load < best_load or (load == best_load and asym < best_asym):
            best_load, best_asym, best = load, asym, list(grid)
    return best


def split_position(k, A, B):
    """B-layout position of frequency k along one axis of a two-stage split
    L = A*B: k = k1 + B*k2 sits at k2 + A*k1 (four-step identity)."""
    return k // B + A * (k % B)


def twiddle(y, l, L):
    """exp(-2 pi i y l / L) with the exact quarter-turn values of
    split_fft.c:373-393."""
    q = (y * l) % L
    if q == 0:
        return complex(1.0, 0.0)
    if 4 * q == L:
        return complex(0.0, -1.0)
    if 2 * q == L:
        return complex(-1.0, 0.0)
    if 4 * q == 3 * L:
        return complex(-0.0, 1.0)
    angle = -(2.0 * np.pi) * ((y * l) / L)
    return complex(np.cos(angle), np.sin(angle))
