"""Pins the oracle (oracle/oracle.c) before anything is compared with it.

Each check reaches the oracle's answer by a different road:
  * descriptors: brute force over every coordinate against tests/refmodel.py,
    plus identities between different descriptor kinds;
  * reshape: the reference's own test.c known answers (test.c:118-231) and a
    descriptor zoo, both placed by global-coordinate pairing
    (tests/golden/*.npz), against the oracle's per-message pairing;
  * processor grid: the values the reference's search yields for the
    benchmark shapes, and flat brute force;
  * twiddles: the closed form with exact quarter turns;
  * split FFT: numpy's DFT (pocketfft) with the four-step output placement,
    and the unnormalised round trip N * x.
"""
import itertools
import os

import numpy as np
import pytest

import oracle
import refmodel as rm
from helpers import oracle_spec

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


DESCS = [
    ("block", 10, 3), ("block", 12, 5), ("block", 64, 2), ("block", 24, 3, 4), ("block", 16, 4, 2),
    ("local", 7), ("single", 9, 2, 4),
    ("rotate", 3, ("block", 10, 4)), ("rotate", -2, ("local", 6)), ("rotate", 11, ("block", 8, 3)),
    ("interleave", ("block", 4, 2), 8), ("interleave", ("block", 16, 4), 32), ("interleave", ("local", 3), 5),
    ("frev", 2, 4, ("block", 8, 2)), ("frev", 16, 32, ("block", 512, 4, 16)), ("frev", 3, 4, ("local", 12)),
    ("rotate", 1, ("interleave", ("block", 6, 3), 2)),
    ("frev", 4, 3, ("rotate", 5, ("block", 12, 3))),
]


@pytest.mark.parametrize("d", DESCS, ids=str)
def test_descriptor_brute_force(d):
    spec = oracle_spec(d)
    assert oracle.coord_size(spec) == rm.coord_size(d)
    assert oracle.proc_size(spec) == rm.proc_size(d)
    for g in range(rm.coord_size(d)):
        assert oracle.rank_offset(spec, g) == rm.rank_off(d, g), (d, g)


@pytest.mark.parametrize("L,procs,factor", [(10, 3, 1), (64, 2, 1), (512, 8, 1), (24, 3, 4), (17, 4, 1)])
def test_block_is_contiguous_blocks(L, procs, factor):
    """Block placement: rank * blocksize + offset recovers g, offsets count
    up from 0 inside each rank, ranks never decrease."""
    spec = oracle.Block(L, procs, factor)
    pairs = [oracle.rank_offset(spec, g) for g in range(L)]
    bs = max(o for _, o in pairs) + 1
    assert all(r * bs + o == g for g, (r, o) in enumerate(pairs))
    assert all(pairs[g][0] <= pairs[g + 1][0] for g in range(L - 1))
    assert bs % factor == 0


@pytest.mark.parametrize("A,B", [(2, 4), (4, 2), (16, 32), (3, 5), (1, 8)])
def test_factor_reverse_identities(A, B):
    """FactorReverse(B, A) undoes FactorReverse(A, B); FactorReverse(A, B)
    over Local equals Interleave(Local(A), B) (two descriptor kinds, one
    permutation: the transpose of an A x B grid)."""
    L = A * B
    twice = oracle.FactorReverse(B, A, oracle.FactorReverse(A, B, oracle.Local(L)))
    fr = oracle.FactorReverse(A, B, oracle.Local(L))
    il = oracle.Interleave(oracle.Local(A), B)
    offs = [oracle.rank_offset(fr, g)[1] for g in range(L)]
    assert sorted(offs) == list(range(L))
    for g in range(L):
        assert oracle.rank_offset(twice, g) == (0, g)
        assert oracle.rank_offset(il, g) == oracle.rank_offset(fr, g)


def test_rotate_composes():
    for a, b in [(1, 2), (3, -5), (7, 9)]:
        two = oracle.Rotate(a, oracle.Rotate(b, oracle.Block(12, 4)))
        one = oracle.Rotate(a + b, oracle.Block(12, 4))
        assert [oracle.rank_offset(two, g) for g in range(12)] == [oracle.rank_offset(one, g) for g in range(12)]


def tc_layout(npe, which):
    S = 10
    if which == "orig":
        return oracle.Decomp(npe, [(oracle.Single(S, 0, 1), 1), (oracle.Single(S, 0, npe), 1)], strides=[1, S])
    if which == "new":
        return oracle.Decomp(npe, [(oracle.Rotate(2, oracle.Local(S)), 1),
                                   (oracle.Rotate(1, oracle.Block(S, npe)), 1)], strides=[1, S])
    return oracle.Decomp(npe, [(oracle.Rotate(3, oracle.Block(S, 2)), 1),
                               (oracle.Rotate(3, oracle.Block(S, npe // 2)), 2)], strides=[1, S // 2])


@pytest.mark.parametrize("npe", [1, 2, 3, 4, 6])
def test_reshape_reference_test_c(npe):
    """test.c:118-231 known answer: after each reshape every local (j, i) of
    the destination holds j + 10 i at getOffset, and no other slot changes
    (test.c:16-45); the whole buffer equals the golden placement."""
    f = gold("reshape_testc.npz")
    pairs = [("orig", "new")] + ([("orig", "d2"), ("new", "d2")] if npe > 1 and npe % 2 == 0 else [])
    for s, d in pairs:
        key = f"npe{npe}_{s}_{d}"
        src, init, want = f[key + "_src"], f[key + "_init"], f[key + "_want"]
        so, do = tc_layout(npe, s), tc_layout(npe, d)
        got = init.copy()
        oracle.reshape(so, do, np.ascontiguousarray(src).view(np.uint8), got.view(np.uint8), 4)
        assert np.array_equal(got, want), key
        touched = np.zeros_like(got, bool)
        for i in range(10):
            for j in range(10):
                r, off = do.owner((j, i))
                assert got[r, off] == j + 10 * i
                touched[r, off] = True
        assert np.array_equal(got[~touched], init[~touched])


def test_reshape_descriptor_zoo():
    """Self-index arrays (every element its own global linear index)
    reshaped into every descriptor kind: the oracle's message pairing equals
    brute-force global pairing on every destination element."""
    import make_golden
    f = gold("reshape_zoo.npz")
    for i, (P, sd, dd, drs, dst) in enumerate(make_golden.ZOO):
        rs_s = rm.Layout(P, sd).rs
        so = oracle.Decomp(P, [(oracle_spec(x), rs_s[k]) for k, x in enumerate(sd)])
        rs_d = drs or rm.Layout(P, dd).rs
        do = oracle.Decomp(P, [(oracle_spec(x), rs_d[k]) for k, x in enumerate(dd)], strides=dst)
        src, want = f[f"zoo{i}_src"], f[f"zoo{i}_want"]
        got = np.full_like(want, -2)
        oracle.reshape(so, do, np.ascontiguousarray(src).view(np.uint8), got.view(np.uint8), 8)
        assert np.array_equal(got, want), f"zoo case {i}"


@pytest.mark.parametrize("P,tasks,want", [
    (4, (64, 64, 64), [1, 2, 2]), (32, (512, 512, 512), [2, 4, 4]), (32, (16, 16, 16), [2, 4, 4]),
    (32, (32, 32, 32), [2, 4, 4]), (1, (8, 8, 8), [1, 1, 1]), (4, (8, 8, 8), [1, 2, 2]),
])
def test_processor_grid_known(P, tasks, want):
    assert oracle.processor_grid(P, list(tasks)) == want
    assert rm.processor_grid(P, list(tasks)) == want


def test_processor_grid_brute_force():
    for P in (1, 2, 3, 4, 5, 6, 8, 12):
        for tasks in itertools.product((1, 2, 3, 4, 6, 8), repeat=3):
            assert oracle.processor_grid(P, list(tasks)) == rm.processor_grid(P, list(tasks)), (P, tasks)
        for tasks in itertools.product((2, 5, 7, 16), repeat=2):
            assert oracle.processor_grid(P, list(tasks)) == rm.processor_grid(P, list(tasks)), (P, tasks)


@pytest.mark.parametrize("L", [8, 12, 16, 64, 512])
@pytest.mark.parametrize("P", [1, 2])
def test_twiddle_closed_form(L, P):
    """The oracle's per-rank twiddle table (split_fft.c:348-420) against the
    closed form: exact quarter turns bit for bit, the rest within 2 ulp."""
    f = gold("twiddle.npz")
    A = int(f[f"tw{L}_A"])
    want = f[f"tw{L}"]
    B = L // A
    ga = P if A % P == 0 else 1
    d = oracle.Decomp(P, [(oracle.Interleave(oracle.Block(A, ga), B), 1)])
    for r in range(ga):
        tab = oracle.twiddle_table(P, d, r, [A], [L], 0)
        for g in range(L):
            q, local = d.owner((g,))
            if q != r:
                continue
            w = want[g]
            if (g % A) * (g // A) % L * 4 % L == 0:  # quarter turn
                assert tab[local].real == w.real and tab[local].imag == w.imag
                assert np.signbit(tab[local].real) == np.signbit(w.real)
            else:
                assert abs(tab[local] - w) <= 4.5e-16


def test_decomp_scale_applies_product_of_axes():
    """decomp_scale (dimscale.c:7-67) on ones leaves at each owned slot the
    product over axes of the closed-form twiddles (conjugated for sign +1)
    and leaves unused slots alone."""
    P, L, A = 2, [8, 12], [2, 3]
    B = [L[0] // A[0], L[1] // A[1]]
    d = oracle.Decomp(P, [(oracle.Interleave(oracle.Block(A[0], 2), B[0]), 1),
                          (oracle.Interleave(oracle.Block(A[1], 1), B[1]), 1)])
    for sign in (-1, 1):
        for r in range(P):
            data = np.ones(int(d.extents[r]) + 5, np.complex128)
            oracle.decomp_scale(P, d, r, A, L, data, sign)
            seen = np.zeros(len(data), bool)
            for g in itertools.product(range(L[0]), range(L[1])):
                q, off = d.owner(g)
                if q != r:
                    continue
                w = rm.twiddle(g[0] % A[0], g[0] // A[0], L[0]) * rm.twiddle(g[1] % A[1], g[1] // A[1], L[1])
                if sign == 1:
                    w = np.conj(w)
                assert abs(data[off] - w) < 1e-15
                seen[off] = True
            assert np.all(data[~seen] == 1)


def start_decomp(P, shape, grid):
    rs, r = [0] * len(shape), 1
    for i in reversed(range(len(shape))):
        rs[i] = r
        r *= grid[i]
    return oracle.Decomp(P, [(oracle.Block(n, g), rs[i]) for i, (n, g) in enumerate(zip(shape, grid))])


def run_oracle_split(P, x, A, start_grid):
    shape = x.shape
    n = len(shape)
    B = [shape[i] // A[i] for i in range(n)]
    ga, gb = oracle.processor_grid(P, A), oracle.processor_grid(P, B)
    start = start_decomp(P, shape, start_grid)
    ext = oracle.split_extent(P, A, B, ga, gb)
    dat = np.zeros((P, int(max(start.extents))), np.complex128)
    for g in np.ndindex(*shape):
        r, off = start.owner(g)
        dat[r, off] = x[g]
    work = np.zeros((P, 2 * ext + 16), np.complex128)
    oracle.split_fft(start, A, B, ga, gb, dat, work, ext, -1)
    Ad, Bd = oracle.split_decomps(P, A, B, ga, gb)
    two_stage = any(a > 1 for a in A)
    out_d = Bd if two_stage else Ad
    spec = np.zeros(shape, np.complex128)
    for k in np.ndindex(*shape):
        pos = tuple(rm.split_position(k[i], A[i], B[i]) if two_stage else k[i] for i in range(n))
        r, off = out_d.owner(pos)
        spec[k] = work[r, off]
    oracle.split_fft(start, A, B, ga, gb, dat, work, ext, +1)
    back = np.zeros(shape, np.complex128)
    for g in np.ndindex(*shape):
        r, off = start.owner(g)
        back[g] = dat[r, off]
    return spec, back, (out_d, work, A, B)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("P", [1, 4])
def test_split_fft_matches_numpy(case, P):
    f = gold("split_fft.npz")
    x, X, A = f[f"split{case}_x"], f[f"split{case}_X"], [int(a) for a in f[f"split{case}_A"]]
    grid = oracle.processor_grid(P, list(x.shape))
    spec, back, _ = run_oracle_split(P, x, A, grid)
    assert rel(spec, X) < 1e-13
    assert rel(back / x.size, x) < 1e-13


def test_split_fft_placement_is_checked():
    """The check above distinguishes the placement k -> k/B + A (k%B) from
    the reference's FactorReverse(A, B) argument order when A != B (the two
    agree only for A == B)."""
    f = gold("split_fft.npz")
    x, X = f["split0_x"], f["split0_X"]
    A = [int(a) for a in f["split0_A"]]
    spec, _, (out_d, work, A, B) = run_oracle_split(4, x, A, oracle.processor_grid(4, list(x.shape)))
    wrong = np.zeros_like(spec)
    for k in np.ndindex(*x.shape):
        pos = tuple(k[i] // A[i] + B[i] * (k[i] % A[i]) for i in range(3))
        r, off = out_d.owner(pos)
        wrong[k] = work[r, off]
    assert rel(spec, X) < 1e-13
    assert rel(wrong, X) > 0.5


def test_single_stage_split_is_plain_dft():
    """A = 1 on every axis: one stage, output in the A layout, natural order."""
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (8, 6, 4)) + 1j * rng.uniform(-1, 1, (8, 6, 4))
    spec, back, _ = run_oracle_split(2, x, [1, 1, 1], [1, 2, 1])
    assert rel(spec, np.fft.fftn(x)) < 1e-13
    assert rel(back / x.size, x) < 1e-13


@pytest.mark.parametrize("P,shape,grid,A", [(4, (16, 16, 16), (1, 2, 2), [16, 16, 2]),
                                            (8, (16, 16, 16), (2, 2, 2), [16, 4, 16]),
                                            (8, (16, 16, 16), (1, 2, 4), [16, 16, 1]),
                                            (8, (8, 12, 16), (2, 2, 2), [1, 3, 16])])
def test_split_fft_mixed_unit_factors_matches_numpy(P, shape, grid, A):
    """Plans with B_d = 1 on some axes beside B_d > 1 on others (the family
    B = (1, 1, 256) that the 512^3 autotuner picks): the oracle still equals
    numpy's DFT, placed by k -> k/B + A (k%B), and round-trips."""
    x = np.random.default_rng(5).uniform(-1, 1, shape) + 1j * np.random.default_rng(6).uniform(-1, 1, shape)
    spec, back, _ = run_oracle_split(P, x, A, list(grid))
    assert rel(spec, np.fft.fftn(x)) < 1e-13
    assert rel(back / x.size, x) < 1e-13
