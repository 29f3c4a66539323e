"""Pins the oracle (CPU, no GPU needed).

The oracle is checked against: the reference's known-answer vectors
(tests/golden/kat.json, produced by the reference's own SeqMdiscr), the
reference SeqMdiscr itself (oracle/_ref, when built), a closed form that
shares no code with it (numpy stable argsort by class), brute force over all
tiny class strings, and the defining invariants.  A mutated copy of the
oracle's result must fail these checks (planted mistakes, on the CPU).
"""
import itertools
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import fnv_fast

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


def closed_form(x, m):
    """Stable partition by class = numpy's stable (merge) sort on the class."""
    c = x % m
    perm = np.argsort(c, kind="stable")
    return x[perm], np.bincount(c, minlength=m).astype(np.uint32)


def invariants_hold(x, out, counts, m):
    c_in = x % m
    c_out = out % m
    if not np.array_equal(np.sort(x), np.sort(out)):
        return False                                   # permutation
    if np.any(np.diff(c_out.astype(np.int64)) < 0):
        return False                                   # classes non-decreasing
    if not np.array_equal(np.bincount(c_in, minlength=m), counts):
        return False
    for k in range(m):                                 # order kept inside a class
        if not np.array_equal(x[c_in == k], out[c_out == k]):
            return False
    return True


@pytest.mark.parametrize("case", KAT["plain"], ids=lambda c: f"m{c['m']}n{len(c['input'])}")
def test_oracle_kat_plain(case):
    x = np.array(case["input"], dtype=np.uint32)
    out, cnt, sizes = oracle.seq_mdiscr(x, case["m"], oracle.MOD, sizes=True)
    assert out.tolist() == case["out"]
    assert sizes.tolist() == case["sizes"]
    assert sizes.tolist() == oracle.sizes_from_counts(cnt, len(x)).tolist()


def test_oracle_kat_segmented():
    case = KAT["segmented"][0]
    x = np.array(case["input"], dtype=np.uint32)
    out, new_sizes = oracle.segm_flags(x, np.array(case["segment_sizes"]), case["m"])
    assert out.tolist() == case["out"]
    assert new_sizes.tolist() == case["new_sizes"]
    # The same partition through segment offsets.
    f = np.array(case["segment_sizes"])
    lengths = f[f > 0]
    out2, cnt = oracle.seq_mdiscr_segm(x, W.offsets_from_lengths(lengths), case["m"])
    assert out2.tolist() == case["out"]
    assert cnt.sum() == len(x)


@pytest.mark.parametrize("case", KAT["seeded"], ids=lambda c: c["name"])
def test_oracle_kat_seeded(case):
    x = W.uniform_u31(case["n"], case["seed"])
    assert x[:8].tolist() == case["input_head"]       # generator reproduces the fixture input
    out, cnt, sizes = oracle.seq_mdiscr(x, case["m"], oracle.MOD, sizes=True)
    assert out[:16].tolist() == case["out_head"]
    assert out[-16:].tolist() == case["out_tail"]
    assert cnt.tolist() == case["counts"]
    assert fnv_fast(out) == case["out_digest"]
    assert fnv_fast(sizes) == case["sizes_digest"]


@needs_ref
@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7, 8, 9, 16, 31, 32, 63, 64])
def test_oracle_matches_reference_seqmdiscr(m):
    rng = np.random.default_rng(m)
    for n in (0, 1, 2, 17, 1000, 65537):
        x = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
        out, cnt, sizes = oracle.seq_mdiscr(x, m, oracle.MOD, sizes=True)
        rout, rsizes = oracle.ref_seq_mdiscr(x, m)
        assert np.array_equal(out, rout)
        assert np.array_equal(sizes, rsizes)
        if m in (4, 8):   # Mod4 / Mod8 specialisations (x & 3, x & 7)
            o2, c2 = oracle.seq_mdiscr(x, m, oracle.MASK)
            s_out, s_sz = oracle.ref_seq_mdiscr(x, m, specialized=True)
            assert np.array_equal(o2, s_out)


@needs_ref
def test_sizes_from_counts_matches_reference():
    rng = np.random.default_rng(5)
    for m in (1, 3, 8, 64):
        for _ in range(20):
            counts = rng.integers(0, 4, size=m).astype(np.uint32)
            counts[rng.integers(0, m)] += 1
            n = int(counts.sum())
            assert np.array_equal(oracle.sizes_from_counts(counts, n),
                                  oracle.ref_tuple_to_size_array(counts, n))


@pytest.mark.parametrize("m", [1, 2, 3, 4, 8, 13, 64])
def test_oracle_matches_closed_form(m):
    x = W.uniform_u31(50000, 99 + m)
    out, cnt = oracle.seq_mdiscr(x, m, oracle.MOD)
    cf_out, cf_cnt = closed_form(x, m)
    assert np.array_equal(out, cf_out)
    assert np.array_equal(cnt, cf_cnt)
    assert invariants_hold(x, out, cnt, m)


def test_oracle_brute_force_tiny():
    """Every class string of length <= 6 over m <= 3 (values encode position)."""
    for m in (1, 2, 3):
        for n in range(0, 7):
            for cls in itertools.product(range(m), repeat=n):
                x = np.array([c + m * i for i, c in enumerate(cls)], dtype=np.uint32)
                out, cnt = oracle.seq_mdiscr(x, m, oracle.MOD)
                expect = [v for k in range(m) for v in x.tolist() if v % m == k]
                assert out.tolist() == expect


def test_planted_mistakes_are_caught():
    """Plausible oracle bugs (unstable order, wrong offset, swapped classes)
    must fail the closed-form and invariant checks."""
    x = W.uniform_u31(4096, 123)
    m = 8
    out, cnt = oracle.seq_mdiscr(x, m, oracle.MOD)
    cf, _ = closed_form(x, m)
    # 1. order reversed inside class 3
    bad = out.copy()
    sel = np.flatnonzero(bad % m == 3)
    bad[sel] = bad[sel[::-1]]
    assert not np.array_equal(bad, cf) and not invariants_hold(x, bad, cnt, m)
    # 2. off-by-one offset: class runs shifted by one slot
    bad = np.roll(out, 1)
    assert not np.array_equal(bad, cf) and not invariants_hold(x, bad, cnt, m)
    # 3. classes 0 and 1 emitted in the wrong order
    c = out % m
    bad = np.concatenate([out[c == 1], out[c == 0], out[c >= 2]])
    assert not np.array_equal(bad, cf) and not invariants_hold(x, bad, cnt, m)


def test_soa_oracle():
    n, m = 20000, 16
    key = W.uniform_u31(n, 1)
    pay = W.normal_f32_bits(n, 2)
    idx = np.arange(n, dtype=np.uint32)
    b8 = (key & 0xFF).astype(np.uint8)
    outs, cnt = oracle.seq_mdiscr_soa([key, pay, idx, b8], m, key_col=0)
    perm = np.argsort(key % m, kind="stable")
    assert np.array_equal(outs[0], key[perm])
    assert np.array_equal(outs[1], pay[perm])
    assert np.array_equal(outs[2], perm.astype(np.uint32))
    assert np.array_equal(outs[3], b8[perm])
    assert np.array_equal(cnt, np.bincount(key % m, minlength=m))
    # key not in column 0
    outs2, _ = oracle.seq_mdiscr_soa([pay, key], m, key_col=1)
    assert np.array_equal(outs2[0], pay[perm])


def test_segmented_oracle_matches_per_segment_closed_form():
    L = W.lognormal_lengths(200, 3, median=300)
    off = W.offsets_from_lengths(L)
    n = int(off[-1])
    x = W.uniform_u31(n, 4)
    out, cnt = oracle.seq_mdiscr_segm(x, off, 4)
    for s in range(len(L)):
        a, b = int(off[s]), int(off[s + 1])
        cf, cc = closed_form(x[a:b], 4)
        assert np.array_equal(out[a:b], cf)
        assert np.array_equal(cnt[s], cc)
    f = W.flags_from_lengths(L, n)
    out2, ns = oracle.segm_flags(x, f, 4)
    assert np.array_equal(out2, out)


def test_openmp_matches_seq():
    x = W.uniform_u31(200000, 11)
    for m in (1, 4, 7, 8):
        out, sizes = oracle.openmp_mdiscr(x, m, oracle.MOD, nthreads=4)
        so, _, ss = oracle.seq_mdiscr(x, m, oracle.MOD, sizes=True)
        assert np.array_equal(out, so)
        assert np.array_equal(sizes.view(np.uint32), ss)


def test_oracle_rejects_bad_arguments():
    x = np.arange(10, dtype=np.uint32)
    with pytest.raises(ValueError):
        oracle.seq_mdiscr(x, 0)
    with pytest.raises(ValueError):
        oracle.seq_mdiscr(x, 6, oracle.MASK)       # MASK needs a power of two
    with pytest.raises(ValueError):
        oracle.sizes_from_counts(np.array([1, 2], dtype=np.uint32), 4)  # counts != n
    with pytest.raises(ValueError):                   # flags that do not tile [0,n)
        oracle.segm_flags(x, np.array([3, 0, 0, 4, 0, 0, 0, 5, 0, 0]), 2)


def test_divmagic_exhaustive_divisors(tmp_path):
    """The Granlund-Montgomery magic used by the Mod<N> kernels equals x % d
    for every d in [1, 64] (plus large d) on edge and sampled x."""
    src = os.path.join(HERE, "c", "test_divmagic.c")
    exe = str(tmp_path / "test_divmagic")
    inc = os.path.join(os.path.dirname(HERE), "mdiscriminator_amd", "csrc")
    subprocess.run(["gcc", "-O2", "-std=c11", "-I", inc, "-o", exe, src], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("length", [1, 2, 3, 8])
def test_arrays_oracle_closed_form(length):
    """Array elements (MdiscrHost_arrays.cu.h:62-86): rows of `length` words,
    class from word 0, whole rows in stable class order."""
    n, m = 3001, 7
    x = W.uniform_u31(n * length, length)
    rows = x.reshape(n, length)
    out, cnt = oracle.seq_mdiscr_arrays(x, length, m)
    perm = np.argsort(rows[:, 0] % m, kind="stable")
    assert np.array_equal(out.reshape(n, length), rows[perm])
    assert np.array_equal(cnt, np.bincount(rows[:, 0] % m, minlength=m))
    # the same partition through the struct-of-arrays oracle, column by column
    cols = [np.ascontiguousarray(rows[:, j]) for j in range(length)]
    outs, cnt2 = oracle.seq_mdiscr_soa(cols, m)
    assert np.array_equal(np.stack(outs, axis=1), out.reshape(n, length))
    assert np.array_equal(cnt, cnt2)


def test_soa_oracle_one_byte_class_key():
    """A 1-byte key column is a class array: class = value (values < m)."""
    n, m = 20000, 13
    cls = (W.uniform_u31(n, 5) % m).astype(np.uint8)
    pay = W.uniform_u31(n, 6)
    outs, cnt = oracle.seq_mdiscr_soa([pay, cls], m, key_col=1)
    perm = np.argsort(cls, kind="stable")
    assert np.array_equal(outs[0], pay[perm])
    assert np.array_equal(outs[1], cls[perm])
    assert np.array_equal(cnt, np.bincount(cls, minlength=m))


# ---- two-way filter (ParFilter.hs:8-21) and the select discriminator ---------

def numpy_pred(x, op, a, b):
    """Each predicate stated with numpy's own comparisons on uint64 copies."""
    x = x.astype(np.uint64)
    a, b = np.uint64(a), np.uint64(b)
    return {oracle.PRED_LT: x < a, oracle.PRED_LE: x <= a, oracle.PRED_GT: x > a,
            oracle.PRED_GE: x >= a, oracle.PRED_EQ: x == a, oracle.PRED_NE: x != a,
            oracle.PRED_BITS: (x & a) == b, oracle.PRED_RANGE: (a <= x) & (x < b)}[op]


PRED_CASES = [
    (oracle.PRED_LT, 1 << 30, 0), (oracle.PRED_LT, 0, 0), (oracle.PRED_LE, 2**32 - 1, 0),
    (oracle.PRED_LE, 12345, 0), (oracle.PRED_GT, 2**32 - 1, 0), (oracle.PRED_GT, 1 << 29, 0),
    (oracle.PRED_GE, 0, 0), (oracle.PRED_GE, 3 << 29, 0), (oracle.PRED_EQ, 7, 0),
    (oracle.PRED_NE, 7, 0), (oracle.PRED_BITS, 1, 0), (oracle.PRED_BITS, 0xF0, 0x30),
    (oracle.PRED_BITS, 0x0F, 0x30), (oracle.PRED_RANGE, 1 << 28, 1 << 30),
    (oracle.PRED_RANGE, 100, 100), (oracle.PRED_RANGE, 9, 3),
]


def test_par_filter_oracle_hand_example():
    """Worked by hand from ParFilter.hs:8-21 and Helpers.hs write/permute with
    cond = even: isT = [0,1,2,2,3], i = 3, isF = [4,4,4,5,5],
    inds = [3,0,1,4,2], flags = write [0,3] [3,2] (replicate 5 0)."""
    x = np.array([5, 2, 8, 3, 4], dtype=np.uint32)
    out, flags, cnt = oracle.par_filter(x, oracle.PRED_BITS, 1, 0)
    assert out.tolist() == [2, 8, 4, 5, 3]
    assert flags.tolist() == [3, 0, 0, 2, 0]
    assert cnt.tolist() == [3, 2]
    # i = n: the write at index n matches no position.
    out, flags, _ = oracle.par_filter(np.array([2, 4], dtype=np.uint32), oracle.PRED_BITS, 1, 0)
    assert out.tolist() == [2, 4] and flags.tolist() == [2, 0]
    # i = 0: write [0,0] [0,2] -- the later write to position 0 wins.
    out, flags, _ = oracle.par_filter(np.array([1, 3], dtype=np.uint32), oracle.PRED_BITS, 1, 0)
    assert out.tolist() == [1, 3] and flags.tolist() == [2, 0]
    out, flags, cnt = oracle.par_filter(np.array([], dtype=np.uint32), oracle.PRED_EQ, 0)
    assert len(out) == 0 and len(flags) == 0 and cnt.tolist() == [0, 0]


@pytest.mark.parametrize("op,a,b", PRED_CASES)
def test_par_filter_oracle_closed_form(op, a, b):
    x = W.uniform_u31(20000, 31 + op) * np.uint32(2) + (W.uniform_u31(20000, 7) & 1)
    x[:4] = [0, 7, 2**32 - 1, 0x30]
    p = numpy_pred(x, op, a, b)
    out, flags, cnt = oracle.par_filter(x, op, a, b)
    assert np.array_equal(out, np.concatenate([x[p], x[~p]]))
    i = int(p.sum())
    assert cnt.tolist() == [i, len(x) - i]
    assert np.array_equal(flags, oracle.sizes_from_counts(cnt, len(x)))


@needs_ref
def test_par_filter_flags_are_reference_size_arrays():
    """flags = write [0,i] [i,n-i] equals the reference's tupleToSizeArray of [i, n-i]."""
    for x in (W.uniform_u31(1000, 3), np.zeros(17, dtype=np.uint32), np.ones(9, dtype=np.uint32)):
        _, flags, cnt = oracle.par_filter(x, oracle.PRED_BITS, 1, 0)
        assert np.array_equal(flags, oracle.ref_tuple_to_size_array(cnt, len(x)))


def test_par_filter_brute_force_tiny():
    """Every keep/drop pattern of length <= 8 (values encode position and bit)."""
    for n in range(0, 9):
        for bits in itertools.product((0, 1), repeat=n):
            x = np.array([2 * i + b for i, b in enumerate(bits)], dtype=np.uint32)
            out, flags, cnt = oracle.par_filter(x, oracle.PRED_BITS, 1, 1)
            keep = [v for v in x.tolist() if v & 1]
            assert out.tolist() == keep + [v for v in x.tolist() if not v & 1]
            expect = [0] * n
            if n:
                expect[0] = len(keep)
            if len(keep) < n:
                expect[len(keep)] = n - len(keep)
            assert flags.tolist() == expect


def test_par_filter_planted_mistakes_are_caught():
    x = W.uniform_u31(4096, 77)
    p = numpy_pred(x, oracle.PRED_LT, 1 << 30, 0)
    out, _, cnt = oracle.par_filter(x, oracle.PRED_LT, 1 << 30)
    cf = np.concatenate([x[p], x[~p]])
    assert np.array_equal(out, cf)
    bad = np.concatenate([x[~p], x[p]])                    # rejected elements first
    assert not np.array_equal(bad, cf)
    bad = np.concatenate([x[p][::-1], x[~p]])              # unstable kept order
    assert not np.array_equal(bad, cf)
    # an off-by-one bound (<= for <) keeps the element equal to the bound
    x2 = x.copy()
    x2[0] = 1 << 30
    le = numpy_pred(x2, oracle.PRED_LE, 1 << 30, 0)
    assert not np.array_equal(oracle.par_filter(x2, oracle.PRED_LT, 1 << 30)[0],
                              np.concatenate([x2[le], x2[~le]]))


def test_filter_soa_oracle():
    n = 30000
    key = W.uniform_u31(n, 8)
    pay = W.splitmix64(n, 9)
    b8 = (key & 0xFF).astype(np.uint8)
    p = numpy_pred(key, oracle.PRED_RANGE, 1 << 29, 3 << 29)
    outs, cnt = oracle.filter_soa([pay, key, b8], oracle.PRED_RANGE, 1 << 29, 3 << 29, key_col=1)
    perm = np.concatenate([np.flatnonzero(p), np.flatnonzero(~p)])
    for o, c in zip(outs, [pay, key, b8]):
        assert np.array_equal(o, c[perm])
    assert cnt.tolist() == [int(p.sum()), n - int(p.sum())]


@pytest.mark.parametrize("m", [4, 5, 8, 64])
def test_select_oracle_closed_form(m):
    """z ? x % N : y % N over the reference's tuple (uint x, uchar y, bool z)."""
    n = 25000
    x = W.uniform_u31(n, m)
    y = (W.uniform_u31(n, m + 1) & 0xFF).astype(np.uint8)
    z = (W.uniform_u31(n, m + 2) & 1).astype(np.uint8)
    outs, cnt = oracle.seq_mdiscr_soa_select([x, y, z], m, sel_col=2, a_col=0, b_col=1)
    cls = np.where(z != 0, x % m, y.astype(np.uint32) % m)
    perm = np.argsort(cls, kind="stable")
    for o, c in zip(outs, [x, y, z]):
        assert np.array_equal(o, c[perm])
    assert np.array_equal(cnt, np.bincount(cls, minlength=m))
    # all-true / all-false selectors reduce to the plain SoA partition on one key
    ones = np.ones(n, dtype=np.uint8)
    o1, _ = oracle.seq_mdiscr_soa_select([x, y, ones], m, 2, 0, 1)
    r1, _ = oracle.seq_mdiscr_soa([x, y, ones], m, key_col=0)
    assert all(np.array_equal(a, b) for a, b in zip(o1, r1))
    zeros = np.zeros(n, dtype=np.uint8)
    o0, _ = oracle.seq_mdiscr_soa_select([x, y, zeros], m, 2, 0, 1)
    r0, _ = oracle.seq_mdiscr_soa([x, y, zeros], m, key_col=1)
    assert all(np.array_equal(a, b) for a, b in zip(o0, r0))
