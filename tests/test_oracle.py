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
