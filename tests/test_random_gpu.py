"""Seeded random cases over every entry point, bit-exact against the oracle:
sizes up to a few hundred thousand (so tail tiles of every geometry occur),
random class counts and discriminator kinds, random record shapes, array
lengths and alignments, segment lengths (empty ones included), predicates
and batches.  Each case prints its parameters before it runs."""
import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff

pytestmark = pytest.mark.gpu

CASES = 24


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    signed = {np.dtype(np.uint8): np.uint8, np.dtype(np.uint16): np.int16,
              np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}
    return torch.from_numpy(a.view(signed[a.dtype])).to("cuda:0")


def host(t, like):
    return t.cpu().numpy().view(like.dtype)


def rand_size(rng):
    return int(rng.choice([rng.integers(0, 70), rng.integers(0, 5000), rng.integers(0, 300000)]))


def rand_discr(rng, md):
    m = int(rng.integers(1, 65))
    if m & (m - 1) == 0 and rng.integers(0, 2):
        return md.Mask(m), oracle.MASK
    return md.Mod(m), oracle.MOD


def keys(rng, n, seed):
    """Uniform 31-bit keys, or full 32-bit keys, or few distinct values."""
    k = int(rng.integers(0, 3))
    if k == 0:
        return W.uniform_u31(n, seed)
    if k == 1:
        return W.splitmix64(n, seed).astype(np.uint32)
    return (W.uniform_u31(n, seed) % np.uint32(int(rng.integers(1, 9)))).astype(np.uint32)


@pytest.mark.parametrize("case", range(CASES))
def test_random_plain(ctx, case):
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(1000 + case)
    n = rand_size(rng)
    d, kind = rand_discr(rng, md)
    x = keys(rng, n, 2000 + case)
    print(f"plain case {case}: n={n} m={d.m} kind={kind}", flush=True)
    d_in = dev(x) if n else torch.empty(0, dtype=torch.int32, device="cuda:0")
    d_out = torch.full_like(d_in, -1)
    d_cnt = torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_u32(d_in, d_out, d, d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr(x, d.m, kind)
    assert first_diff(host(d_out, x), ref) < 0
    assert np.array_equal(host(d_cnt, ref_cnt), ref_cnt)


@pytest.mark.parametrize("case", range(CASES))
def test_random_filter(ctx, case):
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(3000 + case)
    n = max(1, rand_size(rng))
    op = int(rng.integers(0, 8))
    edges = [0, 1, 7, 2**31, 2**32 - 1]
    a = int(rng.choice(edges)) if rng.integers(0, 3) == 0 else int(rng.integers(0, 2**32))
    b = int(rng.choice(edges)) if rng.integers(0, 3) == 0 else int(rng.integers(0, 2**32))
    if op == oracle.PRED_BITS:
        b &= a if rng.integers(0, 2) else 0xFFFFFFFF
    x = keys(rng, n, 4000 + case)
    print(f"filter case {case}: n={n} op={op} a={a} b={b}", flush=True)
    d_in = dev(x)
    d_out = torch.full_like(d_in, -1)
    d_cnt = torch.full((2,), -1, dtype=torch.int32, device="cuda:0")
    ctx.filter_u32(d_in, d_out, md.Pred(op, a, b), d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, _, ref_cnt = oracle.par_filter(x, op, a, b)
    assert first_diff(host(d_out, x), ref) < 0
    assert np.array_equal(host(d_cnt, ref_cnt), ref_cnt)


@pytest.mark.parametrize("case", range(CASES))
def test_random_soa(ctx, case):
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(5000 + case)
    n = rand_size(rng)
    d, kind = rand_discr(rng, md)
    ncols = int(rng.integers(1, 6))
    widths = [int(rng.choice([1, 2, 4, 8])) for _ in range(ncols)]
    key_col = int(rng.integers(0, ncols))
    widths[key_col] = int(rng.choice([1, 4]))
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
    cols = [W.splitmix64(n, 6000 + 10 * case + j).astype(dt[w]) for j, w in enumerate(widths)]
    if widths[key_col] == 1:
        cols[key_col] = (cols[key_col] % np.uint8(d.m)).astype(np.uint8)
    else:
        cols[key_col] = keys(rng, n, 7000 + case)
    print(f"soa case {case}: n={n} m={d.m} kind={kind} widths={widths} key={key_col}", flush=True)
    d_in = [dev(c) for c in cols]
    d_out = [torch.empty_like(t) for t in d_in]
    d_cnt = torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_soa(d_in, d_out, d, key_col=key_col, d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_soa(cols, d.m, key_col=key_col, kind=kind)
    assert np.array_equal(host(d_cnt, ref_cnt), ref_cnt)
    for j, (o, r) in enumerate(zip(d_out, ref)):
        assert first_diff(host(o, cols[j]), r) < 0, f"column {j}"


@pytest.mark.parametrize("case", range(CASES))
def test_random_arrays(ctx, case):
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(8000 + case)
    n = rand_size(rng)
    length = int(rng.integers(1, 9))
    d, kind = rand_discr(rng, md)
    off = int(rng.integers(0, 4))          # words past a 16-byte boundary
    x = keys(rng, n * length, 9000 + case)
    print(f"arrays case {case}: n={n} len={length} m={d.m} kind={kind} offset={off}", flush=True)
    big_in = torch.zeros(n * length + 8, dtype=torch.int32, device="cuda:0")
    big_in[off:off + n * length] = dev(x)
    big_out = torch.full_like(big_in, -1)
    d_cnt = torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_arrays(big_in[off:off + n * length], big_out[off:off + n * length], length, d,
                         d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_arrays(x, length, d.m, kind)
    assert first_diff(host(big_out[off:off + n * length], x), ref) < 0
    assert np.array_equal(host(d_cnt, ref_cnt), ref_cnt)
    assert bool((big_out[:off] == -1).all()) and bool((big_out[off + n * length:] == -1).all())


@pytest.mark.parametrize("case", range(CASES))
def test_random_segmented(ctx, case):
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(10000 + case)
    nseg = int(rng.integers(1, 300))
    scale = float(rng.choice([30, 3000, 60000]))
    lens = np.floor(rng.exponential(scale, nseg)).astype(np.int64)
    lens[rng.random(nseg) < 0.1] = 0
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    n = int(off[-1])
    m = int(rng.choice([1, 2, 3, 4, 5, 8, 16, 33, 64]))
    d = md.Mask(m) if m & (m - 1) == 0 else md.Mod(m)
    kind = oracle.MASK if m & (m - 1) == 0 else oracle.MOD
    x = keys(rng, n, 11000 + case)
    print(f"segmented case {case}: nseg={nseg} n={n} m={m}", flush=True)
    assert n > 0, "every seed here draws at least one non-empty segment"
    d_in = dev(x)
    d_out = torch.full_like(d_in, -1)
    d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
    d_cnt = torch.full((nseg * m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_segmented(d_in, d_out, d_off, d, h_seg_off=off, d_seg_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_segm(x, off, m, kind)
    assert first_diff(host(d_out, x), ref) < 0
    assert np.array_equal(host(d_cnt, ref_cnt).reshape(nseg, m), ref_cnt)


@pytest.mark.parametrize("case", range(CASES))
def test_random_batch(ctx, case):
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(12000 + case)
    count = int(rng.integers(1, 40))
    reqs, xs, ms, kinds = [], [], [], []
    for i in range(count):
        n = rand_size(rng)
        d, kind = rand_discr(rng, md)
        x = keys(rng, n, 13000 + 100 * case + i)
        d_in = dev(x) if n else torch.empty(0, dtype=torch.int32, device="cuda:0")
        reqs.append((d_in, torch.full_like(d_in, -1), d,
                     torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0")))
        xs.append(x)
        ms.append(d.m)
        kinds.append(kind)
    print(f"batch case {case}: {count} requests, n total {sum(len(x) for x in xs)}", flush=True)
    ctx.partition_batch(reqs)
    torch.cuda.synchronize()
    for i, x in enumerate(xs):
        ref, ref_cnt = oracle.seq_mdiscr(x, ms[i], kinds[i])
        assert first_diff(host(reqs[i][1], x), ref) < 0, f"request {i}"
        assert np.array_equal(host(reqs[i][3], ref_cnt), ref_cnt), f"request {i}"


def rand_record_size(rng):
    """Sizes near multiples of every SoA tile (3072, 4096, 6144, 12288) or random."""
    k = int(rng.integers(0, 3))
    if k == 0:
        t = int(rng.choice([3072, 4096, 6144, 12288]))
        return max(0, t * int(rng.integers(1, 40)) + int(rng.integers(-2, 3)))
    return rand_size(rng)


@pytest.mark.parametrize("case", range(CASES))
def test_random_select(ctx, case):
    """z ? D(x) : D(y) with random widths (1 or 4 bytes) for the three class
    columns, which may coincide, plus a random payload column."""
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(14000 + case)
    n = rand_record_size(rng)
    d, kind = rand_discr(rng, md)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
    widths = [int(rng.choice([1, 4])) for _ in range(3)] + [int(rng.choice([1, 2, 4, 8]))]
    cols = [W.splitmix64(n, 15000 + 10 * case + j).astype(dt[w]) for j, w in enumerate(widths)]
    sel, a, b = (int(v) for v in rng.integers(0, 4, size=3))
    while widths[sel] not in (1, 4) or widths[a] not in (1, 4) or widths[b] not in (1, 4):
        sel, a, b = (int(v) for v in rng.integers(0, 4, size=3))
    if rng.integers(0, 2):  # a 0/1 selector, as the reference's bool column
        cols[sel] = (cols[sel] & dt[widths[sel]](1)).astype(dt[widths[sel]])
    print(f"select case {case}: n={n} m={d.m} kind={kind} widths={widths} sel={sel} a={a} b={b}",
          flush=True)
    d_in = [dev(c) for c in cols]
    d_out = [torch.empty_like(t) for t in d_in]
    d_cnt = torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_soa_select(d_in, d_out, d, sel_col=sel, a_col=a, b_col=b, d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_soa_select(cols, d.m, sel, a, b, kind=kind)
    assert np.array_equal(host(d_cnt, ref_cnt), ref_cnt)
    for j, (o, r) in enumerate(zip(d_out, ref)):
        assert first_diff(host(o, cols[j]), r) < 0, f"column {j}"


@pytest.mark.parametrize("case", range(CASES))
def test_random_filter_records(ctx, case):
    """The filter over records: random predicate on a 4-byte key column."""
    import torch
    import mdiscriminator_amd as md
    rng = np.random.default_rng(16000 + case)
    n = rand_record_size(rng)
    op = int(rng.integers(0, 8))
    a, b = int(rng.integers(0, 2**32)), int(rng.integers(0, 2**32))
    if op == oracle.PRED_BITS:
        b &= a
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
    ncols = int(rng.integers(1, 5))
    widths = [int(rng.choice([1, 2, 4, 8])) for _ in range(ncols)]
    key_col = int(rng.integers(0, ncols))
    widths[key_col] = 4
    cols = [W.splitmix64(n, 17000 + 10 * case + j).astype(dt[w]) for j, w in enumerate(widths)]
    cols[key_col] = keys(rng, n, 18000 + case)
    print(f"filter records case {case}: n={n} op={op} a={a} b={b} widths={widths} key={key_col}",
          flush=True)
    d_in = [dev(c) for c in cols]
    d_out = [torch.empty_like(t) for t in d_in]
    d_cnt = torch.full((2,), -1, dtype=torch.int32, device="cuda:0")
    ctx.filter_soa(d_in, d_out, md.Pred(op, a, b), key_col=key_col, d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.filter_soa(cols, op, a, b, key_col=key_col)
    assert np.array_equal(host(d_cnt, ref_cnt), ref_cnt)
    for j, (o, r) in enumerate(zip(d_out, ref)):
        assert first_diff(host(o, cols[j]), r) < 0, f"column {j}"
