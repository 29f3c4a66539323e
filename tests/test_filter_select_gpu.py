"""General discriminators on the GPU, bit-exact against the oracle:

* the two-way filter (parFilter, haskell_code/ParFilter.hs:8-21) over plain
  uint32 arrays and over struct-of-arrays records, for every predicate kind at
  its edge parameters (empty and full windows, bounds at 0 and 2^32-1);
* the multi-column select discriminator, the alternative body of
  ModTuple4/ModTuple<N>::apply (MdiscrHost_tuples.cu.h:27,46):
  z ? x % N : y % N over the reference's tuple (uint x, uchar y, bool z).
"""
import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff

pytestmark = pytest.mark.gpu

SIZES = (1, 4095, 4096, 4097, 100003, 1 << 20)

PREDS = [
    ("lt", oracle.PRED_LT, 1 << 30, 0), ("lt0", oracle.PRED_LT, 0, 0),
    ("le_max", oracle.PRED_LE, 2**32 - 1, 0), ("le", oracle.PRED_LE, 12345, 0),
    ("gt_max", oracle.PRED_GT, 2**32 - 1, 0), ("gt", oracle.PRED_GT, 1 << 29, 0),
    ("ge0", oracle.PRED_GE, 0, 0), ("ge", oracle.PRED_GE, 3 << 29, 0),
    ("eq", oracle.PRED_EQ, 7, 0), ("ne", oracle.PRED_NE, 7, 0),
    ("even", oracle.PRED_BITS, 1, 0), ("bits", oracle.PRED_BITS, 0xF0, 0x30),
    ("bits_never", oracle.PRED_BITS, 0x0F, 0x30), ("range", oracle.PRED_RANGE, 1 << 28, 1 << 30),
    ("range_empty", oracle.PRED_RANGE, 100, 100), ("range_inverted", oracle.PRED_RANGE, 9, 3),
]


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    signed = {np.dtype(np.uint8): np.uint8, np.dtype(np.uint16): np.int16,
              np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}
    return torch.from_numpy(a.view(signed.get(a.dtype, a.dtype))).to("cuda:0")


def host(t, like):
    return t.cpu().numpy().view(like.dtype)


def filter_input(n, seed):
    """Random keys over the full 32-bit range plus the values the predicates'
    bounds single out."""
    x = W.uniform_u31(n, seed) * np.uint32(2) + (W.uniform_u31(n, seed + 1) & 1)
    special = np.array([0, 7, 2**32 - 1, 0x30, 12345, 1 << 28, (1 << 30) - 1, 1 << 30,
                        3 << 29, 100, 9], dtype=np.uint32)
    k = min(n, len(special))
    x[:k] = special[:k]
    return x


@pytest.mark.parametrize("name,op,a,b", PREDS, ids=[p[0] for p in PREDS])
def test_filter_u32(ctx, name, op, a, b):
    import torch
    import mdiscriminator_amd as md
    pred = md.Pred(op, a, b)
    for n in SIZES:
        x = filter_input(n, 50 + op + n % 97)
        d_in = dev(x)
        d_out = torch.full_like(d_in, -1)
        d_cnt = torch.full((2,), -1, dtype=torch.int32, device="cuda:0")
        print(f"filter {name} n={n}", flush=True)
        ctx.filter_u32(d_in, d_out, pred, d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, _, ref_cnt = oracle.par_filter(x, op, a, b)
        assert first_diff(host(d_out, x), ref) < 0, f"{name} n={n}"
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)


def test_par_filter_module_call_with_flags(ctx):
    """par_filter(cond, arr) -> (arr', flags), flags as write [0,i] [i,n-i]."""
    import torch
    import mdiscriminator_amd as md
    for n, seed in ((5, 1), (100003, 2), (1 << 20, 3)):
        x = filter_input(n, seed)
        d_in = dev(x)
        d_out = torch.empty_like(d_in)
        d_flags = torch.full_like(d_in, -1)
        kept = md.par_filter(md.Pred.lt(1 << 30), d_in, d_out, d_flags)
        ref, ref_flags, ref_cnt = oracle.par_filter(x, oracle.PRED_LT, 1 << 30)
        assert kept == ref_cnt[0]
        assert np.array_equal(host(d_out, x), ref)
        assert np.array_equal(host(d_flags, x), ref_flags)
    # all kept / none kept: the flag array is [n, 0, ...] either way
    for x in (np.zeros(4097, dtype=np.uint32), np.full(4097, 2**32 - 1, dtype=np.uint32)):
        d_flags = torch.full((len(x),), -1, dtype=torch.int32, device="cuda:0")
        md.par_filter(md.Pred.eq(0), dev(x), torch.empty(len(x), dtype=torch.int32,
                                                          device="cuda:0"), d_flags)
        _, ref_flags, _ = oracle.par_filter(x, oracle.PRED_EQ, 0)
        assert np.array_equal(host(d_flags, x), ref_flags)


def test_filter_u32_empty_and_unaligned(ctx):
    import torch
    import mdiscriminator_amd as md
    d_cnt = torch.full((2,), -1, dtype=torch.int32, device="cuda:0")
    e = torch.empty(0, dtype=torch.int32, device="cuda:0")
    ctx.filter_u32(e, torch.empty_like(e), md.Pred.lt(5), d_counts=d_cnt)
    torch.cuda.synchronize()
    assert d_cnt.cpu().tolist() == [0, 0]
    # a 4-byte-aligned view that is not 16-byte aligned takes the scalar count path
    x = filter_input(70001, 9)
    big = dev(np.concatenate([np.zeros(1, dtype=np.uint32), x]))
    out = torch.empty(70001, dtype=torch.int32, device="cuda:0")
    ctx.filter_u32(big[1:], out, md.Pred.range(1 << 28, 1 << 30), d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, _, ref_cnt = oracle.par_filter(x, oracle.PRED_RANGE, 1 << 28, 1 << 30)
    assert np.array_equal(host(out, x), ref)
    assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)


@pytest.mark.parametrize("wide", [False, True], ids=["4byte", "8byte"])
def test_filter_soa(ctx, wide):
    import torch
    import mdiscriminator_amd as md
    for n in (1, 8191, 8192, 100003, 1 << 20):
        key = filter_input(n, 70 + n % 13)
        idx = np.arange(n, dtype=np.uint32)
        pay = W.splitmix64(n, 71) if wide else W.normal_f32_bits(n, 71)
        b8 = (key >> 3 & 0xFF).astype(np.uint8)
        cols = [pay, key, idx, b8]
        d_in = [dev(c) for c in cols]
        d_out = [torch.empty_like(t) for t in d_in]
        d_cnt = torch.empty(2, dtype=torch.int32, device="cuda:0")
        print(f"filter_soa wide={wide} n={n}", flush=True)
        ctx.filter_soa(d_in, d_out, md.Pred.ge(3 << 29), key_col=1, d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.filter_soa(cols, oracle.PRED_GE, 3 << 29, key_col=1)
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
        for j, (o, r) in enumerate(zip(d_out, ref)):
            assert first_diff(host(o, cols[j]), r) < 0, f"n={n} column {j}"


@pytest.mark.parametrize("m", [2, 4, 5, 8, 13, 64])
def test_select_reference_tuple(ctx, m):
    """(uint x, uchar y, bool z), class = z ? x % N : y % N."""
    import torch
    import mdiscriminator_amd as md
    for n in (1, 4095, 4097, 100003, 1 << 20):
        x = W.uniform_u31(n, 80 + m)
        y = (W.uniform_u31(n, 81 + m) & 0xFF).astype(np.uint8)
        z = (W.uniform_u31(n, 82 + m) & 1).astype(np.uint8)
        cols = [x, y, z]
        d_in = [dev(c) for c in cols]
        d_out = [torch.empty_like(t) for t in d_in]
        d_cnt = torch.empty(m, dtype=torch.int32, device="cuda:0")
        print(f"select m={m} n={n}", flush=True)
        ctx.partition_soa_select(d_in, d_out, md.Mod(m), sel_col=2, a_col=0, b_col=1,
                                 d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.seq_mdiscr_soa_select(cols, m, 2, 0, 1)
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
        for j, (o, r) in enumerate(zip(d_out, ref)):
            assert first_diff(host(o, cols[j]), r) < 0, f"m={m} n={n} column {j}"


def test_select_mixed_widths_and_shared_columns(ctx):
    """4-byte selector, 1-byte first key, 4-byte second key, an 8-byte payload,
    and a select whose selector is also a key column."""
    import torch
    import mdiscriminator_amd as md
    n = 300007
    for m, kind in ((8, md.Mask(8)), (11, md.Mod(11)), (32, md.Mask(32))):
        s32 = W.uniform_u31(n, m) % np.uint32(3)
        a8 = (W.uniform_u31(n, m + 1) & 0xFF).astype(np.uint8)
        b32 = W.uniform_u31(n, m + 2)
        wide = W.splitmix64(n, m + 3)
        cols = [wide, s32, a8, b32]
        d_in = [dev(c) for c in cols]
        d_out = [torch.empty_like(t) for t in d_in]
        ctx.partition_soa_select(d_in, d_out, kind, sel_col=1, a_col=2, b_col=3)
        torch.cuda.synchronize()
        ref, _ = oracle.seq_mdiscr_soa_select(cols, m, 1, 2, 3, kind=kind.kind)
        for j, (o, r) in enumerate(zip(d_out, ref)):
            assert first_diff(host(o, cols[j]), r) < 0, f"m={m} column {j}"
        # selector doubles as the first key
        cols2 = [b32, s32]
        d_in = [dev(c) for c in cols2]
        d_out = [torch.empty_like(t) for t in d_in]
        ctx.partition_soa_select(d_in, d_out, kind, sel_col=1, a_col=1, b_col=0)
        torch.cuda.synchronize()
        ref, _ = oracle.seq_mdiscr_soa_select(cols2, m, 1, 1, 0, kind=kind.kind)
        for o, r, c in zip(d_out, ref, cols2):
            assert np.array_equal(host(o, c), r)


def test_general_discriminators_invalid(ctx):
    import torch
    import mdiscriminator_amd as md
    a = torch.zeros(100, dtype=torch.int32, device="cuda:0")
    o = torch.zeros(100, dtype=torch.int32, device="cuda:0")
    b8 = torch.zeros(100, dtype=torch.uint8, device="cuda:0")
    b16 = torch.zeros(100, dtype=torch.int16, device="cuda:0")
    with pytest.raises(md.MdiscrError):
        ctx.filter_u32(a, o, md.Pred(99, 0, 0))                          # unknown predicate
    with pytest.raises(md.MdiscrError):
        ctx.filter_u32(a, a, md.Pred.lt(3))                              # aliasing
    with pytest.raises(md.MdiscrError):
        ctx.filter_soa([b8, a], [torch.empty_like(b8), o], md.Pred.lt(3), key_col=0)  # 1-byte key
    with pytest.raises(md.MdiscrError):
        ctx.partition_soa_select([a, b16], [o, torch.empty_like(b16)], md.Mod(4), 0, 1, 0)
    with pytest.raises(md.MdiscrError):
        ctx.partition_soa_select([a, b8], [o, torch.empty_like(b8)], md.Mod(4), 0, 1, 2)


def test_filter_soa_four_byte_records(ctx):
    """The filter over records of contiguous 4-byte columns (one staging pass)."""
    import torch
    import mdiscriminator_amd as md
    for n in (1, 4097, 100003, 1 << 20):
        key = filter_input(n, 90 + n % 5)
        cols = [np.arange(n, dtype=np.uint32), key, W.normal_f32_bits(n, 91)]
        d_in = [dev(c) for c in cols]
        d_out = [torch.empty_like(t) for t in d_in]
        d_cnt = torch.empty(2, dtype=torch.int32, device="cuda:0")
        ctx.filter_soa(d_in, d_out, md.Pred.bits(3, 1), key_col=1, d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.filter_soa(cols, oracle.PRED_BITS, 3, 1, key_col=1)
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
        for j, (o, r) in enumerate(zip(d_out, ref)):
            assert first_diff(host(o, cols[j]), r) < 0, f"n={n} column {j}"


def test_filter_full_size_2p28(ctx):
    """parFilter over the plain configuration's 2^28 keys, against the oracle."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 28
    d_in = W.uniform_u31_torch(n, 0x5EED0000 + 2, "cuda:0")
    x = d_in.cpu().numpy().view(np.uint32)
    d_out = torch.empty_like(d_in)
    d_cnt = torch.empty(2, dtype=torch.int32, device="cuda:0")
    print("filter 2^28", flush=True)
    ctx.filter_u32(d_in, d_out, md.Pred.lt(1 << 30), d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, _, ref_cnt = oracle.par_filter(x, oracle.PRED_LT, 1 << 30)
    assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
    assert first_diff(d_out.cpu().numpy().view(np.uint32), ref) < 0


def test_select_full_size_2p26(ctx):
    """The select discriminator over 2^26 (uint, uchar, bool) records."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 26
    x = W.uniform_u31(n, 31)
    y = (W.uniform_u31(n, 32) & 0xFF).astype(np.uint8)
    z = (W.uniform_u31(n, 33) & 1).astype(np.uint8)
    cols = [x, y, z]
    d_in = [dev(c) for c in cols]
    d_out = [torch.empty_like(t) for t in d_in]
    d_cnt = torch.empty(8, dtype=torch.int32, device="cuda:0")
    print("select 2^26", flush=True)
    ctx.partition_soa_select(d_in, d_out, md.Mask(8), 2, 0, 1, d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_soa_select(cols, 8, 2, 0, 1, kind=oracle.MASK)
    assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
    for j, (o, r) in enumerate(zip(d_out, ref)):
        assert first_diff(host(o, cols[j]), r) < 0, f"column {j}"


def test_select_misaligned_columns(ctx):
    """Select columns off their natural word boundaries: a 4-byte key that is
    not 16-byte aligned and byte columns that are not 4-byte aligned."""
    import torch
    import mdiscriminator_amd as md
    n, m = 70001, 8
    x = W.uniform_u31(n, 41)
    y = (W.uniform_u31(n, 42) & 0xFF).astype(np.uint8)
    z = (W.uniform_u31(n, 43) & 1).astype(np.uint8)
    cols = [x, y, z]

    def place(a, off):
        raw = np.zeros(a.nbytes + 16, dtype=np.uint8)
        raw[off:off + a.nbytes] = a.view(np.uint8)
        v = torch.from_numpy(raw).to("cuda:0")[off:off + a.nbytes]
        return v if a.dtype == np.uint8 else v.view(torch.int32)

    for off_x, off_b in ((4, 0), (0, 1), (8, 3)):
        print(f"select x+{off_x} bytes+{off_b}", flush=True)
        d_in = [place(x, off_x), place(y, off_b), place(z, off_b)]
        d_out = [torch.empty(n, dtype=torch.int32, device="cuda:0"),
                 torch.empty(n, dtype=torch.uint8, device="cuda:0"),
                 torch.empty(n, dtype=torch.uint8, device="cuda:0")]
        d_cnt = torch.empty(m, dtype=torch.int32, device="cuda:0")
        ctx.partition_soa_select(d_in, d_out, md.Mod(m), sel_col=2, a_col=0, b_col=1,
                                 d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.seq_mdiscr_soa_select(cols, m, 2, 0, 1)
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
        for j, (o, r) in enumerate(zip(d_out, ref)):
            assert first_diff(host(o, cols[j]), r) < 0, f"x+{off_x} b+{off_b} column {j}"
