"""Struct-of-arrays partition on the GPU, bit-exact against the oracle
(restated MdiscrHost_tuples.cu.h:81-112): the class comes from the key column,
every column moves to the same index as raw bits, and stability is checked
through an explicit original-index column."""
import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    signed = {np.dtype(np.uint8): np.uint8, np.dtype(np.uint16): np.int16,
              np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}
    return torch.from_numpy(a.view(signed.get(a.dtype, a.dtype))).to("cuda:0")


def host(t, like):
    return t.cpu().numpy().view(like.dtype)


def run_soa(ctx, cols, discr, key_col=0):
    import torch
    d_in = [dev(c) for c in cols]
    d_out = [torch.empty_like(t) for t in d_in]
    d_cnt = torch.full((discr.m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_soa(d_in, d_out, discr, key_col=key_col, d_counts=d_cnt)
    torch.cuda.synchronize()
    return [host(o, c) for o, c in zip(d_out, cols)], d_cnt.cpu().numpy().view(np.uint32)


def check_soa(ctx, cols, discr, key_col=0, label=""):
    outs, cnt = run_soa(ctx, cols, discr, key_col)
    ref_outs, ref_cnt = oracle.seq_mdiscr_soa(cols, discr.m, key_col=key_col, kind=discr.kind)
    assert np.array_equal(cnt, ref_cnt), label
    for j, (o, r) in enumerate(zip(outs, ref_outs)):
        i = first_diff(o, r)
        assert i < 0, f"{label} column {j} first mismatch at {i}"
    return outs


@pytest.mark.parametrize("m", [16, 32, 64])
def test_soa_large_m_records(ctx, m):
    """The large-m record layout: (int32 key, float32 payload bits, int32 index)."""
    import mdiscriminator_amd as md
    for n in (1, 4095, 4096, 4097, 100003, 1 << 20):
        key = W.uniform_u31(n, 300 + m)
        pay = W.normal_f32_bits(n, 400 + m)
        idx = np.arange(n, dtype=np.uint32)
        print(f"soa m={m} n={n}", flush=True)
        outs = check_soa(ctx, [key, pay, idx], md.Mask(m), label=f"m={m} n={n}")
        # stability through the index column
        c = outs[0] & (m - 1)
        for k in range(m):
            s = outs[2][c == k]
            assert np.all(np.diff(s.astype(np.int64)) > 0)


def test_soa_nan_payload_bits_survive(ctx):
    """Payloads move as integers: signalling/quiet NaN patterns and -0.0 keep their bits."""
    import mdiscriminator_amd as md
    n = 20000
    key = W.uniform_u31(n, 1)
    pay = np.array([0x7FA00001, 0xFFC00000, 0x80000000, 0x7F800000, 0x00000001],
                   dtype=np.uint32)[np.arange(n) % 5]
    check_soa(ctx, [key, pay], md.Mod(10), label="nan bits")


def test_soa_reference_tuple_types(ctx):
    """The reference's own tuple: (unsigned int, unsigned char, bool), class from column 1
    (ModTuple<N>, MdiscrHost_tuples.cu.h:31-48), plus 2- and 8-byte columns and a key that
    is not the first column."""
    import mdiscriminator_amd as md
    n = 77777
    key = W.uniform_u31(n, 9)
    u8 = (W.uniform_u31(n, 10) & 0xFF).astype(np.uint8)
    b = (W.uniform_u31(n, 11) & 1).astype(np.uint8)
    for m in (4, 5, 8):
        check_soa(ctx, [key, u8, b], md.Mod(m), label=f"tuple m={m}")
    u16 = (W.uniform_u31(n, 12) & 0xFFFF).astype(np.uint16)
    u64 = W.splitmix64(n, 13)
    check_soa(ctx, [u16, u64, key], md.Mod(13), key_col=2, label="key last")


def test_soa_module_level_tuple_call(ctx):
    import torch
    import mdiscriminator_amd as md
    n = 50000
    key = W.uniform_u31(n, 21)
    u8 = (key & 0xFF).astype(np.uint8)
    b = (key >> 3 & 1).astype(np.uint8)
    d1, d2, d3 = dev(key), dev(u8), dev(b)
    o1, o2, o3 = torch.empty_like(d1), torch.empty_like(d2), torch.empty_like(d3)
    counts = md.mdiscr_tuples(md.Mod(8), n, 1 << 16, d1, d2, d3, o1, o2, o3)
    ref, ref_cnt = oracle.seq_mdiscr_soa([key, u8, b], 8)
    assert counts == ref_cnt.tolist()
    assert np.array_equal(host(o1, key), ref[0])
    assert np.array_equal(host(o2, u8), ref[1])
    assert np.array_equal(host(o3, b), ref[2])


@pytest.mark.parametrize("m", [16, 32, 64])
def test_soa_full_size_2p26(ctx, m):
    """Config-size run: 2^26 (key, f32 bits, index) records at each class
    count of the configuration, checked against the oracle."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 26
    seed = 0x5EED0000 + 3
    print(f"soa full size m={m}", flush=True)
    d_key = W.uniform_u31_torch(n, seed, "cuda:0")
    key = d_key.cpu().numpy().view(np.uint32)
    pay = W.normal_f32_bits(n, seed + 1)
    d_pay = dev(pay)
    d_idx = torch.arange(n, dtype=torch.int32, device="cuda:0")
    outs = [torch.full_like(d_key, -1), torch.full_like(d_pay, -1), torch.full_like(d_idx, -1)]
    d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_soa([d_key, d_pay, d_idx], outs, md.Mask(m), d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_soa([key, pay, np.arange(n, dtype=np.uint32)], m,
                                         kind=oracle.MASK)
    assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
    for o, r in zip(outs, ref):
        assert first_diff(o.cpu().numpy().view(np.uint32), r) < 0


def test_soa_invalid(ctx):
    import torch
    import mdiscriminator_amd as md
    a = torch.zeros(100, dtype=torch.int32, device="cuda:0")
    b16 = torch.zeros(100, dtype=torch.int16, device="cuda:0")
    o = torch.zeros(100, dtype=torch.int32, device="cuda:0")
    with pytest.raises(md.MdiscrError):
        ctx.partition_soa([b16, a], [torch.empty_like(b16), o], md.Mod4, key_col=0)  # 2-byte key
    with pytest.raises(md.MdiscrError):
        ctx.partition_soa([a], [a], md.Mod4)                                        # aliasing


@pytest.mark.parametrize("length", [1, 2, 3, 4, 5, 6, 7, 8])
def test_array_elements(ctx, length):
    """Array-valued elements (mdiscr<DISCR,LEN>, MdiscrHost_arrays.cu.h:250-470),
    around the row kernel's tile sizes (1024, 2048 and 4096 elements)."""
    import torch
    import mdiscriminator_amd as md
    for n, m in ((1, 4), (1023, 2), (1025, 3), (2049, 1), (4097, 16), (8191, 8), (100003, 5),
                 (1 << 18, 64)):
        x = W.uniform_u31(n * length, 900 + length + n)
        d_in = torch.from_numpy(x.view(np.int32)).to("cuda:0")
        d_out = torch.full_like(d_in, -1)
        d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
        print(f"arrays len={length} n={n} m={m}", flush=True)
        ctx.partition_arrays(d_in, d_out, length, md.Mod(m), d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.seq_mdiscr_arrays(x, length, m)
        assert first_diff(d_out.cpu().numpy().view(np.uint32), ref) < 0
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)


def test_class_array_key(ctx):
    """A 1-byte class column as the discriminator, with 4- and 8-byte payloads."""
    import mdiscriminator_amd as md
    for m, n in ((3, 50001), (16, 1 << 20), (64, 300000)):
        cls = (W.uniform_u31(n, m) % m).astype(np.uint8)
        pay = W.uniform_u31(n, m + 1)
        wide = W.splitmix64(n, m + 2)
        check_soa(ctx, [pay, cls], md.Mod(m), key_col=1, label=f"class key m={m}")
        check_soa(ctx, [cls, wide, pay], md.Mod(m), key_col=0, label=f"class key wide m={m}")


@pytest.mark.parametrize("ncols,key_col", [(2, 0), (2, 1), (3, 2), (4, 0), (4, 3)])
def test_four_byte_records_every_shape(ctx, ncols, key_col):
    """Records of 2..4 contiguous 4-byte columns (one staging pass for all
    columns), key in any column, tail tiles and large m included."""
    import mdiscriminator_amd as md
    for n, m in ((1, 4), (4095, 7), (4097, 16), (100003, 64), (1 << 20, 32)):
        cols = [W.uniform_u31(n, 500 + 10 * j + n % 7) for j in range(ncols)]
        cols[(key_col + 1) % ncols] = np.arange(n, dtype=np.uint32)   # an index column
        print(f"w4 ncols={ncols} key={key_col} n={n} m={m}", flush=True)
        check_soa(ctx, cols, md.Mod(m), key_col=key_col, label=f"w4 {ncols}/{key_col} n={n}")


def test_four_byte_records_unaligned_key(ctx):
    """A 4-byte key column that is not 16-byte aligned takes the scalar count path."""
    import torch
    import mdiscriminator_amd as md
    n, m = 70001, 16
    key = W.uniform_u31(n, 61)
    pay = W.normal_f32_bits(n, 62)
    big = dev(np.concatenate([np.zeros(1, dtype=np.uint32), key]))
    d_key = big[1:]
    d_pay = dev(pay)
    outs = [torch.empty(n, dtype=torch.int32, device="cuda:0"), torch.empty_like(d_pay)]
    ctx.partition_soa([d_key, d_pay], outs, md.Mask(m))
    torch.cuda.synchronize()
    ref, _ = oracle.seq_mdiscr_soa([key, pay], m, kind=oracle.MASK)
    assert np.array_equal(host(outs[0], key), ref[0])
    assert np.array_equal(host(outs[1], pay), ref[1])


def test_array_elements_full_size(ctx):
    """2^26 array elements of 4 words (1 GiB), m=16, against the oracle."""
    import torch
    import mdiscriminator_amd as md
    n, length, m = 1 << 26, 4, 16
    d_in = W.uniform_u31_torch(n * length, 77, "cuda:0")
    x = d_in.cpu().numpy().view(np.uint32)
    d_out = torch.empty_like(d_in)
    d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
    print("arrays 2^26 x 4", flush=True)
    ctx.partition_arrays(d_in, d_out, length, md.Mask(m), d_counts=d_cnt)
    torch.cuda.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr_arrays(x, length, m, kind=oracle.MASK)
    assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
    assert first_diff(d_out.cpu().numpy().view(np.uint32), ref) < 0


def offset_view(a, off):
    """A device copy of `a` that starts `off` bytes into a larger buffer."""
    import torch
    raw = np.zeros(a.nbytes + 16, dtype=np.uint8)
    raw[off:off + a.nbytes] = np.ascontiguousarray(a).view(np.uint8)
    big = torch.from_numpy(raw).to("cuda:0")
    v = big[off:off + a.nbytes]
    return v if a.dtype == np.uint8 else v.view(dev(a[:0]).dtype)


@pytest.mark.parametrize("m", [1, 2, 3, 8, 33, 64])
def test_byte_columns_every_m(ctx, m):
    """1-byte columns: a 4-byte key with byte payloads, and records of byte
    columns only with a byte class key, at tile and wave boundaries."""
    import mdiscriminator_amd as md
    for n in (1, 3, 255, 256, 1000, 4097, 6143, 6144, 6145, 12289, 100003):
        key = W.uniform_u31(n, 700 + m)
        cls = (W.uniform_u31(n, 701 + m) % m).astype(np.uint8)
        b1 = (W.uniform_u31(n, 702 + m) & 0xFF).astype(np.uint8)
        b2 = (W.uniform_u31(n, 703 + m) & 1).astype(np.uint8)
        idx = np.arange(n, dtype=np.uint32)
        print(f"bytes m={m} n={n}", flush=True)
        check_soa(ctx, [key, b1, idx, b2], md.Mod(m), label=f"u32 key m={m} n={n}")
        check_soa(ctx, [b1, cls, b2], md.Mod(m), key_col=1, label=f"byte key m={m} n={n}")


def test_byte_columns_misaligned(ctx):
    """Byte columns, and a byte class key, whose input or output does not
    start on a 4-byte boundary."""
    import torch
    import mdiscriminator_amd as md
    n, m = 50003, 8
    key = W.uniform_u31(n, 91)
    cls = (W.uniform_u31(n, 92) % m).astype(np.uint8)
    b1 = (W.uniform_u31(n, 93) & 0xFF).astype(np.uint8)
    for off_in, off_out in ((1, 0), (0, 1), (2, 3), (3, 3), (0, 0)):
        print(f"misaligned in+{off_in} out+{off_out}", flush=True)
        cols = [key, b1, cls]
        d_in = [dev(key), offset_view(b1, off_in), offset_view(cls, off_in)]
        d_out = [torch.empty_like(d_in[0]), offset_view(np.zeros_like(b1), off_out),
                 offset_view(np.zeros_like(cls), off_out)]
        for key_col in (0, 2):
            ctx.partition_soa(d_in, d_out, md.Mod(m), key_col=key_col)
            torch.cuda.synchronize()
            ref, _ = oracle.seq_mdiscr_soa(cols, m, key_col=key_col)
            for j, (o, r) in enumerate(zip(d_out, ref)):
                assert first_diff(host(o, cols[j]), r) < 0, \
                    f"in+{off_in} out+{off_out} key {key_col} column {j}"


def test_soa_maximum_size(ctx):
    """n = MDISCR_MAX_ELEMS records of (u32 key = i, u8 payload = (i >> 3) & 0xFF),
    m = 16: record indexes cross 2^31.  Closed form: class k holds k, k + 16,
    ... in order, so key_out[off_k + j] = 16j + k and the payload follows it."""
    import torch
    import mdiscriminator_amd as md
    n, m = 0xFFFF0000, 16
    chunk = 1 << 27
    print("soa max size: generating", flush=True)
    d_key = torch.empty(n, dtype=torch.int32, device="cuda:0")
    d_pay = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        i = torch.arange(s, e, dtype=torch.int64, device="cuda:0")
        d_key[s:e] = i.to(torch.int32)
        d_pay[s:e] = ((i >> 3) & 0xFF).to(torch.uint8)
    o_key, o_pay = torch.empty_like(d_key), torch.empty_like(d_pay)
    d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
    print("soa max size: partition", flush=True)
    ctx.partition_soa([d_key, d_pay], [o_key, o_pay], md.Mask(m), d_counts=d_cnt)
    torch.cuda.synchronize()
    assert np.array_equal(d_cnt.cpu().numpy().astype(np.int64), np.full(m, n // m))
    print("soa max size: checking", flush=True)
    per = n // m
    for k in range(m):
        for j0 in range(0, per, chunk):
            j1 = min(per, j0 + chunk)
            i = torch.arange(j0, j1, dtype=torch.int64, device="cuda:0") * m + k
            a = k * per
            assert torch.equal(o_key[a + j0:a + j1], i.to(torch.int32)), f"key class {k}"
            assert torch.equal(o_pay[a + j0:a + j1], ((i >> 3) & 0xFF).to(torch.uint8)), \
                f"payload class {k}"
    del d_key, d_pay, o_key, o_pay
    torch.cuda.empty_cache()


@pytest.mark.parametrize("length", [2, 4, 8])
def test_array_elements_word_aligned_only(ctx, length):
    """Rows whose buffers start 4 bytes past a 16-byte boundary (so rows of 2,
    4 and 8 words cannot move as whole vectors), input and output each, tail
    tiles included, against the oracle."""
    import torch
    import mdiscriminator_amd as md
    for n, m in ((4097, 5), (70001, 16)):
        x = W.uniform_u31(n * length, 900 + length + n % 7)
        big_in = torch.zeros(n * length + 4, dtype=torch.int32, device="cuda:0")
        big_in[1:1 + n * length] = torch.from_numpy(x.view(np.int32)).to("cuda:0")
        d_in = big_in[1:1 + n * length]
        big_out = torch.full((n * length + 4,), -1, dtype=torch.int32, device="cuda:0")
        d_out = big_out[1:1 + n * length]
        d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
        print(f"arrays word-aligned len={length} n={n} m={m}", flush=True)
        ctx.partition_arrays(d_in, d_out, length, md.Mod(m), d_counts=d_cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.seq_mdiscr_arrays(x, length, m)
        assert first_diff(d_out.cpu().numpy().view(np.uint32), ref) < 0
        assert np.array_equal(d_cnt.cpu().numpy().view(np.uint32), ref_cnt)
        assert int(big_out[0]) == -1 and int(big_out[1 + n * length]) == -1   # nothing outside
