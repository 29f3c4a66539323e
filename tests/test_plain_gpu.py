"""Plain partition on the GPU through the C ABI, bit-exact against the oracle
(restated SeqMdiscr.h:29-49), stability included.  Mirrors the reference's
end-to-end validation (MdiscrTest_optimized_clean.cu:58-88 ->
validateOneSegment, HelpersCommon.h:317-340) with seeded inputs."""
import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff, from_dev, to_dev

pytestmark = pytest.mark.gpu

TILE = 4096  # elements per partition workgroup


def run_plain(ctx, x, discr):
    import torch
    md = __import__("mdiscriminator_amd")
    d_in = to_dev(x)
    d_out = torch.full_like(d_in, -1)
    d_cnt = torch.full((discr.m,), -1, dtype=torch.int32, device="cuda:0")
    ctx.partition_u32(d_in, d_out, discr, d_counts=d_cnt)
    torch.cuda.synchronize()
    return from_dev(d_out), from_dev(d_cnt), md


def check(ctx, x, discr, label=""):
    out, cnt, _ = run_plain(ctx, x, discr)
    ref_out, ref_cnt = oracle.seq_mdiscr(x, discr.m, discr.kind)
    assert np.array_equal(cnt, ref_cnt), f"{label} counts {cnt[:8]} vs {ref_cnt[:8]}"
    i = first_diff(out, ref_out)
    assert i < 0, f"{label} first mismatch at {i}: {out[i]} vs {ref_out[i]}"


# TILE: the ballot path's tile; QTILE: the 16-byte-load path's (up to 16
# classes) and the ballot path's from 48 classes on.
QTILE = 8192
EDGE_N = [0, 1, 2, 63, 64, 65, 1023, 1024, 1025, TILE - 1, TILE, TILE + 1, QTILE - 1, QTILE,
          QTILE + 1, 3 * TILE + 17, 2 * QTILE + 3, 100003]


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5, 7, 8, 13, 16, 31, 32, 33, 63, 64])
def test_plain_mod_edges(ctx, m):
    import mdiscriminator_amd as md
    for n in EDGE_N:
        x = W.uniform_u31(n, 1000 + m * 7 + n)
        print(f"plain mod m={m} n={n}", flush=True)
        check(ctx, x, md.Mod(m), f"m={m} n={n}")


@pytest.mark.parametrize("m", [1, 2, 4, 8, 16, 32, 64])
def test_plain_mask(ctx, m):
    import mdiscriminator_amd as md
    for n in (TILE * 5 + 3, 1 << 20):
        x = W.uniform_u31(n, 77 + m)
        print(f"plain mask m={m} n={n}", flush=True)
        check(ctx, x, md.Mask(m), f"mask m={m} n={n}")


def test_full_32bit_values(ctx):
    """Keys above 2^31 (the reference's InType is unsigned int)."""
    import mdiscriminator_amd as md
    x = W.splitmix64(300001, 5).astype(np.uint32)
    for m in (3, 8, 10, 64):
        check(ctx, x, md.Mod(m), f"full32 m={m}")


def test_degenerate_class_patterns(ctx):
    import mdiscriminator_amd as md
    n = 7 * TILE + 5
    i = np.arange(n, dtype=np.uint32)
    cases = {
        "all one class": np.full(n, 8 * 12345 + 5, dtype=np.uint32),
        "only first and last class": (i * 8) + np.where(i % 3 == 0, 0, 7).astype(np.uint32),
        "descending classes": ((n - i) // 100 % 8 + 8 * i).astype(np.uint32),
        "empty middle classes": (i * 8 + (i % 2) * 7).astype(np.uint32),
        "one element of class 7 at the end": np.where(i == n - 1, 7, 0).astype(np.uint32) + 8 * i,
    }
    for name, x in cases.items():
        check(ctx, x, md.Mod8, name)
        check(ctx, x, md.Mod(8), name + " (mod)")


def test_misaligned_pointers(ctx):
    """Input and output views that start 1..3 elements into an allocation."""
    import torch
    import mdiscriminator_amd as md
    n = 50001
    x = W.uniform_u31(n + 8, 3)
    d_all = to_dev(x)
    for off in (1, 2, 3):
        d_in = d_all[off:off + n]
        d_out_all = torch.full((n + 8,), -1, dtype=torch.int32, device="cuda:0")
        d_out = d_out_all[off:off + n]
        ctx.partition_u32(d_in, d_out, md.Mod(6))
        torch.cuda.synchronize()
        ref, _ = oracle.seq_mdiscr(x[off:off + n], 6)
        assert np.array_equal(from_dev(d_out), ref)
        assert np.all(from_dev(d_out_all[:off]) == 0xFFFFFFFF)        # no write outside
        assert np.all(from_dev(d_out_all[off + n:]) == 0xFFFFFFFF)


def test_reference_shaped_call(ctx):
    """mdiscr<Mod4>(n, num_hwd_thds, d_in, d_out) returns the counts."""
    import torch
    import mdiscriminator_amd as md
    x = W.uniform_u31(1 << 20, W.SEED_BASE + 1)
    d_in = to_dev(x)
    d_out = torch.empty_like(d_in)
    counts = ctx.mdiscr(md.Mod4, len(x), 1 << 16, d_in, d_out)
    ref, ref_cnt = oracle.seq_mdiscr(x, 4, oracle.MASK)
    assert counts == ref_cnt.tolist()
    assert np.array_equal(from_dev(d_out), ref)
    # module-level mirror on the default context
    d_out2 = torch.empty_like(d_in)
    assert md.mdiscr(md.Mod4, len(x), 1 << 16, d_in, d_out2) == ref_cnt.tolist()
    assert np.array_equal(from_dev(d_out2), ref)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 8191, 8192, 8193])
def test_one_tile_inputs(ctx, n):
    """Inputs of at most one tile take a single launch that counts its own
    tile; the counts come back through both the synchronous reference-shaped
    call and the stream-ordered one, for every class count up to 64."""
    import torch
    import mdiscriminator_amd as md
    for m in (1, 2, 3, 5, 8, 13, 16, 33, 64):
        x = W.uniform_u31(n, W.SEED_BASE + 7 * n + m)
        d_in = to_dev(x)
        for d, kind in ((md.Mod(m), oracle.MOD), (md.Mask(1 << (m.bit_length() - 1)), oracle.MASK)):
            print(f"one tile n={n} m={d.m} kind={kind}", flush=True)
            ref, ref_cnt = oracle.seq_mdiscr(x, d.m, kind)
            d_out = torch.full_like(d_in, -1)
            assert ctx.mdiscr(d, n, 1 << 16, d_in, d_out) == ref_cnt.tolist()
            assert np.array_equal(from_dev(d_out), ref)
            d_out = torch.full_like(d_in, -1)
            d_cnt = torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0")
            ctx.partition_u32(d_in, d_out, d, d_counts=d_cnt)
            torch.cuda.synchronize()
            assert np.array_equal(from_dev(d_out), ref)
            assert np.array_equal(from_dev(d_cnt), ref_cnt)


def test_sizes_from_counts_device(ctx):
    import torch
    for counts in ([3, 0, 0, 3], [0, 5, 0, 2], [4, 4, 0, 0], [1], [0, 0, 9]):
        c = np.array(counts, dtype=np.uint32)
        n = int(c.sum())
        d_c = to_dev(c)
        d_s = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
        ctx.sizes_from_counts(d_c, len(c), n, d_s)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(d_s), oracle.sizes_from_counts(c, n))


def test_invalid_arguments(ctx):
    import torch
    import mdiscriminator_amd as md
    d = torch.zeros(1000, dtype=torch.int32, device="cuda:0")
    o = torch.zeros(1000, dtype=torch.int32, device="cuda:0")
    for bad in (md.Mod(0), md.Mod(65), md.Mask(6)):
        with pytest.raises(md.MdiscrError) as e:
            ctx.partition_u32(d, o, bad)
        assert e.value.status == 1
    with pytest.raises(md.MdiscrError):
        ctx.partition_u32(d, d[10:], md.Mod4, n=990)          # overlapping output
    with pytest.raises(md.MdiscrError):                       # n beyond the limit, seen by
        ctx.partition_u32(d.data_ptr(), o.data_ptr(), md.Mod4, n=1 << 40)   # the library
    # tensors are checked before the call: strided, short, or short counts
    with pytest.raises(md.MdiscrError):
        ctx.partition_u32(d[::2], o[:500], md.Mod4)
    with pytest.raises(md.MdiscrError):
        ctx.partition_u32(d, o[:10], md.Mod4)
    with pytest.raises(md.MdiscrError):
        ctx.partition_u32(d, o, md.Mod(16), d_counts=torch.zeros(8, dtype=torch.int32,
                                                                 device="cuda:0"))
    # the context still works after rejected calls
    x = W.uniform_u31(1000, 1)
    check(ctx, x, md.Mod4, "after errors")


def test_repeated_calls_reuse_workspace(ctx):
    """Large then small then large: results stay exact as the workspace is reused."""
    import mdiscriminator_amd as md
    for n in (1 << 22, 777, 1 << 21, 5):
        x = W.uniform_u31(n, n)
        check(ctx, x, md.Mod(8), f"n={n}")


@pytest.mark.parametrize("variant", ["uniform", "zipf"])
def test_full_size_2p28_m8(ctx, variant):
    """The headline workload at its full size: 2^28 int32, m=8, bit-exact
    against the oracle, plus size-independent invariants."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 28
    seed = W.SEED_BASE + 2
    print(f"2^28 {variant}: generating", flush=True)
    if variant == "uniform":
        d_in = W.uniform_u31_torch(n, seed, "cuda:0")
    else:
        d_in = W.zipf8_values_torch(n, seed, "cuda:0")
    d_out = torch.empty_like(d_in)
    d_cnt = torch.empty(8, dtype=torch.int32, device="cuda:0")
    ctx.partition_u32(d_in, d_out, md.Mod8, d_counts=d_cnt)
    torch.cuda.synchronize()
    # invariants on the device: classes non-decreasing, counts, checksum kept
    c_out = d_out & 7
    assert bool((c_out[1:] >= c_out[:-1]).all())
    cnt = from_dev(d_cnt)
    assert np.array_equal(cnt, torch.bincount(d_in & 7, minlength=8).cpu().numpy())
    assert int(d_out.to(torch.int64).sum()) == int(d_in.to(torch.int64).sum())
    # bit-exact against the oracle
    x = from_dev(d_in)
    ref, ref_cnt = oracle.seq_mdiscr(x, 8, oracle.MASK)
    assert np.array_equal(cnt, ref_cnt)
    out = from_dev(d_out)
    i = first_diff(out, ref)
    assert i < 0, f"first mismatch at {i}"


@pytest.mark.parametrize("discr", ["mask64", "mod33", "mod47"])
def test_full_size_2p28_many_classes(ctx, discr):
    """2^28 int32 over 33..64 classes (the two-digit path: low 3 class bits,
    then the high bits, both ranked in LDS), bit-exact against the oracle,
    with the class offsets."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 28
    m = int(discr[4:] if discr.startswith("mask") else discr[3:])
    d, kind = (md.Mask(m), oracle.MASK) if discr.startswith("mask") else (md.Mod(m), oracle.MOD)
    print(f"2^28 {discr}: generating", flush=True)
    d_in = W.uniform_u31_torch(n, W.SEED_BASE + 3 + m, "cuda:0")
    d_out = torch.empty_like(d_in)
    d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
    d_off = torch.full((m + 1,), -1, dtype=torch.int64, device="cuda:0")
    ctx.partition_u32_ex(d_in, d_out, d, counts=d_cnt, offsets=d_off)
    torch.cuda.synchronize()
    x = from_dev(d_in)
    ref, ref_cnt = oracle.seq_mdiscr(x, m, kind)
    assert np.array_equal(from_dev(d_cnt), ref_cnt)
    off = d_off.cpu().numpy()
    assert off[0] == 0 and off[m] == n
    assert np.array_equal(np.diff(off), ref_cnt.astype(np.int64))
    i = first_diff(from_dev(d_out), ref)
    assert i < 0, f"{discr}: first mismatch at {i}"


@pytest.mark.parametrize("m", [8, 7])
def test_maximum_size(ctx, m):
    """n = MDISCR_MAX_ELEMS (0xFFFF0000, about 4.29 G elements, 17 GB each way)
    with x[i] = i, so element offsets, tile indexes and values cross 2^31 and
    run up to the 32-bit limit.  The exact output has a closed form: class k
    holds k, k + m, k + 2m, ... in order, so out[off_k + j] = m*j + k, checked
    on the device chunk by chunk (placement, stability and the permutation)."""
    import torch
    import mdiscriminator_amd as md
    n = 0xFFFF0000
    chunk = 1 << 27
    print(f"max size m={m}: generating", flush=True)
    d_in = torch.empty(n, dtype=torch.int32, device="cuda:0")
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        d_in[s:e] = torch.arange(s, e, dtype=torch.int64, device="cuda:0").to(torch.int32)
    d_out = torch.empty_like(d_in)
    d_cnt = torch.empty(m, dtype=torch.int32, device="cuda:0")
    discr = md.Mask(m) if m & (m - 1) == 0 else md.Mod(m)
    print(f"max size m={m}: partition", flush=True)
    ctx.partition_u32(d_in, d_out, discr, d_counts=d_cnt)
    torch.cuda.synchronize()
    cnt = from_dev(d_cnt).astype(np.int64)
    want = np.array([(n - k + m - 1) // m for k in range(m)], dtype=np.int64)
    assert np.array_equal(cnt, want)
    print(f"max size m={m}: checking", flush=True)
    off = 0
    for k in range(m):
        for j0 in range(0, int(want[k]), chunk):
            j1 = min(int(want[k]), j0 + chunk)
            exp = (torch.arange(j0, j1, dtype=torch.int64, device="cuda:0") * m + k).to(torch.int32)
            assert torch.equal(d_out[off + j0:off + j1], exp), f"class {k}, j in [{j0}, {j1})"
        off += int(want[k])
    del d_in, d_out
    torch.cuda.empty_cache()
