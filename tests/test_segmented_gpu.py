"""Batched segmented partition on the GPU, bit-exact against the oracle applied
per segment (MdiscrSegmTest.cu:21-38), through both entry points: segment
offsets (mdiscr_partition_segmented) and the reference's flag arrays
(mdiscr_segm_flags, i.e. mdiscrSegm(n, in, segment_sizes, out, new_sizes))."""
import json
import os

import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff, from_dev, to_dev

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat.json")))
NUM_CONFIGS = 24


def run_offsets(ctx, x, off, discr, stream=None, counts=True):
    import torch
    d_in = to_dev(x)
    d_out = torch.full_like(d_in, -1)
    d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
    nseg = len(off) - 1
    d_cnt = torch.full((max(nseg * discr.m, 1),), -1, dtype=torch.int32, device="cuda:0") \
        if counts else None
    ctx.partition_segmented(d_in, d_out, d_off, discr, h_seg_off=off.astype(np.uint64),
                            d_seg_counts=d_cnt, stream=stream)
    torch.cuda.synchronize()
    cnt = from_dev(d_cnt)[:nseg * discr.m].reshape(nseg, discr.m) if counts else None
    return from_dev(d_out), cnt


def check_offsets(ctx, x, off, discr, label="", stream=None):
    out, cnt = run_offsets(ctx, x, off, discr, stream=stream)
    ref, ref_cnt = oracle.seq_mdiscr_segm(x, off, discr.m, discr.kind)
    i = first_diff(out, ref)
    assert i < 0, f"{label}: first mismatch at {i} ({out[i]} vs {ref[i]})"
    assert np.array_equal(cnt, ref_cnt), f"{label}: segment counts differ"


def lengths_mix(seed, nseg, median, sigma=1.5, zeros=True):
    L = W.lognormal_lengths(nseg, seed, median=median, sigma=sigma, lo=0 if zeros else 1)
    if zeros:
        L[::17] = 0
    return L


def test_reference_kat_flags(ctx):
    """haskell_code/testsSegm.hs inputs; expected from the reference SeqMdiscr per segment."""
    import torch
    import mdiscriminator_amd as md
    case = KAT["segmented"][0]
    x = np.array(case["input"], dtype=np.uint32)
    f = np.array(case["segment_sizes"], dtype=np.int32)
    d_in = to_dev(x)
    d_out = torch.empty_like(d_in)
    d_new = torch.full_like(d_in, -7)
    md.mdiscr_segm(md.Mod(case["m"]), len(x), d_in, torch.from_numpy(f).to("cuda:0"), d_out,
                   d_new)
    assert from_dev(d_out).tolist() == case["out"]
    assert d_new.cpu().tolist() == case["new_sizes"]


@pytest.mark.parametrize("m", [1, 2, 3, 4, 7, 8, 16, 33, 64])
def test_segmented_lognormal_mix(ctx, m):
    import mdiscriminator_amd as md
    for median, nseg in ((30, 400), (900, 300), (5000, 120), (40000, 20)):
        L = lengths_mix(m * 100 + median, nseg, median)
        off = W.offsets_from_lengths(L)
        x = W.uniform_u31(int(off[-1]), median + m)
        print(f"segmented m={m} median={median} n={off[-1]}", flush=True)
        check_offsets(ctx, x, off, md.Mod(m), f"m={m} median={median}")


@pytest.mark.parametrize("cfg", list(range(NUM_CONFIGS)))
def test_every_tuned_config(ctx, cfg):
    """Every candidate configuration of the autotuner gives the exact result,
    for power-of-two (mask) and other (magic-divide mod) class counts up to 64."""
    import mdiscriminator_amd as md
    os.environ["MDISCR_SEG_CONFIG"] = str(cfg)
    try:
        for discr in (md.Mask(4), md.Mod(5), md.Mask(16), md.Mod(33), md.Mask(64)):
            m = discr.m
            L = lengths_mix(cfg * 31 + m, 300, 2000)
            L[5] = 70000    # a many-tile segment
            L[6] = 1        # and neighbours of every size
            L[7] = 2047
            L[8] = 40000    # a tile-multiple of every candidate (1024 * 2^k)
            L[9] = 32768
            off = W.offsets_from_lengths(L)
            x = W.uniform_u31(int(off[-1]), cfg + m)
            print(f"segmented cfg={cfg} m={m} n={off[-1]}", flush=True)
            check_offsets(ctx, x, off, discr, f"cfg={cfg} m={m}")
    finally:
        del os.environ["MDISCR_SEG_CONFIG"]


def test_autotuner_buckets_and_cache(ctx, tmp_path):
    """A call of a new (kind, m, size) bucket times every candidate on its own
    data, keeps the fastest and records it, with its name, in
    MDISCR_TUNE_CACHE; the result is exact whichever candidate won."""
    import mdiscriminator_amd as md
    cache = tmp_path / "tune.txt"
    os.environ["MDISCR_TUNE_CACHE"] = str(cache)
    try:
        c2 = md.Context(0)   # a fresh tuning table
        try:
            for discr in (md.Mod(33), md.Mask(8)):
                L = lengths_mix(400 + discr.m, 500, 3000)
                off = W.offsets_from_lengths(L)
                x = W.uniform_u31(int(off[-1]), 7 + discr.m)
                print(f"autotune m={discr.m} n={off[-1]}", flush=True)
                check_offsets(c2, x, off, discr, f"tuned m={discr.m}")
        finally:
            c2.close()
    finally:
        del os.environ["MDISCR_TUNE_CACHE"]
    lines = cache.read_text().splitlines()
    keys = {ln.split("=")[0] for ln in lines}
    assert any(k.startswith("u32 mod m33 ") for k in keys), lines
    assert any(k.startswith("u32 mask m8 ") for k in keys), lines
    for ln in lines:
        idx, name = ln.split("=")[1].split()
        assert 0 <= int(idx) < NUM_CONFIGS and name.startswith("b")


def test_segment_shapes(ctx):
    import mdiscriminator_amd as md
    shapes = {
        "one segment": np.array([123457]),
        "all length one": np.ones(5000, dtype=np.int64),
        "all empty but one": np.array([0, 0, 0, 9, 0, 0]),
        "exact tiles": np.array([4096, 8192, 1024, 64, 4096 * 5]),
        "tile +- 1": np.array([4095, 4097, 1023, 1025, 63, 65, 8191, 8193]),
    }
    for name, L in shapes.items():
        off = W.offsets_from_lengths(L)
        x = W.uniform_u31(int(off[-1]), len(L))
        for m in (4, 5):
            check_offsets(ctx, x, off, md.Mod(m), f"{name} m={m}")


def test_graph_replay_sees_new_contents(ctx):
    """A call on a non-null stream is captured once and replayed; a replay
    with new input contents in the same buffers gives the new result."""
    import torch
    import mdiscriminator_amd as md
    L = lengths_mix(5, 500, 3000)
    off = W.offsets_from_lengths(L)
    n = int(off[-1])
    s = torch.cuda.Stream()
    d_in = torch.empty(n, dtype=torch.int32, device="cuda:0")
    d_out = torch.empty_like(d_in)
    d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
    for rep in range(4):
        x = W.uniform_u31(n, 1000 + rep)
        d_in.copy_(torch.from_numpy(x.view(np.int32)))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            ctx.partition_segmented(d_in, d_out, d_off, md.Mod4, stream=s)
        s.synchronize()
        torch.cuda.synchronize()
        ref, _ = oracle.seq_mdiscr_segm(x, off, 4)
        assert first_diff(from_dev(d_out), ref) < 0, f"replay {rep}"


def test_graph_replay_after_workspace_growth():
    """A call captured into a graph, then a larger call on the same context
    (which reallocates the shared workspace), then the first call again: the
    graph captured against the old workspace must not be replayed."""
    import torch
    import mdiscriminator_amd as md
    c = md.Context()   # its own, small workspace
    try:
        L = lengths_mix(7, 300, 2000)
        off = W.offsets_from_lengths(L)
        n = int(off[-1])
        s = torch.cuda.Stream()
        d_in = torch.empty(n, dtype=torch.int32, device="cuda:0")
        d_out = torch.empty_like(d_in)
        d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
        big = W.uniform_u31(1 << 24, 77)
        d_big = to_dev(big)
        d_big_out = torch.empty_like(d_big)
        for rep in range(3):
            x = W.uniform_u31(n, 2000 + rep)
            d_in.copy_(torch.from_numpy(x.view(np.int32)))
            torch.cuda.synchronize()
            print(f"workspace growth replay {rep}", flush=True)
            with torch.cuda.stream(s):
                c.partition_segmented(d_in, d_out, d_off, md.Mod4, stream=s)
            s.synchronize()
            ref, _ = oracle.seq_mdiscr_segm(x, off, 4)
            assert first_diff(from_dev(d_out), ref) < 0, f"call {rep}"
            if rep == 0:
                # grows the context's workspace well past the segmented call's
                c.partition_u32(d_big, d_big_out, md.Mod(7))
                torch.cuda.synchronize()
                ref_big, _ = oracle.seq_mdiscr(big, 7)
                assert first_diff(from_dev(d_big_out), ref_big) < 0
    finally:
        c.close()


def test_flags_interface_random(ctx):
    import torch
    import mdiscriminator_amd as md
    L = lengths_mix(9, 3000, 700, zeros=False)
    n = int(L.sum())
    x = W.uniform_u31(n, 10)
    f = W.flags_from_lengths(L, n)
    d_in = to_dev(x)
    d_out = torch.empty_like(d_in)
    d_new = torch.empty_like(d_in)
    ctx.segm_flags(md.Mod(6), n, d_in, torch.from_numpy(f).to("cuda:0"), d_out, d_new)
    torch.cuda.synchronize()
    ref, ref_new = oracle.segm_flags(x, f, 6)
    assert first_diff(from_dev(d_out), ref) < 0
    assert np.array_equal(d_new.cpu().numpy(), ref_new)


def test_flags_interface_rejects_bad_flags(ctx):
    import torch
    import mdiscriminator_amd as md
    x = to_dev(np.arange(10, dtype=np.uint32))
    o = torch.empty_like(x)
    ns = torch.empty_like(x)
    for bad in ([3, 0, 0, 4, 0, 0, 0, 5, 0, 0],    # last segment runs past n
                [0, 3, 0, 0, 6, 0, 0, 0, 0, 0],    # no segment starts at 0
                [2, 0, 3, 0, 0, 0, 4, 0, 0, 0],    # length disagrees with next start
                [10, 0, 0, 0, 0, -1, 0, 0, 0, 0]):  # negative flag
        with pytest.raises(md.MdiscrError):
            ctx.segm_flags(md.Mod4, 10, x, torch.tensor(bad, dtype=torch.int32,
                                                         device="cuda:0"), o, ns)


def test_segmented_invalid_offsets(ctx):
    import torch
    import mdiscriminator_amd as md
    d_in = to_dev(np.arange(100, dtype=np.uint32))
    d_out = torch.empty_like(d_in)
    bad = np.array([0, 50, 40, 100], dtype=np.uint64)
    with pytest.raises(md.MdiscrError):
        ctx.partition_segmented(d_in, d_out, torch.from_numpy(bad.astype(np.int64)).to("cuda:0"),
                                md.Mod4, h_seg_off=bad)


def test_segmented_unchecked_bad_offsets_stay_in_bounds(ctx):
    """Device offsets that decrease, run past n or overlap, with no host copy
    to check them: each segment is clamped into [0, n), and a plan whose
    segments add up to more than the workspace holds is left empty.  The
    call returns, and nothing outside d_out is written (sentinels around it)."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 20
    x = W.uniform_u31(n, 77)
    cases = {
        "decreasing": np.array([0, 500000, 400000, n], dtype=np.uint64),
        "past n": np.array([0, 300000, n + 5000, 3 * n], dtype=np.uint64),
        "overlapping": np.array([0, n, 0, n], dtype=np.uint64),
        "overlapping beyond the queue": np.array([0, n] * 6 + [0], dtype=np.uint64),
    }
    for name, off in cases.items():
        print(f"unchecked offsets: {name}", flush=True)
        guard = 4096
        d_all = torch.full((n + 2 * guard,), -7, dtype=torch.int32, device="cuda:0")
        d_out = d_all[guard:guard + n]
        d_in = to_dev(x)
        d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
        nseg = len(off) - 1
        d_cnt = torch.full((nseg * 4 + 2 * guard,), -7, dtype=torch.int32, device="cuda:0")
        ctx.partition_segmented(d_in, d_out, d_off, md.Mod4,
                                d_seg_counts=d_cnt[guard:guard + nseg * 4])
        torch.cuda.synchronize()
        assert torch.all(d_all[:guard] == -7) and torch.all(d_all[guard + n:] == -7), name
        assert torch.all(d_cnt[:guard] == -7) and torch.all(d_cnt[guard + nseg * 4:] == -7), name
    # the context works normally afterwards
    off = W.offsets_from_lengths(lengths_mix(5, 200, 3000))
    check_offsets(ctx, W.uniform_u31(int(off[-1]), 6), off, md.Mod4, "after bad offsets")


def test_full_size_10k_lognormal(ctx):
    """The batched workload at its stated parameters: 10,000 segments with
    lognormal lengths (median 4096, sigma 1.5, clipped to [1, 4Mi]), m=4."""
    import torch
    import mdiscriminator_amd as md
    L = W.lognormal_lengths(10000, W.SEED_BASE + 4)
    off = W.offsets_from_lengths(L)
    n = int(off[-1])
    print(f"10k lognormal: n={n}", flush=True)
    d_in = W.uniform_u31_torch(n, W.SEED_BASE + 4, "cuda:0")
    d_out = torch.empty_like(d_in)
    d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
    d_cnt = torch.empty(10000 * 4, dtype=torch.int32, device="cuda:0")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ctx.partition_segmented(d_in, d_out, d_off, md.Mod4, d_seg_counts=d_cnt, stream=s)
        ctx.partition_segmented(d_in, d_out, d_off, md.Mod4, d_seg_counts=d_cnt, stream=s)
    s.synchronize()
    x = from_dev(d_in)
    ref, ref_cnt = oracle.seq_mdiscr_segm(x, off, 4)
    assert first_diff(from_dev(d_out), ref) < 0
    assert np.array_equal(from_dev(d_cnt).reshape(10000, 4), ref_cnt)


def test_full_size_10k_lognormal_flags(ctx):
    """The same 10,000 lognormal segments given as the reference's start-flag
    array (segment_sizes, MdiscrSegmTest.cu:52-53); out and new_sizes against
    the oracle's restatement of the flag form."""
    import torch
    import mdiscriminator_amd as md
    L = W.lognormal_lengths(10000, W.SEED_BASE + 4)
    n = int(L.sum())
    f = W.flags_from_lengths(L, n)
    print(f"10k lognormal flags: n={n}", flush=True)
    d_in = W.uniform_u31_torch(n, W.SEED_BASE + 5, "cuda:0")
    d_out = torch.empty_like(d_in)
    d_new = torch.empty_like(d_in)
    ctx.segm_flags(md.Mod4, n, d_in, torch.from_numpy(f).to("cuda:0"), d_out, d_new)
    torch.cuda.synchronize()
    x = from_dev(d_in)
    ref, ref_new = oracle.segm_flags(x, f, 4)
    assert first_diff(from_dev(d_out), ref) < 0
    assert first_diff(d_new.cpu().numpy(), ref_new) < 0


def test_segmented_maximum_size(ctx):
    """n = MDISCR_MAX_ELEMS in three segments, the first past 2^31 elements long,
    x[i] = i, m = 4: in segment [a, b), class k holds the i = k (mod 4) in
    order, so its j-th element is first_k + 4j, checked on the device."""
    import torch
    import mdiscriminator_amd as md
    n, m = 0xFFFF0000, 4
    off = np.array([0, (1 << 31) + 5, 0xF0000003, n], dtype=np.uint64)
    chunk = 1 << 27
    print("segmented max size: generating", flush=True)
    d_in = torch.empty(n, dtype=torch.int32, device="cuda:0")
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        d_in[s:e] = torch.arange(s, e, dtype=torch.int64, device="cuda:0").to(torch.int32)
    d_out = torch.empty_like(d_in)
    d_off = torch.from_numpy(off.astype(np.int64)).to("cuda:0")
    d_cnt = torch.empty(3 * m, dtype=torch.int32, device="cuda:0")
    print("segmented max size: partition", flush=True)
    ctx.partition_segmented(d_in, d_out, d_off, md.Mod4, h_seg_off=off, d_seg_counts=d_cnt)
    torch.cuda.synchronize()
    cnt = d_cnt.cpu().numpy().astype(np.int64).reshape(3, m)
    print("segmented max size: checking", flush=True)
    for sg in range(3):
        a, b = int(off[sg]), int(off[sg + 1])
        pos = a
        for k in range(m):
            first = a + ((k - a) % m)
            count = (b - first + m - 1) // m if first < b else 0
            assert cnt[sg, k] == count, f"segment {sg} class {k} count"
            for j0 in range(0, count, chunk):
                j1 = min(count, j0 + chunk)
                exp = (torch.arange(j0, j1, dtype=torch.int64, device="cuda:0") * m + first)
                assert torch.equal(d_out[pos + j0:pos + j1], exp.to(torch.int32)), \
                    f"segment {sg} class {k}"
            pos += count
        assert pos == b
    del d_in, d_out
    torch.cuda.empty_cache()
