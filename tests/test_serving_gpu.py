"""Batched independent requests and the continuous-batching server on the GPU.
Every result must be bit-identical alone (one plain call) and in the mix
(batched launch / served alongside other requests), and equal to the oracle."""
import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff, from_dev, to_dev

pytestmark = pytest.mark.gpu


def make_requests(seed, count, lo=0, hi=200000):
    """Mixed sizes (log-uniform, some empty) and m in {1,2,3,4,8,16,64}."""
    r = W.splitmix64(3 * count, seed)
    sizes = []
    for i in range(count):
        u = float(r[3 * i] >> np.uint64(11)) / float(1 << 53)
        n = int(round(np.exp(np.log(lo + 1) + u * (np.log(hi + 1) - np.log(lo + 1))) - 1))
        sizes.append(n if i % 11 else 0)
    ms = [(1, 2, 3, 4, 8, 16, 64)[int(r[3 * i + 1] % np.uint64(7))] for i in range(count)]
    xs = [W.uniform_u31(n, seed * 1000 + i) for i, n in enumerate(sizes)]
    return xs, ms


def test_batch_matches_alone_and_oracle(ctx):
    import torch
    import mdiscriminator_amd as md
    xs, ms = make_requests(5, 60)
    d_in = [to_dev(x) if len(x) else torch.empty(0, dtype=torch.int32, device="cuda:0")
            for x in xs]
    d_out = [torch.full_like(t, -1) for t in d_in]
    d_cnt = [torch.full((m,), -1, dtype=torch.int32, device="cuda:0") for m in ms]
    reqs = [(d_in[i], d_out[i], md.Mod(ms[i]), d_cnt[i]) for i in range(len(xs))]
    ctx.partition_batch(reqs)
    torch.cuda.synchronize()
    for i, x in enumerate(xs):
        ref, ref_cnt = oracle.seq_mdiscr(x, ms[i])
        assert first_diff(from_dev(d_out[i]), ref) < 0, f"request {i} (n={len(x)}, m={ms[i]})"
        assert np.array_equal(from_dev(d_cnt[i]), ref_cnt)
        if len(x):   # and alone, through the plain path
            alone = torch.empty_like(d_in[i])
            ctx.partition_u32(d_in[i], alone, md.Mod(ms[i]))
            torch.cuda.synchronize()
            assert torch.equal(alone, d_out[i])


@pytest.mark.parametrize("cfg", list(range(24)))
def test_batch_every_forced_config(ctx, cfg):
    """MDISCR_SEG_CONFIG applies to batch calls too: every configuration gives
    the oracle's result for a power-of-two mix and for a mix of other class
    counts, with a many-tile request."""
    import os
    import torch
    import mdiscriminator_amd as md
    xs, ms = make_requests(40 + cfg, 24, hi=300000)
    for label, mk in (("pow2", lambda m: md.Mask((2, 4, 8, 16)[m % 4])), ("mod", md.Mod)):
        discrs = [mk(m) for m in ms]
        print(f"batch cfg={cfg} {label}", flush=True)
        d_in = [to_dev(x) if len(x) else torch.empty(0, dtype=torch.int32, device="cuda:0")
                for x in xs]
        d_out = [torch.full_like(t, -1) for t in d_in]
        d_cnt = [torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0") for d in discrs]
        os.environ["MDISCR_SEG_CONFIG"] = str(cfg)
        try:
            ctx.partition_batch([(d_in[i], d_out[i], discrs[i], d_cnt[i]) for i in range(len(xs))])
            torch.cuda.synchronize()
        finally:
            del os.environ["MDISCR_SEG_CONFIG"]
        for i, x in enumerate(xs):
            d = discrs[i]
            ref, ref_cnt = oracle.seq_mdiscr(x, d.m, oracle.MASK if d.kind == md.DISCR_MASK
                                             else oracle.MOD)
            assert first_diff(from_dev(d_out[i]), ref) < 0, f"{label} request {i} (n={len(x)})"
            assert np.array_equal(from_dev(d_cnt[i]), ref_cnt)


def test_batch_sequence_on_stream(ctx):
    """Successive batches on one non-null stream (the server's case), reusing
    the context's pinned descriptor staging: other tensors, request counts,
    sizes, class counts and contents in turn, each equal to the oracle,
    counts included."""
    import torch
    import mdiscriminator_amd as md
    s = torch.cuda.Stream()
    plans = [(7, 3, 50000), (7, 4, 120000), (5, 5, 80000), (70, 6, 30000), (7, 7, 90000),
             (3, 8, 2000000), (7, 9, 60000), (7, 10, 60000), (1, 11, 5000)]
    for count, seed, hi in plans:
        xs, ms = make_requests(100 + seed, count, hi=hi)
        print(f"batch on stream: {count} requests, seed {seed}", flush=True)
        d_in = [to_dev(x) if len(x) else torch.empty(0, dtype=torch.int32, device="cuda:0")
                for x in xs]
        d_out = [torch.full_like(t, -1) for t in d_in]
        d_cnt = [torch.full((m,), -1, dtype=torch.int32, device="cuda:0") for m in ms]
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            ctx.partition_batch([(d_in[i], d_out[i], md.Mod(ms[i]), d_cnt[i])
                                 for i in range(count)], stream=s)
        s.synchronize()
        for i, x in enumerate(xs):
            ref, ref_cnt = oracle.seq_mdiscr(x, ms[i])
            assert first_diff(from_dev(d_out[i]), ref) < 0, f"seed {seed} request {i}"
            assert np.array_equal(from_dev(d_cnt[i]), ref_cnt), f"seed {seed} request {i}"


def test_batch_power_of_two_mix_uses_mask_kernels(ctx):
    import torch
    import mdiscriminator_amd as md
    xs, _ = make_requests(6, 40, lo=1000, hi=3000000)
    ms = [(2, 4, 8, 16)[i % 4] for i in range(len(xs))]
    d_in = [to_dev(x) if len(x) else torch.empty(0, dtype=torch.int32, device="cuda:0")
            for x in xs]
    d_out = [torch.empty_like(t) for t in d_in]
    ctx.partition_batch([(d_in[i], d_out[i], md.Mask(ms[i])) for i in range(len(xs))])
    torch.cuda.synchronize()
    for i, x in enumerate(xs):
        ref, _ = oracle.seq_mdiscr(x, ms[i], oracle.MASK)
        assert first_diff(from_dev(d_out[i]), ref) < 0


# (waves per workgroup, elements per lane) of the 24 segmented / batch
# candidates, in MDISCR_SEG_CONFIG order.
SEG_CANDIDATES = [(4, 16), (4, 16), (8, 8), (8, 8), (4, 8), (8, 16), (16, 4), (4, 32),
                  (4, 4), (8, 4), (8, 32), (16, 8), (16, 16), (16, 32), (4, 4), (4, 8),
                  (4, 32), (8, 4), (8, 16), (8, 32), (16, 4), (16, 8), (16, 16), (16, 32)]


@pytest.mark.parametrize("cfg", list(range(24)))
def test_batch_of_one_tile_requests(ctx, cfg):
    """Batches whose requests each fit one tile of the configuration launch no
    count or scan: requests at 0, 1 and the tile and wave-share boundaries,
    then the same batch with one request a key past the tile (count rows
    again).  Outputs and class counts (pre-filled with -1) equal the oracle."""
    import os
    import torch
    import mdiscriminator_amd as md
    nw, r = SEG_CANDIDATES[cfg]
    tile, wt = nw * 64 * r, 64 * r
    for extra in ([], [tile + 1]):
        sizes = [0, 1, 63, wt - 1, wt, wt + 1, tile - 1, tile, 5, tile] + extra
        ms = [(2, 3, 4, 7, 8, 16, 64)[i % 7] for i in range(len(sizes))]
        discrs = [md.Mask(m) if m & (m - 1) == 0 else md.Mod(m) for m in ms]
        xs = [W.uniform_u31(n, 9000 + 97 * cfg + i) for i, n in enumerate(sizes)]
        print(f"one-tile batch cfg={cfg} tile={tile} extra={extra}", flush=True)
        d_in = [to_dev(x) if len(x) else torch.empty(0, dtype=torch.int32, device="cuda:0")
                for x in xs]
        d_out = [torch.full_like(t, -1) for t in d_in]
        d_cnt = [torch.full((d.m,), -1, dtype=torch.int32, device="cuda:0") for d in discrs]
        os.environ["MDISCR_SEG_CONFIG"] = str(cfg)
        try:
            ctx.partition_batch([(d_in[i], d_out[i], discrs[i], d_cnt[i])
                                 for i in range(len(xs))])
            torch.cuda.synchronize()
        finally:
            del os.environ["MDISCR_SEG_CONFIG"]
        for i, x in enumerate(xs):
            d = discrs[i]
            ref, ref_cnt = oracle.seq_mdiscr(x, d.m, oracle.MASK if d.kind == md.DISCR_MASK
                                             else oracle.MOD)
            assert first_diff(from_dev(d_out[i]), ref) < 0, f"request {i} (n={len(x)})"
            assert np.array_equal(from_dev(d_cnt[i]), ref_cnt), f"request {i} (n={len(x)})"


def test_server_results_same_alone_and_in_mix():
    import torch
    import mdiscriminator_amd as md
    xs, ms = make_requests(7, 150, hi=2000000)
    d_in = [to_dev(x) if len(x) else torch.empty(0, dtype=torch.int32, device="cuda:0")
            for x in xs]
    srv = md.Server(budget_bytes=1 << 30, nstreams=3, max_batch_requests=32)
    try:
        tickets = [srv.submit(d_in[i], md.Mod(ms[i])) for i in range(len(xs))]
        ctx = md.default_context(0)
        for i, t in enumerate(tickets):
            ptr, cnt, lat = srv.wait(t)
            got = srv.output_tensor(t, ptr)
            srv.release(t)
            alone = torch.empty_like(d_in[i])
            if len(xs[i]):
                ctx.partition_u32(d_in[i], alone, md.Mod(ms[i]))
            torch.cuda.synchronize()
            assert torch.equal(got, alone), f"request {i}"
            ref, ref_cnt = oracle.seq_mdiscr(xs[i], ms[i])
            assert first_diff(from_dev(got), ref) < 0
            assert cnt == ref_cnt.tolist()
            assert lat >= 0
        st = srv.stats()
        assert st["completed"] == len(xs)
        assert st["batches"] >= 1
        assert st["p99_s"] >= st["p50_s"] >= 0
    finally:
        srv.close()


def test_server_admission_under_tiny_budget():
    """A budget that holds only a few outputs: requests wait for admission and
    complete as earlier outputs are released, in FIFO order."""
    import torch
    import mdiscriminator_amd as md
    n = 100000
    xs = [W.uniform_u31(n, 50 + i) for i in range(12)]
    d_in = [to_dev(x) for x in xs]
    budget = 3 * (1 << 19)  # three 512 KiB size classes (each output needs 400 KB)
    srv = md.Server(budget_bytes=budget, nstreams=2, max_batch_requests=8)
    try:
        tickets = [srv.submit(t, md.Mod4) for t in d_in]
        for i, t in enumerate(tickets):
            ptr, cnt, _ = srv.wait(t)
            got = srv.output_tensor(t, ptr)
            srv.release(t)
            ref, _ = oracle.seq_mdiscr(xs[i], 4)
            assert first_diff(from_dev(got), ref) < 0
        assert srv.stats()["peak_bytes"] <= budget
    finally:
        srv.close()


def test_server_rejects_and_backpressure():
    import torch
    import mdiscriminator_amd as md
    srv = md.Server(budget_bytes=1 << 20, nstreams=1, max_batch_requests=4, max_pending=3)
    try:
        d = torch.zeros(1 << 18, dtype=torch.int32, device="cuda:0")   # needs 1 MiB
        big = torch.zeros(1 << 19, dtype=torch.int32, device="cuda:0")  # needs 2 MiB
        with pytest.raises(md.MdiscrError) as e:
            srv.submit(big, md.Mod4)            # can never fit the budget
        assert e.value.status == 2
        with pytest.raises(md.MdiscrError):
            srv.submit(d, md.Mod(0))            # bad discriminator
        t = [srv.submit(d, md.Mod4) for _ in range(3)]   # one admitted, two queued
        busy = None
        for _ in range(50):
            try:
                t.append(srv.submit(d, md.Mod4))
            except md.MdiscrError as err:
                busy = err
                break
        assert busy is not None and busy.status == 4       # EBUSY
        for x in t:
            srv.wait(x)
            srv.release(x)
    finally:
        srv.close()
