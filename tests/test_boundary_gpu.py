"""Boundary behaviour of the C ABI on the GPU: refusal under stream capture,
the offsets output (device and host), every input/output alignment of the
plain path, and withdrawing a queued serving request.  Results are compared
with the oracle (restated SeqMdiscr.h:29-49) bit for bit."""
import ctypes
import threading

import numpy as np
import pytest

import oracle
from mdiscriminator_amd import workloads as W
from tests.helpers import first_diff, from_dev, to_dev

pytestmark = pytest.mark.gpu

TILE = 4096
EUNSUPPORTED = 5


def _hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.hipGraphGetNodes.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_size_t)]
    lib.hipGraphDestroy.argtypes = [ctypes.c_void_p]
    return lib


def test_calls_under_capture_are_refused(ctx):
    """A stream being captured into a graph is refused before anything is
    allocated or launched: every entry point returns MDISCR_EUNSUPPORTED and
    the captured graph holds no node.  The same stream works afterwards."""
    import torch
    import mdiscriminator_amd as md
    n = 3 * TILE + 5
    x = W.uniform_u31(n, 91)
    d_in = to_dev(x)
    d_out = torch.full_like(d_in, -1)
    d_cnt = torch.full((8,), -1, dtype=torch.int32, device="cuda:0")
    d_off = torch.zeros(2, dtype=torch.int64, device="cuda:0")
    d_off[1] = n
    d_sizes = torch.empty_like(d_in)
    cols = [d_in, torch.arange(n, dtype=torch.int32, device="cuda:0")]
    outs = [torch.empty_like(c) for c in cols]
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    hip = _hip()
    calls = {
        "partition_u32": lambda: ctx.partition_u32(d_in, d_out, md.Mod8, d_counts=d_cnt, stream=s),
        "partition_u32_ex": lambda: ctx.partition_u32_ex(d_in, d_out, md.Mod(7), counts=d_cnt,
                                                         stream=s),
        "filter_u32": lambda: ctx.filter_u32(d_in, d_out, md.Pred.lt(1 << 30), stream=s),
        "partition_soa": lambda: ctx.partition_soa(cols, outs, md.Mask(16), stream=s),
        "partition_arrays": lambda: ctx.partition_arrays(d_in[:n - 1], d_out[:n - 1], 2,
                                                         md.Mod4, n=(n - 1) // 2, stream=s),
        "partition_segmented": lambda: ctx.partition_segmented(d_in, d_out, d_off, md.Mod4,
                                                               stream=s),
        "partition_batch": lambda: ctx.partition_batch([(d_in, d_out, md.Mod4)], stream=s),
        "sizes_from_counts": lambda: ctx.sizes_from_counts(d_cnt, 8, n, d_sizes, stream=s),
    }
    print("capture: begin (thread-local mode)", flush=True)
    assert hip.hipStreamBeginCapture(ctypes.c_void_p(s.cuda_stream), 1) == 0
    refused = {}
    try:
        for name, fn in calls.items():
            try:
                fn()
                refused[name] = None
            except md.MdiscrError as e:
                refused[name] = e.status
    finally:
        graph = ctypes.c_void_p()
        rc = hip.hipStreamEndCapture(ctypes.c_void_p(s.cuda_stream), ctypes.byref(graph))
    print(f"capture: end rc={rc} statuses={refused}", flush=True)
    assert rc == 0
    nodes = ctypes.c_size_t(0)
    assert hip.hipGraphGetNodes(graph, None, ctypes.byref(nodes)) == 0
    hip.hipGraphDestroy(graph)
    assert all(v == EUNSUPPORTED for v in refused.values()), refused
    assert nodes.value == 0
    assert torch.all(d_out == -1)
    # the stream is usable again, and the result is the oracle's
    ctx.partition_u32(d_in, d_out, md.Mod8, d_counts=d_cnt, stream=s)
    s.synchronize()
    ref, ref_cnt = oracle.seq_mdiscr(x, 8, oracle.MASK)
    assert first_diff(from_dev(d_out), ref) < 0
    assert np.array_equal(from_dev(d_cnt), ref_cnt)


def _ref_offsets(cnt, n):
    return np.concatenate([[0], np.cumsum(cnt.astype(np.uint64))]).astype(np.uint64)


@pytest.mark.parametrize("m", [1, 3, 8, 16, 33, 64])
def test_offsets_device_and_host(ctx, m):
    import torch
    import mdiscriminator_amd as md
    for n in (0, 1, 100, TILE, TILE + 1, 5 * TILE + 77, 1 << 20):
        x = W.uniform_u31(n, 4000 + 17 * m + n)
        ref, ref_cnt = oracle.seq_mdiscr(x, m)
        ref_off = _ref_offsets(ref_cnt, n)
        assert ref_off[-1] == n
        print(f"offsets m={m} n={n}", flush=True)
        d_in = to_dev(x) if n else torch.empty(0, dtype=torch.int32, device="cuda:0")
        d_out = torch.full_like(d_in, -1)
        d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
        d_off = torch.full((m + 1,), -1, dtype=torch.int64, device="cuda:0")
        ctx.partition_u32_ex(d_in, d_out, md.Mod(m), counts=d_cnt, offsets=d_off)
        torch.cuda.synchronize()
        assert first_diff(from_dev(d_out), ref) < 0
        assert np.array_equal(from_dev(d_cnt), ref_cnt)
        assert np.array_equal(d_off.cpu().numpy().astype(np.uint64), ref_off)
        d_out.fill_(-1)
        hc, ho = ctx.partition_u32_ex(d_in, d_out, md.Mod(m), host=True)
        assert first_diff(from_dev(d_out), ref) < 0
        assert hc == ref_cnt.tolist()
        assert ho == ref_off.tolist()


def test_offsets_unaligned_input(ctx):
    """An input that is not 16-byte aligned takes the 4-byte-load kernel, which
    writes the offsets too."""
    import torch
    import mdiscriminator_amd as md
    n = 7 * TILE + 3
    x = W.uniform_u31(n + 1, 12)
    d_all = to_dev(x)
    d_in = d_all[1:]
    d_out = torch.full_like(d_in, -1)
    hc, ho = ctx.partition_u32_ex(d_in, d_out, md.Mask(8), host=True)
    ref, ref_cnt = oracle.seq_mdiscr(x[1:], 8, oracle.MASK)
    assert first_diff(from_dev(d_out), ref) < 0
    assert hc == ref_cnt.tolist()
    assert ho == _ref_offsets(ref_cnt, n).tolist()


def test_ex_rejects_unknown_flags(ctx):
    import torch
    import mdiscriminator_amd as md
    d = torch.zeros(16, dtype=torch.int32, device="cuda:0")
    st = ctx.lib.mdiscr_partition_u32_ex(ctx.h, d.data_ptr(), d.data_ptr() + 4096, 16,
                                         md.Mod4.c(), None, None, 2, None)
    assert st == 1


@pytest.mark.parametrize("m", [2, 3, 8, 13, 16])
def test_every_alignment(ctx, m):
    """Input and output views at every 4-byte offset mod 16: the 16-byte-load
    kernel (aligned input) stages class runs for the output's own alignment."""
    import torch
    import mdiscriminator_amd as md
    n = 9 * TILE + 1234
    x = W.uniform_u31(n + 4, 500 + m)
    d_all = to_dev(x)
    for off_in in range(4):
        d_in = d_all[off_in:off_in + n]
        ref, ref_cnt = oracle.seq_mdiscr(x[off_in:off_in + n], m)
        for off_out in range(4):
            print(f"alignment m={m} in+{off_in} out+{off_out}", flush=True)
            d_out_all = torch.full((n + 8,), -1, dtype=torch.int32, device="cuda:0")
            d_out = d_out_all[off_out:off_out + n]
            d_cnt = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
            ctx.partition_u32(d_in, d_out, md.Mod(m), d_counts=d_cnt)
            torch.cuda.synchronize()
            i = first_diff(from_dev(d_out), ref)
            assert i < 0, f"in+{off_in} out+{off_out}: first mismatch at {i}"
            assert np.array_equal(from_dev(d_cnt), ref_cnt)
            assert np.all(from_dev(d_out_all[:off_out]) == 0xFFFFFFFF)
            assert np.all(from_dev(d_out_all[off_out + n:]) == 0xFFFFFFFF)


def test_server_cancel_queued_request():
    """With a budget that holds one output, later requests queue for budget; a
    queued one can be withdrawn (its waiter returns EINVAL), a launched or
    finished one cannot, and the rest complete equal to the oracle."""
    import torch
    import mdiscriminator_amd as md
    n = 1 << 14
    xs = [W.uniform_u31(n, 900 + i) for i in range(4)]
    d_in = [to_dev(x) for x in xs]
    srv = md.Server(budget_bytes=n * 4, nstreams=1, max_batch_requests=4)
    try:
        t = [srv.submit(d, md.Mod(5)) for d in d_in]
        waited = {}

        def waiter():
            try:
                srv.wait(t[2])
                waited["status"] = 0
            except md.MdiscrError as e:
                waited["status"] = e.status

        th = threading.Thread(target=waiter)
        th.start()
        assert srv.cancel(t[2]) == 0           # queued behind t[0]'s output
        th.join(timeout=30)
        assert not th.is_alive() and waited["status"] == 1
        ptr, cnt, _ = srv.wait(t[0])
        assert srv.cancel(t[0]) == 4           # finished: EBUSY, release instead
        got = srv.output_tensor(t[0], ptr)
        srv.release(t[0])
        ref, ref_cnt = oracle.seq_mdiscr(xs[0], 5)
        assert first_diff(from_dev(got), ref) < 0 and cnt == ref_cnt.tolist()
        for i in (1, 3):
            ptr, cnt, _ = srv.wait(t[i])
            got = srv.output_tensor(t[i], ptr)
            srv.release(t[i])
            ref, ref_cnt = oracle.seq_mdiscr(xs[i], 5)
            assert first_diff(from_dev(got), ref) < 0 and cnt == ref_cnt.tolist()
        with pytest.raises(md.MdiscrError) as e:
            srv.cancel(t[2])                   # no longer known
        assert e.value.status == 1
        assert srv.stats()["completed"] == 3
    finally:
        srv.close()


def test_server_mix_with_16m_request():
    """A mixed batch holding a 16M-element request (the largest serving size)
    beside small ones: each served result equals the oracle."""
    import torch
    import mdiscriminator_amd as md
    sizes = [1 << 24, 1000, 70000, 4096, 1, 300001]
    ms = [16, 2, 4, 8, 2, 16]
    xs = [W.uniform_u31(nn, 700 + i) for i, nn in enumerate(sizes)]
    d_in = [to_dev(x) for x in xs]
    srv = md.Server(budget_bytes=1 << 30, nstreams=2, max_batch_requests=16)
    try:
        t = [srv.submit(d_in[i], md.Mask(ms[i])) for i in range(len(xs))]
        for i, tk in enumerate(t):
            print(f"served request {i}: n={sizes[i]} m={ms[i]}", flush=True)
            ptr, cnt, _ = srv.wait(tk)
            got = srv.output_tensor(tk, ptr)
            srv.release(tk)
            ref, ref_cnt = oracle.seq_mdiscr(xs[i], ms[i], oracle.MASK)
            assert first_diff(from_dev(got), ref) < 0, f"request {i}"
            assert cnt == ref_cnt.tolist()
    finally:
        srv.close()
