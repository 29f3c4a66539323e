"""The no-allocation plain entry point (mdiscr_partition_u32_noalloc) through
the C ABI: two calls in a row, the second on a smaller input, each equal to
mdiscr_partition_u32 on the same input bit for bit (output, counts) and to
the oracle; offsets equal the counts' exclusive scan; inputs too large for the
fixed scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest

import oracle
from mdiscriminator_amd import _lib, workloads as W
from tests.helpers import first_diff, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _noalloc(ctx, d_in, d_out, n, discr, d_counts, d_offsets, stream):
    return ctx.lib.mdiscr_partition_u32_noalloc(
        ctx.h, d_in.data_ptr(), d_out.data_ptr(), n, discr.c(),
        d_counts.data_ptr() if d_counts is not None else None,
        d_offsets.data_ptr() if d_offsets is not None else None, stream)


@pytest.mark.parametrize("m,kind", [(8, "mask"), (7, "mod"), (64, "mask")])
def test_noalloc_matches_partition_u32(ctx, m, kind):
    import torch
    import mdiscriminator_amd as md
    discr = md.Mask(m) if kind == "mask" else md.Mod(m)
    stream = torch.cuda.current_stream().cuda_stream
    for call, n in enumerate((5 * (1 << 20) + 333, 70001)):
        print(f"noalloc m={m} {kind} call {call} n={n}", flush=True)
        x = W.uniform_u31(n, 6000 + 31 * m + call)
        d_in = to_dev(x)
        out_a = torch.full_like(d_in, -1)
        out_b = torch.full_like(d_in, -1)
        cnt_a = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
        cnt_b = torch.full((m,), -1, dtype=torch.int32, device="cuda:0")
        off_b = torch.full((m + 1,), -1, dtype=torch.int64, device="cuda:0")
        ctx.partition_u32(d_in, out_a, discr, d_counts=cnt_a)
        st = _noalloc(ctx, d_in, out_b, n, discr, cnt_b, off_b, stream)
        assert st == _lib.OK, ctx.lib.mdiscr_last_error(ctx.h)
        torch.cuda.synchronize()
        assert torch.equal(out_a, out_b)
        assert torch.equal(cnt_a, cnt_b)
        ref, ref_cnt = oracle.seq_mdiscr(x, m, oracle.MASK if kind == "mask" else oracle.MOD)
        assert first_diff(from_dev(out_b), ref) < 0
        assert np.array_equal(from_dev(cnt_b), ref_cnt)
        ref_off = np.concatenate([[0], np.cumsum(ref_cnt.astype(np.uint64))])
        assert np.array_equal(off_b.cpu().numpy().astype(np.uint64), ref_off)


def test_noalloc_refuses_beyond_scratch(ctx):
    """Temporaries beyond the fixed scratch: ENOMEM, nothing written (the
    check runs on the host before any launch)."""
    import torch
    import mdiscriminator_amd as md
    n_big = 1 << 27   # m = 64: 2 * 32768 * 64 * 4 B = 16 MiB of temporaries
    d_in = torch.zeros(n_big, dtype=torch.int32, device="cuda:0")
    d_out = torch.full((n_big,), -1, dtype=torch.int32, device="cuda:0")
    st = ctx.lib.mdiscr_partition_u32_noalloc(ctx.h, d_in.data_ptr(), d_out.data_ptr(),
                                              n_big, md.Mask(64).c(), None, None, None)
    assert st == _lib.ENOMEM
    assert b"no-allocation" in ctx.lib.mdiscr_last_error(ctx.h)
    torch.cuda.synchronize()
    assert torch.all(d_out == -1)
