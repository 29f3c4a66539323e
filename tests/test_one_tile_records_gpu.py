"""Inputs at one tile of the record and array paths (n = TILE-1, TILE,
TILE+1) on the GPU, against the oracle: at most one tile takes the single
launch that counts its own tile, one more takes the full count, scan and
partition.  Class counts start filled with -1, so a count the call did not
write shows up.  Tiles: SoA 4096 (m <= 32, 8-byte columns), 6144 (64
classes); records with a column narrower than 4 bytes, or a select source, up
to 16 classes 2048 (the quad-layout kernel); the select source above 16
classes 3072; array rows 2048 / 2560 / 1536 / 1024 elements for LEN
2 / 3 / 5 / 8."""
import numpy as np
import pytest

import oracle
from mdiscriminator_amd import _lib, workloads as W
from tests.helpers import first_diff

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    a = np.ascontiguousarray(a)
    view = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}.get(a.dtype)
    return torch.from_numpy(a.view(view) if view else a).to("cuda:0")


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def counts_tensor(m):
    import torch
    return torch.full((m,), -1, dtype=torch.int32, device="cuda:0")


def filled(c):
    """An output column like c, every byte 0xEE (a value the call must overwrite)."""
    import torch
    return torch.full(c.shape, 0xEE, dtype=torch.uint8, device=c.device).view(c.dtype) \
        if c.dtype == torch.uint8 else \
        torch.full((c.numel() * c.element_size(),), 0xEE, dtype=torch.uint8,
                   device=c.device).view(c.dtype)


def edges(tile):
    return (tile - 1, tile, tile + 1)


def check_cols(outs, ref_outs, cnt, ref_cnt, label):
    for j, (o, r) in enumerate(zip(outs, ref_outs)):
        i = first_diff(host(o, r.dtype), r)
        assert i < 0, f"{label}: column {j} differs first at {i}"
    assert np.array_equal(host(cnt, np.uint32), ref_cnt), f"{label}: counts"


@pytest.mark.parametrize("m,tile", [(4, 4096), (16, 4096), (64, 6144)])
def test_soa_records_one_tile(ctx, m, tile):
    import torch
    import mdiscriminator_amd as md
    for n in edges(tile):
        cols = [W.uniform_u31(n, 100 + n), W.normal_f32_bits(n, 200 + n),
                np.arange(n, dtype=np.uint32)]
        d_cols = [dev(c) for c in cols]
        outs = [torch.full_like(c, -1) for c in d_cols]
        cnt = counts_tensor(m)
        print(f"soa m={m} n={n}", flush=True)
        ctx.partition_soa(d_cols, outs, md.Mask(m), d_counts=cnt)
        torch.cuda.synchronize()
        ref_outs, ref_cnt = oracle.seq_mdiscr_soa(cols, m, 0, oracle.MASK)
        check_cols(outs, ref_outs, cnt, ref_cnt, f"soa m={m} n={n}")


def test_soa_wide_column_one_tile(ctx):
    """An 8-byte column moves the tile to 8-byte staging (4096 records)."""
    import torch
    import mdiscriminator_amd as md
    for n in edges(4096):
        cols = [W.uniform_u31(n, 300 + n), W.splitmix64(n, 400 + n)]
        d_cols = [dev(c) for c in cols]
        outs = [torch.full_like(c, -1) for c in d_cols]
        cnt = counts_tensor(7)
        ctx.partition_soa(d_cols, outs, md.Mod(7), d_counts=cnt)
        torch.cuda.synchronize()
        ref_outs, ref_cnt = oracle.seq_mdiscr_soa(cols, 7, 0, oracle.MOD)
        check_cols(outs, ref_outs, cnt, ref_cnt, f"wide n={n}")


@pytest.mark.parametrize("m,tile", [(2, 2048), (8, 2048), (16, 2048), (64, 3072)])
def test_select_one_tile(ctx, m, tile):
    """(uint x, uchar y, bool z) records under z ? x % m : y % m."""
    import torch
    import mdiscriminator_amd as md
    for n in edges(tile):
        cols = [W.uniform_u31(n, 500 + n),
                (W.uniform_u31(n, 600 + n) & 0xFF).astype(np.uint8),
                (W.uniform_u31(n, 700 + n) & 1).astype(np.uint8)]
        d_cols = [dev(c) for c in cols]
        outs = [torch.full_like(c, -1) if c.dtype != torch.uint8 else torch.full_like(c, 0xEE)
                for c in d_cols]
        cnt = counts_tensor(m)
        print(f"select m={m} n={n}", flush=True)
        ctx.partition_soa_select(d_cols, outs, md.Mask(m), 2, 0, 1, d_counts=cnt)
        torch.cuda.synchronize()
        ref_outs, ref_cnt = oracle.seq_mdiscr_soa_select(cols, m, 2, 0, 1, oracle.MASK)
        check_cols(outs, ref_outs, cnt, ref_cnt, f"select m={m} n={n}")


@pytest.mark.parametrize("shape", ["tuple_m1", "tuple_m4", "tuple_m16", "bytekey_mod7", "u16_m8",
                                   "filter_bytes"])
def test_quad_records_one_tile(ctx, shape):
    """Records with a column narrower than 4 bytes, up to 16 classes (the
    quad-layout kernel, 2048-record tiles)."""
    import torch
    import mdiscriminator_amd as md
    for n in edges(2048):
        x = W.uniform_u31(n, 1000 + n)
        b = (W.uniform_u31(n, 1100 + n) & 0xFF).astype(np.uint8)
        if shape.startswith("tuple"):
            m = int(shape[len("tuple_m"):])
            cols, key, d, kind = [x, b, (x & 1).astype(np.uint8)], 0, md.Mask(m), oracle.MASK
        elif shape == "bytekey_mod7":
            m = 7
            cols, key, d, kind = [(x % 7).astype(np.uint8), x], 0, md.Mod(7), oracle.MOD
        elif shape == "u16_m8":
            m = 8
            cols = [(x & 0xFFFF).astype(np.uint16), x, b]
            key, d, kind = 1, md.Mask(8), oracle.MASK
        else:
            m = 2
            cols, key = [x, b], 0
        d_cols = [dev(c) for c in cols]
        outs = [filled(c) for c in d_cols]
        cnt = counts_tensor(m)
        print(f"quad records {shape} n={n}", flush=True)
        if shape == "filter_bytes":
            ctx.filter_soa(d_cols, outs, md.Pred.lt(1 << 30), d_counts=cnt)
            torch.cuda.synchronize()
            ref_outs, ref_cnt = oracle.filter_soa(cols, _lib.PRED_LT, 1 << 30)
        else:
            ctx.partition_soa(d_cols, outs, d, key, d_counts=cnt)
            torch.cuda.synchronize()
            ref_outs, ref_cnt = oracle.seq_mdiscr_soa(cols, m, key, kind)
        check_cols(outs, ref_outs, cnt, ref_cnt, f"{shape} n={n}")


@pytest.mark.parametrize("offset,key_offset", [(1, 0), (2, 0), (3, 0), (0, 1), (0, 2)])
def test_quad_records_misaligned(ctx, offset, key_offset):
    """Columns that start off a 4-element boundary (a byte column `offset`
    bytes, the u32 key column `key_offset` words past one) take the
    ballot-ranked kernel; the result is the same."""
    import torch
    import mdiscriminator_amd as md
    n = 3 * 2048 + 5
    x = W.uniform_u31(n + key_offset, 1200 + offset + 7 * key_offset)
    b = (W.uniform_u31(n + offset, 1300 + offset) & 0xFF).astype(np.uint8)
    d_b = dev(b)[offset:]
    d_ob = torch.full((n + offset,), 0xEE, dtype=torch.uint8, device="cuda:0")[offset:]
    d_x = dev(x)[key_offset:]
    d_ox = torch.full((n + key_offset,), -1, dtype=torch.int32, device="cuda:0")[key_offset:]
    cnt = counts_tensor(16)
    ctx.partition_soa([d_x, d_b], [d_ox, d_ob], md.Mask(16), 0, d_counts=cnt)
    torch.cuda.synchronize()
    ref_outs, ref_cnt = oracle.seq_mdiscr_soa([x[key_offset:], b[offset:]], 16, 0, oracle.MASK)
    check_cols([d_ox, d_ob], ref_outs, cnt, ref_cnt, f"misaligned by {offset} / {key_offset}")


def test_filter_records_one_tile(ctx):
    import torch
    import mdiscriminator_amd as md
    for n in edges(4096):
        cols = [W.uniform_u31(n, 800 + n), np.arange(n, dtype=np.uint32)]
        d_cols = [dev(c) for c in cols]
        outs = [torch.full_like(c, -1) for c in d_cols]
        cnt = counts_tensor(2)
        ctx.filter_soa(d_cols, outs, md.Pred.lt(1 << 30), d_counts=cnt)
        torch.cuda.synchronize()
        ref_outs, ref_cnt = oracle.filter_soa(cols, _lib.PRED_LT, 1 << 30)
        check_cols(outs, ref_outs, cnt, ref_cnt, f"filter records n={n}")


@pytest.mark.parametrize("length,tile", [(2, 2048), (3, 2560), (5, 1536), (8, 1024)])
def test_arrays_one_tile(ctx, length, tile):
    import torch
    import mdiscriminator_amd as md
    for n in edges(tile):
        x = W.uniform_u31(n * length, 900 + n + length)
        d_in = dev(x)
        d_out = torch.full_like(d_in, -1)
        cnt = counts_tensor(5)
        print(f"arrays len={length} n={n}", flush=True)
        ctx.partition_arrays(d_in, d_out, length, md.Mod(5), d_counts=cnt)
        torch.cuda.synchronize()
        ref, ref_cnt = oracle.seq_mdiscr_arrays(x, length, 5, oracle.MOD)
        check_cols([d_out], [ref], cnt, ref_cnt, f"arrays len={length} n={n}")
