"""Shared test helpers (digests, device transfer)."""
import numpy as np


def fnv1a64(a):
    """FNV-1a 64 over the little-endian bytes of a uint32 array."""
    h = 0xCBF29CE484222325
    b = np.ascontiguousarray(a, dtype="<u4").tobytes()
    for byte in b:
        h ^= byte
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def fnv_fast(a):
    """64-bit digest of a uint32 array: FNV-1a over per-4096-word sums of the
    words times their index (fast enough for 2^20 words in numpy)."""
    a = np.asarray(a, dtype=np.uint64)
    idx = np.arange(1, len(a) + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        w = a * (idx * np.uint64(0x9E3779B1) + np.uint64(1))
        pad = (-len(w)) % 4096
        w = np.concatenate([w, np.zeros(pad, dtype=np.uint64)]).reshape(-1, 4096).sum(axis=1)
    return fnv1a64(np.concatenate([(w & np.uint64(0xFFFFFFFF)).astype(np.uint32),
                                   (w >> np.uint64(32)).astype(np.uint32)]))


def to_dev(a, device="cuda:0"):
    """numpy uint32/int32 array -> int32 torch tensor on the GPU."""
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to(device)


def from_dev(t):
    """int32 torch tensor -> uint32 numpy array."""
    return t.cpu().numpy().view(np.uint32)


def first_diff(a, b):
    d = np.flatnonzero(np.asarray(a) != np.asarray(b))
    return int(d[0]) if len(d) else -1
