"""Seeded synthetic inputs for the m-discriminator workloads.

Every generator draws from SplitMix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014): output i of the stream seeded
with s is mix(s + (i+1) * 0x9E3779B97F4A7C15).  Workloads are seeded
0x5EED0000 + workload id so every run, test and measurement sees the same data.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
SEED_BASE = 0x5EED0000
_CHUNK = 1 << 22

# Zipf(s=1.5) over 8 classes: p_k proportional to (k+1)^-1.5.
ZIPF8_P = np.array([(k + 1) ** -1.5 for k in range(8)])
ZIPF8_P = ZIPF8_P / ZIPF8_P.sum()


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(n, seed, start=0):
    """n consecutive 64-bit outputs of the SplitMix64 stream `seed`, from `start`."""
    out = np.empty(n, dtype=np.uint64)
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        for b in range(0, n, _CHUNK):
            e = min(n, b + _CHUNK)
            i = np.arange(start + b + 1, start + e + 1, dtype=np.uint64)
            out[b:e] = _mix(s + i * GOLDEN)
    return out


def uniform_u31(n, seed):
    """uint32 values uniform in [0, 2^31)."""
    out = np.empty(n, dtype=np.uint32)
    for b in range(0, n, _CHUNK):
        e = min(n, b + _CHUNK)
        out[b:e] = (splitmix64(e - b, seed, b) >> np.uint64(33)).astype(np.uint32)
    return out


def zipf8_values(n, seed):
    """Skewed m=8 workload: class k ~ Zipf(1.5) over 8 classes by inverse CDF,
    value = k + 8 * (uniform mod 2^28), so value & 7 == k and value < 2^31."""
    cdf = np.cumsum(ZIPF8_P)
    cdf[-1] = 1.0
    out = np.empty(n, dtype=np.uint32)
    for b in range(0, n, _CHUNK):
        e = min(n, b + _CHUNK)
        r = splitmix64(e - b, seed, b)
        u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
        k = np.searchsorted(cdf, u, side="right").astype(np.uint32)
        np.minimum(k, 7, out=k)
        hi = (r & np.uint64((1 << 28) - 1)).astype(np.uint32)
        out[b:e] = k + (hi << np.uint32(3))
    return out


def normal_f32_bits(n, seed):
    """float32 N(0,1) samples (Box-Muller) returned as their raw uint32 bits."""
    r = splitmix64(2 * n, seed)
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))
    u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return z.astype(np.float32).view(np.uint32)


def lognormal_lengths(nseg, seed, median=4096.0, sigma=1.5, lo=1, hi=4 << 20):
    """Segment lengths round(exp(ln(median) + sigma * z)) clipped to [lo, hi]."""
    r = splitmix64(2 * nseg, seed)
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / (1 << 53))
    u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    L = np.rint(np.exp(np.log(median) + sigma * z))
    return np.clip(L, lo, hi).astype(np.int64)


def offsets_from_lengths(lengths):
    off = np.zeros(len(lengths) + 1, dtype=np.uint64)
    np.cumsum(lengths, out=off[1:])
    return off


def flags_from_lengths(lengths, n):
    """Reference segment_sizes form: length at each segment start, 0 elsewhere."""
    f = np.zeros(n, dtype=np.int32)
    off = offsets_from_lengths(lengths)[:-1].astype(np.int64)
    f[off] = np.asarray(lengths, dtype=np.int32)
    return f


def _s64(v):
    """uint64 constant as the int64 with the same bits."""
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def _lsr(z, k):
    """Logical right shift of int64 tensor bits."""
    return (z >> k) & ((1 << (64 - k)) - 1)


def splitmix64_torch(n, seed, device, start=0):
    """Same stream as splitmix64(), computed on `device` (int64 bit patterns;
    int64 multiplication wraps like uint64)."""
    import torch
    i = torch.arange(start + 1, start + n + 1, dtype=torch.int64, device=device)
    z = i * _s64(0x9E3779B97F4A7C15) + _s64(seed)
    z = (z ^ _lsr(z, 30)) * _s64(0xBF58476D1CE4E5B9)
    z = (z ^ _lsr(z, 27)) * _s64(0x94D049BB133111EB)
    return z ^ _lsr(z, 31)


def uniform_u31_torch(n, seed, device, chunk=1 << 24):
    """uniform_u31() generated on the GPU, as an int32 tensor."""
    import torch
    out = torch.empty(n, dtype=torch.int32, device=device)
    for b in range(0, n, chunk):
        e = min(n, b + chunk)
        out[b:e] = _lsr(splitmix64_torch(e - b, seed, device, b), 33).to(torch.int32)
    return out


def zipf8_values_torch(n, seed, device, chunk=1 << 24):
    """zipf8_values() generated on the GPU, as an int32 tensor."""
    import torch
    cdf = torch.tensor(np.cumsum(ZIPF8_P), dtype=torch.float64, device=device)
    cdf[-1] = 1.0
    out = torch.empty(n, dtype=torch.int32, device=device)
    for b in range(0, n, chunk):
        e = min(n, b + chunk)
        r = splitmix64_torch(e - b, seed, device, b)
        u = _lsr(r, 11).to(torch.float64) * (1.0 / (1 << 53))
        k = torch.searchsorted(cdf, u, right=True).clamp_(max=7).to(torch.int64)
        hi = r & ((1 << 28) - 1)
        out[b:e] = (k + (hi << 3)).to(torch.int32)
    return out
