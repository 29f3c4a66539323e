"""mdiscriminator_amd -- m-way stable partition (m-discriminator) on MI355X.

Python host layer over the C ABI of libmdiscr.so (include/mdiscr.h).  It
mirrors the reference's host interface (asbjornvg/mdiscriminator, cuda_code/):

    reference                                           here
    ------------------------------------------------    ---------------------------
    Mod4 / Mod8 / Mod<N>  (HelpersCommon.h:587-596,     Mod4, Mod8, Mod(N), Mask(N)
        840-849, 1200-1210)
    mdiscr<DISCR>(n, num_hwd_thds, d_in, d_out)         mdiscr(DISCR, n, num_hwd_thds,
        -> counts (MdiscrHost_optimized_clean.cu.h:9)       d_in, d_out) -> counts
    mdiscr<DISCR>(n, thds, in1..3, out1..3) -> counts   mdiscr_tuples(...)
        (MdiscrHost_tuples.cu.h:262-272)
    mdiscrSegm<DISCR>(n, in, segment_sizes, out,        mdiscr_segm(DISCR, n, in,
        new_sizes) (MdiscrSegmHost.cu.h:36-42)              segment_sizes, out, new_sizes)
    tupleToSizeArray(counts, sizes, n)                  sizes_from_counts(...)
        (HelpersCommon.h:295-311)
    parFilter cond arr -> (arr', flags)                 par_filter(Pred, d_in, d_out,
        (haskell_code/ParFilter.hs:8-21)                    d_flags) -> kept
    ModTuple<N>: z ? x % N : y % N                      Context.partition_soa_select
        (MdiscrHost_tuples.cu.h:27,46)

Device buffers are torch tensors on the GPU (plumbing only: all compute runs in
the HIP kernels of libmdiscr.so).  Errors raise MdiscrError with the library's
status code; nothing falls back to the CPU.
"""
import ctypes
from dataclasses import dataclass

from . import _lib
from ._lib import MdiscrError, DISCR_MOD, DISCR_MASK, MAX_CLASSES, MAX_ELEMS

__all__ = [
    "Discr", "Mod", "Mask", "Mod4", "Mod8", "Context", "Server", "MdiscrError", "default_context",
    "mdiscr", "mdiscr_tuples", "mdiscr_segm", "sizes_from_counts", "MAX_CLASSES", "Pred",
    "par_filter",
]


@dataclass(frozen=True)
class Discr:
    """A discriminator: class = x % m (MOD) or x & (m-1) (MASK)."""
    kind: int
    m: int

    def c(self):
        return _lib.MdiscrDiscr(self.kind, self.m)

    def apply(self, x):
        """Host-side class of x (numpy arrays or ints)."""
        return (x & (self.m - 1)) if self.kind == DISCR_MASK else (x % self.m)


def Mod(n):
    """Mod<N>: class = x % N (HelpersCommon.h:1200-1210)."""
    return Discr(DISCR_MOD, int(n))


def Mask(n):
    """class = x & (N-1), N a power of two."""
    return Discr(DISCR_MASK, int(n))


Mod4 = Mask(4)  # HelpersCommon.h:587-596
Mod8 = Mask(8)  # HelpersCommon.h:840-849


@dataclass(frozen=True)
class Pred:
    """A two-way filter predicate on unsigned 32-bit keys; elements where it
    holds come first (class 0), the rest after (class 1)."""
    op: int
    a: int = 0
    b: int = 0

    @staticmethod
    def lt(a):
        return Pred(_lib.PRED_LT, a)

    @staticmethod
    def le(a):
        return Pred(_lib.PRED_LE, a)

    @staticmethod
    def gt(a):
        return Pred(_lib.PRED_GT, a)

    @staticmethod
    def ge(a):
        return Pred(_lib.PRED_GE, a)

    @staticmethod
    def eq(a):
        return Pred(_lib.PRED_EQ, a)

    @staticmethod
    def ne(a):
        return Pred(_lib.PRED_NE, a)

    @staticmethod
    def bits(mask, value):
        """(x & mask) == value"""
        return Pred(_lib.PRED_BITS, mask, value)

    @staticmethod
    def range(lo, hi):
        """lo <= x < hi"""
        return Pred(_lib.PRED_RANGE, lo, hi)

    def c(self):
        return _lib.MdiscrPred(self.op, self.a & 0xFFFFFFFF, self.b & 0xFFFFFFFF)


def _ptr(t, nbytes=None):
    """Device address of a tensor argument; an int is taken as an address and
    passed through unchecked.  The kernels read and write dense arrays from the
    address given, so a tensor must be a contiguous GPU tensor of at least
    nbytes: a host tensor, a strided view or a short buffer raises here instead
    of faulting or corrupting memory on the device."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if not getattr(t, "is_cuda", False):
        raise MdiscrError(_lib.EINVAL, f"expected a GPU tensor, got one on {t.device}")
    if not t.is_contiguous():
        raise MdiscrError(_lib.EINVAL, "expected a contiguous tensor")
    have = t.numel() * t.element_size()
    if nbytes is not None and have < nbytes:
        raise MdiscrError(_lib.EINVAL, f"tensor of {have} bytes, the call needs {nbytes}")
    return t.data_ptr()


def _columns(cols, out_cols, n):
    """Pointer arrays and element sizes of SoA columns, each checked to hold
    n elements (an output column has its input's element size)."""
    k = len(cols)
    if len(out_cols) != k:
        raise MdiscrError(_lib.EINVAL, f"{k} input columns but {len(out_cols)} output columns")
    sizes = [c.element_size() for c in cols]
    pin = (ctypes.c_void_p * k)(*[_ptr(c, n * b) for c, b in zip(cols, sizes)])
    pout = (ctypes.c_void_p * k)(*[_ptr(c, n * b) for c, b in zip(out_cols, sizes)])
    return pin, pout, (ctypes.c_uint32 * k)(*sizes)


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Context:
    """Owns an mdiscr_ctx (device workspace, timing records)."""

    def __init__(self, device=-1):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        st = self.lib.mdiscr_ctx_create(ctypes.byref(h), int(device))
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_ctx_create: " +
                              self.lib.mdiscr_status_str(st).decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.mdiscr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != _lib.OK:
            msg = self.lib.mdiscr_last_error(self.h).decode(errors="replace")
            raise MdiscrError(st, f"{what}: {msg}")

    # -- plain ---------------------------------------------------------------
    def partition_u32(self, d_in, d_out, discr, d_counts=None, n=None, stream=None):
        """Stream-ordered stable partition of d_in into d_out (uint32/int32 tensors)."""
        n = d_in.numel() if n is None else n
        st = self.lib.mdiscr_partition_u32(self.h, _ptr(d_in, 4 * n), _ptr(d_out, 4 * n), n,
                                           discr.c(), _ptr(d_counts, 4 * discr.m),
                                           _stream(stream))
        self._check(st, "mdiscr_partition_u32")

    def partition_u32_ex(self, d_in, d_out, discr, counts=None, offsets=None, host=False,
                         n=None, stream=None):
        """The plain partition with counts (m) and offsets (m+1, uint64).
        host=False: counts/offsets are device tensors, written stream-ordered.
        host=True: they are returned as Python lists after the device work."""
        n = d_in.numel() if n is None else n
        if host:
            hc = (ctypes.c_uint32 * discr.m)()
            ho = (ctypes.c_uint64 * (discr.m + 1))()
            st = self.lib.mdiscr_partition_u32_ex(self.h, _ptr(d_in, 4 * n), _ptr(d_out, 4 * n),
                                                  n, discr.c(), hc, ho, _lib.OUT_HOST,
                                                  _stream(stream))
            self._check(st, "mdiscr_partition_u32_ex")
            return list(hc), list(ho)
        st = self.lib.mdiscr_partition_u32_ex(self.h, _ptr(d_in, 4 * n), _ptr(d_out, 4 * n), n,
                                              discr.c(), _ptr(counts, 4 * discr.m),
                                              _ptr(offsets, 8 * (discr.m + 1)), 0,
                                              _stream(stream))
        self._check(st, "mdiscr_partition_u32_ex")
        return None

    def mdiscr(self, discr, num_elems, num_hwd_thds, d_in, d_out):
        """Reference-shaped synchronous call; returns the m class counts."""
        counts = (ctypes.c_uint32 * discr.m)()
        st = self.lib.mdiscr_u32(self.h, num_elems, num_hwd_thds, _ptr(d_in, 4 * num_elems),
                                 _ptr(d_out, 4 * num_elems), discr.c(), counts)
        self._check(st, "mdiscr_u32")
        return list(counts)

    # -- struct of arrays ------------------------------------------------------
    def partition_soa(self, cols, out_cols, discr, key_col=0, d_counts=None, n=None,
                      stream=None):
        """Partitions parallel columns (tensors of 1/2/4/8-byte elements) by the
        class of cols[key_col] (a 32-bit column); payloads move as raw bits."""
        k = len(cols)
        n = cols[key_col].numel() if n is None else n
        pin, pout, nb = _columns(cols, out_cols, n)
        st = self.lib.mdiscr_partition_soa(self.h, k, pin, nb, key_col, pout, n, discr.c(),
                                           _ptr(d_counts, 4 * discr.m), _stream(stream))
        self._check(st, "mdiscr_partition_soa")

    def partition_soa_select(self, cols, out_cols, discr, sel_col, a_col, b_col, d_counts=None,
                             n=None, stream=None):
        """class = cols[sel] != 0 ? D(cols[a]) : D(cols[b]) (the alternative
        ModTuple body, MdiscrHost_tuples.cu.h:27,46); 1- or 4-byte columns."""
        k = len(cols)
        n = cols[0].numel() if n is None else n
        pin, pout, nb = _columns(cols, out_cols, n)
        st = self.lib.mdiscr_partition_soa_select(self.h, k, pin, nb, sel_col, a_col, b_col, pout,
                                                  n, discr.c(), _ptr(d_counts, 4 * discr.m),
                                                  _stream(stream))
        self._check(st, "mdiscr_partition_soa_select")

    # -- two-way filter ----------------------------------------------------------
    def filter_u32(self, d_in, d_out, pred, d_counts=None, n=None, stream=None):
        """parFilter (ParFilter.hs:8-21): kept elements, then the rest, both in
        input order; d_counts (2 entries) = [kept, rest]."""
        n = d_in.numel() if n is None else n
        st = self.lib.mdiscr_filter_u32(self.h, _ptr(d_in, 4 * n), _ptr(d_out, 4 * n), n,
                                        pred.c(), _ptr(d_counts, 8), _stream(stream))
        self._check(st, "mdiscr_filter_u32")

    def filter_soa(self, cols, out_cols, pred, key_col=0, d_counts=None, n=None, stream=None):
        """The filter over records: predicate on the 4-byte key column."""
        k = len(cols)
        n = cols[key_col].numel() if n is None else n
        pin, pout, nb = _columns(cols, out_cols, n)
        st = self.lib.mdiscr_filter_soa(self.h, k, pin, nb, key_col, pout, n, pred.c(),
                                        _ptr(d_counts, 8), _stream(stream))
        self._check(st, "mdiscr_filter_soa")

    # -- array-valued elements --------------------------------------------------
    def partition_arrays(self, d_in, d_out, length, discr, d_counts=None, n=None, stream=None):
        """Elements of `length` uint32 words stored flat; class from word 0
        (MdiscrHost_arrays.cu.h:14-96)."""
        n = d_in.numel() // length if n is None else n
        st = self.lib.mdiscr_partition_arrays(self.h, _ptr(d_in, 4 * n * length),
                                              _ptr(d_out, 4 * n * length), n, length, discr.c(),
                                              _ptr(d_counts, 4 * discr.m), _stream(stream))
        self._check(st, "mdiscr_partition_arrays")

    # -- segmented -------------------------------------------------------------
    def partition_segmented(self, d_in, d_out, d_seg_off, discr, h_seg_off=None,
                            d_seg_counts=None, stream=None):
        """Each segment [off[s], off[s+1]) partitioned independently."""
        nseg = d_seg_off.numel() - 1
        hp = None
        if h_seg_off is not None:
            import numpy as np
            h_seg_off = np.ascontiguousarray(h_seg_off, dtype=np.uint64)
            if h_seg_off.shape != (nseg + 1,):
                raise MdiscrError(_lib.EINVAL, f"h_seg_off has {h_seg_off.size} entries, "
                                               f"the call needs {nseg + 1}")
            hp = h_seg_off.ctypes.data_as(ctypes.c_void_p)
        n = d_in.numel()
        st = self.lib.mdiscr_partition_segmented(self.h, _ptr(d_in, 4 * n), _ptr(d_out, 4 * n), n,
                                                 _ptr(d_seg_off, 8 * (nseg + 1)), hp, nseg,
                                                 discr.c(), _ptr(d_seg_counts, 4 * nseg * discr.m),
                                                 _stream(stream))
        self._check(st, "mdiscr_partition_segmented")

    def segm_flags(self, discr, num_elems, d_in, d_segment_sizes, d_out, d_new_sizes,
                   stream=None):
        nb = 4 * num_elems
        st = self.lib.mdiscr_segm_flags(self.h, num_elems, _ptr(d_in, nb), _ptr(d_segment_sizes, nb),
                                        _ptr(d_out, nb), _ptr(d_new_sizes, nb), discr.c(),
                                        _stream(stream))
        self._check(st, "mdiscr_segm_flags")

    def partition_batch(self, requests, stream=None):
        """requests: list of (d_in, d_out, discr[, d_counts]) tensors; one launch."""
        arr = (_lib.MdiscrRequest * len(requests))()
        for i, r in enumerate(requests):
            d_in, d_out, discr = r[0], r[1], r[2]
            n = d_in.numel()
            arr[i] = _lib.MdiscrRequest(_ptr(d_in, 4 * n), _ptr(d_out, 4 * n), n, discr.c(),
                                        _ptr(r[3], 4 * discr.m) if len(r) > 3 else None)
        st = self.lib.mdiscr_partition_batch(self.h, arr, len(requests), _stream(stream))
        self._check(st, "mdiscr_partition_batch")

    # -- output format ---------------------------------------------------------
    def sizes_from_counts(self, d_counts, m, n, d_sizes, stream=None):
        st = self.lib.mdiscr_sizes_from_counts(self.h, _ptr(d_counts, 4 * m), m, n,
                                               _ptr(d_sizes, 4 * n), _stream(stream))
        self._check(st, "mdiscr_sizes_from_counts")

    # -- timing ----------------------------------------------------------------
    def timing(self, enable):
        self._check(self.lib.mdiscr_timing_enable(self.h, 1 if enable else 0), "timing_enable")

    def timing_read(self, name):
        ms = ctypes.c_double()
        cnt = ctypes.c_uint64()
        self._check(self.lib.mdiscr_timing_read(self.h, name.encode(), ctypes.byref(ms),
                                                ctypes.byref(cnt)), "timing_read")
        return ms.value, cnt.value

    def sync(self, stream=None):
        self._check(self.lib.mdiscr_sync(self.h, _stream(stream)), "mdiscr_sync")


class Server:
    """Continuous-batching server (mdiscr_server_*): submit device-resident
    inputs, wait for device outputs; outputs count against the budget until
    released."""

    def __init__(self, budget_bytes, nstreams=4, max_batch_requests=1024,
                 max_batch_elems=1 << 28, max_pending=0, device=0, reserve_bytes=0):
        self.lib = _lib.load()
        cfg = _lib.ServerConfig(device, budget_bytes, nstreams, max_batch_requests,
                                max_batch_elems, max_pending, reserve_bytes)
        h = ctypes.c_void_p()
        st = self.lib.mdiscr_server_create(ctypes.byref(h), ctypes.byref(cfg))
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_server_create")
        self.h = h
        self._m = {}

    def close(self):
        if self.h:
            self.lib.mdiscr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, d_in, discr, n=None):
        t = ctypes.c_uint64()
        n = d_in.numel() if n is None else n
        st = self.lib.mdiscr_server_submit(self.h, _ptr(d_in, 4 * n), n, discr.c(),
                                           ctypes.byref(t))
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_server_submit")
        self._m[t.value] = (discr.m, n)
        return t.value

    def poll(self, ticket):
        done = ctypes.c_int()
        st = self.lib.mdiscr_server_poll(self.h, ticket, ctypes.byref(done))
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_server_poll")
        return bool(done.value)

    def wait(self, ticket):
        """-> (device output pointer, class counts, latency seconds)."""
        m, n = self._m[ticket]
        p = ctypes.c_void_p()
        cnt = (ctypes.c_uint32 * m)()
        lat = ctypes.c_double()
        st = self.lib.mdiscr_server_wait(self.h, ticket, ctypes.byref(p), cnt, ctypes.byref(lat))
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_server_wait")
        return p.value, list(cnt), lat.value

    def output_tensor(self, ticket, ptr):
        """Copies the request's device output into a new int32 tensor."""
        import torch
        m, n = self._m[ticket]
        out = torch.empty(n, dtype=torch.int32, device="cuda:0")
        if n:
            from ctypes import CDLL
            hip = CDLL("libamdhip64.so")
            hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            rc = hip.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), n * 4, 3)
            if rc != 0:
                raise MdiscrError(_lib.EDEVICE, f"hipMemcpy {rc}")
        return out

    def release(self, ticket):
        st = self.lib.mdiscr_server_release(self.h, ticket)
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_server_release")
        self._m.pop(ticket, None)

    def cancel(self, ticket):
        """Withdraws a request not yet on the GPU; returns the status code
        (OK, or EBUSY once it has been launched)."""
        st = self.lib.mdiscr_server_cancel(self.h, ticket)
        if st == _lib.OK:
            self._m.pop(ticket, None)
        elif st != _lib.EBUSY:
            raise MdiscrError(st, "mdiscr_server_cancel")
        return st

    def stats(self):
        s = _lib.ServerStats()
        st = self.lib.mdiscr_server_stats_get(self.h, ctypes.byref(s))
        if st != _lib.OK:
            raise MdiscrError(st, "mdiscr_server_stats_get")
        return {f: getattr(s, f) for f, _ in _lib.ServerStats._fields_}


_default = {}


def default_context(device=None):
    import torch
    dev = torch.cuda.current_device() if device is None else int(device)
    if dev not in _default:
        _default[dev] = Context(dev)
    return _default[dev]


def mdiscr(discr, num_elems, num_hwd_thds, d_in, d_out):
    """mdiscr<DISCR>(n, num_hwd_thds, d_in, d_out) -> class counts
    (MdiscrHost_optimized_clean.cu.h:9-151)."""
    return default_context(d_in.device.index).mdiscr(discr, num_elems, num_hwd_thds, d_in,
                                                     d_out)


def mdiscr_tuples(discr, num_elems, num_hwd_thds, d_in1, d_in2, d_in3, d_out1, d_out2, d_out3):
    """Three-column form (MdiscrHost_tuples.cu.h:262-429): class from d_in1."""
    import torch
    ctx = default_context(d_in1.device.index)
    counts = torch.empty(discr.m, dtype=torch.int32, device=d_in1.device)
    ctx.partition_soa([d_in1, d_in2, d_in3], [d_out1, d_out2, d_out3], discr, key_col=0,
                      d_counts=counts, n=num_elems)
    return [int(x) for x in counts.cpu().tolist()]


def mdiscr_segm(discr, num_elems, in_array, segment_sizes, out_array, new_sizes):
    """mdiscrSegm<DISCR>(n, in, segment_sizes, out, new_sizes)
    (MdiscrSegmHost.cu.h:36-42)."""
    default_context(in_array.device.index).segm_flags(discr, num_elems, in_array, segment_sizes,
                                                      out_array, new_sizes)


def par_filter(pred, d_in, d_out, d_flags=None):
    """parFilter cond arr -> (arr', flags) (haskell_code/ParFilter.hs:8-21).
    Writes the filtered array to d_out and, when given, the reference's
    n-length flag array (write [0,i] [i,n-i]) to d_flags; returns i, the
    number of elements kept."""
    import torch
    ctx = default_context(d_in.device.index)
    n = d_in.numel()
    counts = torch.empty(2, dtype=torch.int32, device=d_in.device)
    ctx.filter_u32(d_in, d_out, pred, d_counts=counts)
    if d_flags is not None and n:
        ctx.sizes_from_counts(counts, 2, n, d_flags)
    return int(counts[0].item())


def sizes_from_counts(d_counts, m, n, d_sizes):
    """tupleToSizeArray (HelpersCommon.h:295-311) on device counts."""
    default_context(d_counts.device.index).sizes_from_counts(d_counts, m, n, d_sizes)
