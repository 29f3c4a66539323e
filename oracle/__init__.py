"""ctypes access to the oracle (TEST INFRASTRUCTURE ONLY).

The CPU restatement (liboracle.so, from seq_mdiscr.c) and, when built, the
reference's own SeqMdiscr (oracle/_ref/libseqref.so).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
always as the checker or the CPU baseline; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libseqref.so")

MOD, MASK = 0, 1
# two-way filter predicates (oracle.h ORACLE_PRED_*)
PRED_LT, PRED_LE, PRED_GT, PRED_GE, PRED_EQ, PRED_NE, PRED_BITS, PRED_RANGE = range(8)

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_lib = None
_ref = None


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} not built (make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.oracle_seq_mdiscr_u32.argtypes = [_u64, _vp, _vp, _vp, _u32, _u32, _vp]
        L.oracle_seq_mdiscr_soa.argtypes = [_u64, _u32, _vp, _vp, _u32, _vp, _u32, _u32, _vp]
        L.oracle_seq_mdiscr_arrays.argtypes = [_u64, _u32, _vp, _vp, _u32, _u32, _vp]
        L.oracle_par_filter_u32.argtypes = [_u64, _vp, _vp, _vp, _u32, _u32, _u32, _vp]
        L.oracle_filter_soa.argtypes = [_u64, _u32, _vp, _vp, _u32, _vp, _u32, _u32, _u32, _vp]
        L.oracle_seq_mdiscr_soa_select.argtypes = [_u64, _u32, _vp, _vp, _u32, _u32, _u32, _vp,
                                                   _u32, _u32, _vp]
        L.oracle_seq_mdiscr_segm.argtypes = [_u64, _vp, _vp, _u64, _vp, _u32, _u32, _vp]
        L.oracle_segm_flags.argtypes = [_u64, _vp, _vp, _vp, _vp, _u32, _u32]
        L.oracle_sizes_from_counts.argtypes = [_vp, _u32, _u64, _vp]
        L.oracle_openmp_mdiscr_u32.argtypes = [_u64, _vp, _vp, _vp, _u32, _u32, ctypes.c_int]
        L.oracle_now.restype = ctypes.c_double
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_SO)
        R.ref_seq_mdiscr_mod.argtypes = [_u32, _vp, _vp, _vp, _u32]
        R.ref_seq_mdiscr_mod4.argtypes = [_u32, _vp, _vp, _vp]
        R.ref_seq_mdiscr_mod8.argtypes = [_u32, _vp, _vp, _vp]
        R.ref_tuple_to_size_array.argtypes = [_vp, _u32, _u32, _vp]
        _ref = R
    return _ref


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"{what} returned {rc}")


def seq_mdiscr(x, m, kind=MOD, sizes=False):
    """Stable partition of uint32 array x -> (out, counts[, sizes])."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty_like(x)
    cnt = np.zeros(m, dtype=np.uint32)
    sz = np.empty(len(x), dtype=np.uint32) if sizes else None
    _check(lib().oracle_seq_mdiscr_u32(len(x), _p(x), _p(out), _p(sz), kind, m, _p(cnt)),
           "oracle_seq_mdiscr_u32")
    return (out, cnt, sz) if sizes else (out, cnt)


def seq_mdiscr_soa(cols, m, key_col=0, kind=MOD):
    cols = [np.ascontiguousarray(c) for c in cols]
    n = len(cols[key_col])
    outs = [np.empty_like(c) for c in cols]
    k = len(cols)
    pin = (_vp * k)(*[c.ctypes.data for c in cols])
    pout = (_vp * k)(*[o.ctypes.data for o in outs])
    nb = (_u32 * k)(*[c.dtype.itemsize for c in cols])
    cnt = np.zeros(m, dtype=np.uint32)
    _check(lib().oracle_seq_mdiscr_soa(n, k, pin, nb, key_col, pout, kind, m, _p(cnt)),
           "oracle_seq_mdiscr_soa")
    return outs, cnt


def _col_arrays(cols):
    cols = [np.ascontiguousarray(c) for c in cols]
    outs = [np.empty_like(c) for c in cols]
    k = len(cols)
    pin = (_vp * k)(*[c.ctypes.data for c in cols])
    pout = (_vp * k)(*[o.ctypes.data for o in outs])
    nb = (_u32 * k)(*[c.dtype.itemsize for c in cols])
    return cols, outs, pin, pout, nb


def par_filter(x, op, a=0, b=0):
    """parFilter (ParFilter.hs:8-21): (out, flags, [kept, rest])."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty_like(x)
    flags = np.empty(len(x), dtype=np.uint32)
    cnt = np.zeros(2, dtype=np.uint32)
    _check(lib().oracle_par_filter_u32(len(x), _p(x), _p(out), _p(flags), op, a, b, _p(cnt)),
           "oracle_par_filter_u32")
    return out, flags, cnt


def filter_soa(cols, op, a=0, b=0, key_col=0):
    cols, outs, pin, pout, nb = _col_arrays(cols)
    cnt = np.zeros(2, dtype=np.uint32)
    _check(lib().oracle_filter_soa(len(cols[key_col]), len(cols), pin, nb, key_col, pout, op, a,
                                   b, _p(cnt)), "oracle_filter_soa")
    return outs, cnt


def seq_mdiscr_soa_select(cols, m, sel_col, a_col, b_col, kind=MOD):
    """class = cols[sel] != 0 ? D(cols[a]) : D(cols[b]) (MdiscrHost_tuples.cu.h:46)."""
    cols, outs, pin, pout, nb = _col_arrays(cols)
    cnt = np.zeros(m, dtype=np.uint32)
    _check(lib().oracle_seq_mdiscr_soa_select(len(cols[0]), len(cols), pin, nb, sel_col, a_col,
                                              b_col, pout, kind, m, _p(cnt)),
           "oracle_seq_mdiscr_soa_select")
    return outs, cnt


def seq_mdiscr_arrays(x, length, m, kind=MOD):
    """x: uint32 array of n*length words -> (out, counts)."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    n = len(x) // length
    out = np.empty_like(x)
    cnt = np.zeros(m, dtype=np.uint32)
    _check(lib().oracle_seq_mdiscr_arrays(n, length, _p(x), _p(out), kind, m, _p(cnt)),
           "oracle_seq_mdiscr_arrays")
    return out, cnt


def seq_mdiscr_segm(x, seg_off, m, kind=MOD):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    seg_off = np.ascontiguousarray(seg_off, dtype=np.uint64)
    nseg = len(seg_off) - 1
    out = np.empty_like(x)
    cnt = np.zeros(nseg * m, dtype=np.uint32)
    _check(lib().oracle_seq_mdiscr_segm(len(x), _p(x), _p(seg_off), nseg, _p(out), kind, m,
                                        _p(cnt)), "oracle_seq_mdiscr_segm")
    return out, cnt.reshape(nseg, m)


def segm_flags(x, segment_sizes, m, kind=MOD):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    f = np.ascontiguousarray(segment_sizes, dtype=np.int32)
    out = np.empty_like(x)
    ns = np.empty(len(x), dtype=np.int32)
    _check(lib().oracle_segm_flags(len(x), _p(x), _p(f), _p(out), _p(ns), kind, m),
           "oracle_segm_flags")
    return out, ns


def sizes_from_counts(counts, n):
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    s = np.empty(n, dtype=np.uint32)
    _check(lib().oracle_sizes_from_counts(_p(counts), len(counts), n, _p(s)),
           "oracle_sizes_from_counts")
    return s


def openmp_mdiscr(x, m, kind=MOD, nthreads=0):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.empty_like(x)
    sz = np.empty(len(x), dtype=np.int32)
    _check(lib().oracle_openmp_mdiscr_u32(len(x), _p(x), _p(out), _p(sz), kind, m, nthreads),
           "oracle_openmp_mdiscr_u32")
    return out, sz


def ref_seq_mdiscr(x, m, specialized=False):
    """The reference's own seqMdiscr<Mod<m>> (or Mod4/Mod8): (out, sizes[:n])."""
    x = np.ascontiguousarray(x, dtype=np.uint32).copy()
    out = np.empty_like(x)
    sz = np.empty(len(x) + 1, dtype=np.uint32)  # the reference may write sizes[n]
    R = ref()
    if specialized and m == 4:
        rc = R.ref_seq_mdiscr_mod4(len(x), _p(x), _p(out), _p(sz))
    elif specialized and m == 8:
        rc = R.ref_seq_mdiscr_mod8(len(x), _p(x), _p(out), _p(sz))
    else:
        rc = R.ref_seq_mdiscr_mod(len(x), _p(x), _p(out), _p(sz), m)
    _check(rc, "ref_seq_mdiscr")
    return out, sz[:len(x)]


def ref_tuple_to_size_array(counts, n):
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    s = np.empty(n, dtype=np.uint32)
    _check(ref().ref_tuple_to_size_array(_p(counts), len(counts), n, _p(s)),
           "ref_tuple_to_size_array")
    return s


def now():
    return lib().oracle_now()
