/*
 * seq_mdiscr.c -- CPU restatement of the reference m-discriminator.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * path never calls, links or loads this file; it is built into
 * oracle/liboracle.so for tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.
 *
 * Each function cites the reference file:line it restates
 * (asbjornvg/mdiscriminator, paths relative to cuda_code/).
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Discriminator: Mod<N>::apply is x % N (HelpersCommon.h:1207-1209),
 * Mod4/Mod8 are x & 3 / x & 7 (HelpersCommon.h:593-595, 846-848). */
static inline uint32_t discr(uint32_t x, uint32_t kind, uint32_t m) {
    return kind == ORACLE_DISCR_MASK ? (x & (m - 1u)) : (x % m);
}

static int discr_ok(uint32_t kind, uint32_t m) {
    if (m == 0) return 0;
    if (kind == ORACLE_DISCR_MASK) return (m & (m - 1u)) == 0;
    return kind == ORACLE_DISCR_MOD;
}

double oracle_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* SeqMdiscr.h:29-49: histogram, exclusive scan over the m classes, stable
 * scatter with per-class cursors, then the sizes array.  The reference writes
 * sizes[offsets[k]] for every k, which indexes sizes[n] when trailing classes
 * are empty (an out-of-bounds write); here only non-empty classes write, which
 * gives the same in-bounds contents because empty classes precede the
 * non-empty class that shares their offset and are overwritten by it. */
int oracle_seq_mdiscr_u32(uint64_t n, const uint32_t* in, uint32_t* out,
                          uint32_t* sizes, uint32_t kind, uint32_t m,
                          uint32_t* counts) {
    if (!discr_ok(kind, m) || (n && (!in || !out))) return -1;
    uint64_t* red = (uint64_t*)calloc(m, sizeof(uint64_t));
    uint64_t* cur = (uint64_t*)calloc(m, sizeof(uint64_t));
    if (!red || !cur) { free(red); free(cur); return -2; }
    for (uint64_t i = 0; i < n; i++) red[discr(in[i], kind, m)]++;       /* :29-32 */
    uint64_t acc = 0;
    for (uint32_t k = 0; k < m; k++) { cur[k] = acc; acc += red[k]; }    /* :34-38 */
    for (uint64_t i = 0; i < n; i++) {                                   /* :40-45 */
        uint32_t x = in[i];
        out[cur[discr(x, kind, m)]++] = x;
    }
    if (sizes) {                                                         /* :44, :47-49 */
        memset(sizes, 0, n * sizeof(uint32_t));
        acc = 0;
        for (uint32_t k = 0; k < m; k++) {
            if (red[k]) sizes[acc] = (uint32_t)red[k];
            acc += red[k];
        }
    }
    if (counts)
        for (uint32_t k = 0; k < m; k++) counts[k] = (uint32_t)red[k];
    free(red); free(cur);
    return 0;
}

static inline void copy_elem(void* dst, uint64_t di, const void* src,
                             uint64_t si, uint32_t bytes) {
    memcpy((char*)dst + di * bytes, (const char*)src + si * bytes, bytes);
}

/* MdiscrHost_tuples.cu.h:81-112: the class comes from the first (key)
 * column only (ModTuple<N>::apply returns x % N, :44-45), and every column is
 * written to the same index. */
int oracle_seq_mdiscr_soa(uint64_t n, uint32_t ncols, const void* const* cols,
                          const uint32_t* col_bytes, uint32_t key_col,
                          void* const* out_cols, uint32_t kind, uint32_t m,
                          uint32_t* counts) {
    if (!discr_ok(kind, m) || ncols == 0 || key_col >= ncols || !cols ||
        !col_bytes || !out_cols || (col_bytes[key_col] != 4 && col_bytes[key_col] != 1))
        return -1;
    for (uint32_t j = 0; j < ncols; j++) {
        uint32_t b = col_bytes[j];
        if (b != 1 && b != 2 && b != 4 && b != 8) return -1;
        if (n && (!cols[j] || !out_cols[j])) return -1;
    }
    const void* key = cols[key_col];
    const int key8 = col_bytes[key_col] == 1;
#define KEY(i) (key8 ? (uint32_t)((const uint8_t*)key)[i] : ((const uint32_t*)key)[i])
    uint64_t* red = (uint64_t*)calloc(m, sizeof(uint64_t));
    uint64_t* cur = (uint64_t*)calloc(m, sizeof(uint64_t));
    if (!red || !cur) { free(red); free(cur); return -2; }
    for (uint64_t i = 0; i < n; i++) red[discr(KEY(i), kind, m)]++;
    uint64_t acc = 0;
    for (uint32_t k = 0; k < m; k++) { cur[k] = acc; acc += red[k]; }
    for (uint64_t i = 0; i < n; i++) {
        uint64_t d = cur[discr(KEY(i), kind, m)]++;
        for (uint32_t j = 0; j < ncols; j++)
            copy_elem(out_cols[j], d, cols[j], i, col_bytes[j]);
    }
    if (counts)
        for (uint32_t k = 0; k < m; k++) counts[k] = (uint32_t)red[k];
    free(red); free(cur);
#undef KEY
    return 0;
}

/* MdiscrHost_arrays.cu.h:62-86 (seqMdiscr<DISCR,LEN>): element i is the len
 * words in[i*len .. i*len+len), its class is D(word 0) (ModArray,
 * :27-36), and the whole element is copied to its stable position. */
int oracle_seq_mdiscr_arrays(uint64_t n, uint32_t len, const uint32_t* in,
                             uint32_t* out, uint32_t kind, uint32_t m,
                             uint32_t* counts) {
    if (!discr_ok(kind, m) || len == 0 || (n && (!in || !out))) return -1;
    uint64_t* red = (uint64_t*)calloc(m, sizeof(uint64_t));
    uint64_t* cur = (uint64_t*)calloc(m, sizeof(uint64_t));
    if (!red || !cur) { free(red); free(cur); return -2; }
    for (uint64_t i = 0; i < n; i++) red[discr(in[i * len], kind, m)]++;
    uint64_t acc = 0;
    for (uint32_t k = 0; k < m; k++) { cur[k] = acc; acc += red[k]; }
    for (uint64_t i = 0; i < n; i++) {
        uint64_t d = cur[discr(in[i * len], kind, m)]++;
        memcpy(out + d * len, in + i * len, len * sizeof(uint32_t));
    }
    if (counts)
        for (uint32_t k = 0; k < m; k++) counts[k] = (uint32_t)red[k];
    free(red); free(cur);
    return 0;
}

/* Predicates of the two-way filter, each evaluated from its own definition
 * (unsigned 32-bit comparisons).  -1 for an unknown op. */
static int pred_eval(uint32_t op, uint32_t a, uint32_t b, uint32_t x) {
    switch (op) {
        case ORACLE_PRED_LT: return x < a;
        case ORACLE_PRED_LE: return x <= a;
        case ORACLE_PRED_GT: return x > a;
        case ORACLE_PRED_GE: return x >= a;
        case ORACLE_PRED_EQ: return x == a;
        case ORACLE_PRED_NE: return x != a;
        case ORACLE_PRED_BITS: return (x & a) == b;
        case ORACLE_PRED_RANGE: return a <= x && x < b;
    }
    return -1;
}

/* parFilter cond arr (ParFilter.hs:8-21), step by step:
 *   cs  = map cond arr, tfs = 1 where cond holds, ffs = 1 where it does not;
 *   isT = scanInc (+) 0 tfs, i = last isT, isF = map (+ i) (scanInc (+) 0 ffs);
 *   inds = [c ? iT-1 : iF-1], result = permute inds arr;
 *   flags = write [0,i] [i,n-i] (replicate n 0): later writes win, and an
 *   index that matches no position (i == n) writes nothing (Helpers.hs write).
 * counts (2 entries, may be NULL) = [i, n-i]. */
int oracle_par_filter_u32(uint64_t n, const uint32_t* in, uint32_t* out,
                          uint32_t* flags, uint32_t op, uint32_t a, uint32_t b,
                          uint32_t* counts) {
    if (pred_eval(op, a, b, 0) < 0 || (n && (!in || !out))) return -1;
    uint64_t i_true = 0;                                     /* i = last isT */
    for (uint64_t k = 0; k < n; k++) i_true += (uint64_t)pred_eval(op, a, b, in[k]);
    uint64_t isT = 0, isF = 0;
    for (uint64_t k = 0; k < n; k++) {
        if (pred_eval(op, a, b, in[k])) out[++isT - 1] = in[k];
        else out[i_true + ++isF - 1] = in[k];
    }
    if (flags) {
        memset(flags, 0, n * sizeof(uint32_t));
        if (0 < n) flags[0] = (uint32_t)i_true;
        if (i_true < n) flags[i_true] = (uint32_t)(n - i_true);
    }
    if (counts) { counts[0] = (uint32_t)i_true; counts[1] = (uint32_t)(n - i_true); }
    return 0;
}

/* Moves every column of element i to the stable position of class cls[i]. */
static int soa_scatter(uint64_t n, uint32_t ncols, const void* const* cols,
                       const uint32_t* col_bytes, void* const* out_cols,
                       const uint8_t* cls, uint32_t m, uint32_t* counts) {
    uint64_t* red = (uint64_t*)calloc(m, sizeof(uint64_t));
    uint64_t* cur = (uint64_t*)calloc(m, sizeof(uint64_t));
    if (!red || !cur) { free(red); free(cur); return -2; }
    for (uint64_t i = 0; i < n; i++) red[cls[i]]++;
    uint64_t acc = 0;
    for (uint32_t k = 0; k < m; k++) { cur[k] = acc; acc += red[k]; }
    for (uint64_t i = 0; i < n; i++) {
        uint64_t d = cur[cls[i]]++;
        for (uint32_t j = 0; j < ncols; j++)
            copy_elem(out_cols[j], d, cols[j], i, col_bytes[j]);
    }
    if (counts)
        for (uint32_t k = 0; k < m; k++) counts[k] = (uint32_t)red[k];
    free(red); free(cur);
    return 0;
}

static int soa_cols_ok(uint64_t n, uint32_t ncols, const void* const* cols,
                       const uint32_t* col_bytes, void* const* out_cols) {
    if (ncols == 0 || !cols || !col_bytes || !out_cols) return 0;
    for (uint32_t j = 0; j < ncols; j++) {
        uint32_t b = col_bytes[j];
        if (b != 1 && b != 2 && b != 4 && b != 8) return 0;
        if (n && (!cols[j] || !out_cols[j])) return 0;
    }
    return 1;
}

static inline uint32_t col_word(const void* col, uint32_t bytes, uint64_t i) {
    return bytes == 1 ? (uint32_t)((const uint8_t*)col)[i] : ((const uint32_t*)col)[i];
}

/* The two-way filter over records: class 0 where the predicate holds on the
 * 32-bit key column, 1 elsewhere; every column moves. */
int oracle_filter_soa(uint64_t n, uint32_t ncols, const void* const* cols,
                      const uint32_t* col_bytes, uint32_t key_col,
                      void* const* out_cols, uint32_t op, uint32_t a, uint32_t b,
                      uint32_t* counts) {
    if (!soa_cols_ok(n, ncols, cols, col_bytes, out_cols) || key_col >= ncols ||
        col_bytes[key_col] != 4 || pred_eval(op, a, b, 0) < 0)
        return -1;
    uint8_t* cls = (uint8_t*)malloc(n ? n : 1);
    if (!cls) return -2;
    for (uint64_t i = 0; i < n; i++)
        cls[i] = pred_eval(op, a, b, ((const uint32_t*)cols[key_col])[i]) ? 0 : 1;
    int rc = soa_scatter(n, ncols, cols, col_bytes, out_cols, cls, 2, counts);
    free(cls);
    return rc;
}

/* The multi-column discriminator of MdiscrHost_tuples.cu.h:27,46
 * (z ? x % N : y % N): class = cols[sel] != 0 ? D(cols[a]) : D(cols[b]); the
 * three columns hold 1- or 4-byte values. */
int oracle_seq_mdiscr_soa_select(uint64_t n, uint32_t ncols, const void* const* cols,
                                 const uint32_t* col_bytes, uint32_t sel_col,
                                 uint32_t a_col, uint32_t b_col, void* const* out_cols,
                                 uint32_t kind, uint32_t m, uint32_t* counts) {
    if (!discr_ok(kind, m) || !soa_cols_ok(n, ncols, cols, col_bytes, out_cols) ||
        sel_col >= ncols || a_col >= ncols || b_col >= ncols)
        return -1;
    const uint32_t sb = col_bytes[sel_col], ab = col_bytes[a_col], bb = col_bytes[b_col];
    if ((sb != 1 && sb != 4) || (ab != 1 && ab != 4) || (bb != 1 && bb != 4)) return -1;
    uint8_t* cls = (uint8_t*)malloc(n ? n : 1);
    if (!cls) return -2;
    for (uint64_t i = 0; i < n; i++)
        cls[i] = (uint8_t)(col_word(cols[sel_col], sb, i)
                               ? discr(col_word(cols[a_col], ab, i), kind, m)
                               : discr(col_word(cols[b_col], bb, i), kind, m));
    int rc = soa_scatter(n, ncols, cols, col_bytes, out_cols, cls, m, counts);
    free(cls);
    return rc;
}

/* MdiscrSegmTest.cu:21-38: the segmented result is the plain result of each
 * segment placed at that segment's range. */
int oracle_seq_mdiscr_segm(uint64_t n, const uint32_t* in,
                           const uint64_t* seg_off, uint64_t nseg,
                           uint32_t* out, uint32_t kind, uint32_t m,
                           uint32_t* seg_counts) {
    if (!discr_ok(kind, m) || !seg_off) return -1;
    if (seg_off[0] != 0 || seg_off[nseg] != n) return -1;
    for (uint64_t s = 0; s < nseg; s++)
        if (seg_off[s + 1] < seg_off[s]) return -1;
    for (uint64_t s = 0; s < nseg; s++) {
        uint64_t b = seg_off[s], len = seg_off[s + 1] - b;
        int rc = oracle_seq_mdiscr_u32(len, in + b, out + b, NULL, kind, m,
                                       seg_counts ? seg_counts + s * m : NULL);
        if (rc) return rc;
    }
    return 0;
}

/* Flag-array interface of mdiscrSegm (MdiscrSegmHost.cu.h:36-42), checked per
 * segment with the plain oracle as MdiscrSegmTest.cu:21-38 does; new_sizes
 * follows sizesKernelSegm (MdiscrSegmKernels.cu.h:98-112): at segment-relative
 * index iot, the count of the class whose offset is iot, if non-empty. */
int oracle_segm_flags(uint64_t n, const uint32_t* in,
                      const int32_t* segment_sizes, uint32_t* out,
                      int32_t* new_sizes, uint32_t kind, uint32_t m) {
    if (!discr_ok(kind, m) || (n && (!in || !out || !segment_sizes)))
        return -1;
    uint64_t i = 0;
    while (i < n) {
        int32_t len = segment_sizes[i];
        if (len <= 0 || i + (uint64_t)len > n) return -1;
        for (uint64_t j = i + 1; j < i + (uint64_t)len; j++)
            if (segment_sizes[j] != 0) return -1;
        int rc = oracle_seq_mdiscr_u32((uint64_t)len, in + i, out + i,
                                       new_sizes ? (uint32_t*)(new_sizes + i) : NULL,
                                       kind, m, NULL);
        if (rc) return rc;
        i += (uint64_t)len;
    }
    return 0;
}

/* tupleToSizeArray (HelpersCommon.h:295-311): zero the array, then walk the
 * class sizes writing each at the running offset; an empty class writes 0 at
 * the offset that the next class then overwrites (last writer wins). */
int oracle_sizes_from_counts(const uint32_t* counts, uint32_t m, uint64_t n,
                             uint32_t* sizes) {
    if (!counts || m == 0 || (n && !sizes)) return -1;
    uint64_t tot = 0;
    for (uint32_t k = 0; k < m; k++) tot += counts[k];
    if (tot != n) return -1;
    memset(sizes, 0, n * sizeof(uint32_t));
    uint64_t i = 0;
    for (uint32_t j = 0; j < m && i < n; j++) {
        sizes[i] = counts[j];
        i += counts[j];
    }
    return 0;
}

/* OpenmpMdiscr.h:7-83.  The reference file does not compile against its own
 * tuple types (TupleType(0) at :32 and :38), so this restates the algorithm
 * with a flat n x m tuple array: parallel map to class + one-hot tuple
 * (:28-35), a serial inclusive tuple scan (:39-42), offsets (:46-51), parallel
 * scatter (index = scan[k] + offset[k] - 1, :53-61) and flags (:63-70). */
int oracle_openmp_mdiscr_u32(uint64_t n, const uint32_t* in, uint32_t* out,
                             int32_t* sizes, uint32_t kind, uint32_t m,
                             int nthreads) {
    if (!discr_ok(kind, m) || (n && (!in || !out || !sizes))) return -1;
    if (n == 0) return 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int32_t* classes = (int32_t*)malloc(n * sizeof(int32_t));
    int32_t* columns = (int32_t*)malloc(n * m * sizeof(int32_t));
    int32_t* scan = (int32_t*)malloc(n * m * sizeof(int32_t));
    int32_t* red = (int32_t*)calloc(m, sizeof(int32_t));
    int32_t* off = (int32_t*)calloc(m, sizeof(int32_t));
    if (!classes || !columns || !scan || !red || !off) {
        free(classes); free(columns); free(scan); free(red); free(off);
        return -2;
    }
    int64_t nn = (int64_t)n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; i++) {
        int32_t c = (int32_t)discr(in[i], kind, m);
        classes[i] = c;
        int32_t* t = columns + (uint64_t)i * m;
        for (uint32_t k = 0; k < m; k++) t[k] = 0;
        t[c] = 1;
    }
    for (uint32_t k = 0; k < m; k++) scan[k] = columns[k];
    for (uint64_t i = 1; i < n; i++) {
        const int32_t* prev = scan + (i - 1) * m;
        const int32_t* col = columns + i * m;
        int32_t* dst = scan + i * m;
        for (uint32_t k = 0; k < m; k++) dst[k] = prev[k] + col[k];
    }
    const int32_t* last = scan + (n - 1) * m;
    int32_t acc = 0;
    for (uint32_t k = 0; k < m; k++) { red[k] = last[k]; off[k] = acc; acc += last[k]; }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; i++) {
        int32_t k = classes[i];
        int32_t idx = scan[(uint64_t)i * m + k] + off[k] - 1;
        out[idx] = in[i];
        int32_t s = 0;
        for (uint32_t kk = 0; kk < m; kk++)
            if (i == off[kk] && red[kk] > 0) s = red[kk];
        sizes[i] = s;
    }
    free(classes); free(columns); free(scan); free(red); free(off);
    return 0;
}
