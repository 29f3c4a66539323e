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
