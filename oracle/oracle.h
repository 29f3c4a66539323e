/*
 * oracle.h -- CPU restatement of the reference m-discriminator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this code.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and only as the checker (or as the timed CPU baseline), never as the
 * thing measured on the GPU.
 *
 * Semantics (SeqMdiscr.h:29-49 of asbjornvg/mdiscriminator, cuda_code/):
 *   class c(i) = D(x[i]) in [0,m), count[k] = |{i : c(i)=k}|,
 *   off[k] = sum_{j<k} count[j],
 *   out[off[c(i)] + |{j<i : c(j)=c(i)}|] = x[i]          (stable m-way partition)
 *   sizes[i] = 0 except sizes[off[k]] = count[k] for every k with count[k] > 0.
 *
 * Discriminators (HelpersCommon.h:587-596, 840-849, 1200-1210):
 *   ORACLE_DISCR_MOD  : c = x % m   (Mod<N>)
 *   ORACLE_DISCR_MASK : c = x & (m-1), m a power of two (Mod4, Mod8)
 */
#ifndef MDISCR_ORACLE_H
#define MDISCR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_DISCR_MOD = 0, ORACLE_DISCR_MASK = 1 };

/* Plain stable partition.  counts: m entries (may be NULL); sizes: n entries
 * (may be NULL).  Returns 0 on success, -1 on bad arguments. */
int oracle_seq_mdiscr_u32(uint64_t n, const uint32_t* in, uint32_t* out,
                          uint32_t* sizes, uint32_t kind, uint32_t m,
                          uint32_t* counts);

/* Struct-of-arrays partition: the class comes from the key column
 * cols[key_col] (32-bit keys, or 1-byte values); every column is moved to the
 * same destination index as raw bytes (col_bytes[j] in {1,2,4,8}). */
int oracle_seq_mdiscr_soa(uint64_t n, uint32_t ncols, const void* const* cols,
                          const uint32_t* col_bytes, uint32_t key_col,
                          void* const* out_cols, uint32_t kind, uint32_t m,
                          uint32_t* counts);

/* Array-valued elements: element i is in[i*len .. i*len+len), classified by
 * word 0 and moved whole. */
int oracle_seq_mdiscr_arrays(uint64_t n, uint32_t len, const uint32_t* in,
                             uint32_t* out, uint32_t kind, uint32_t m,
                             uint32_t* counts);

/* Two-way filter predicates (unsigned 32-bit x). */
enum {
    ORACLE_PRED_LT = 0,    /* x <  a */
    ORACLE_PRED_LE = 1,    /* x <= a */
    ORACLE_PRED_GT = 2,    /* x >  a */
    ORACLE_PRED_GE = 3,    /* x >= a */
    ORACLE_PRED_EQ = 4,    /* x == a */
    ORACLE_PRED_NE = 5,    /* x != a */
    ORACLE_PRED_BITS = 6,  /* (x & a) == b */
    ORACLE_PRED_RANGE = 7  /* a <= x < b */
};

/* parFilter (ParFilter.hs:8-21): the elements where the predicate holds, in
 * order, then the rest, in order; flags is the reference's n-length flag
 * array (write [0,i] [i,n-i]); counts = [i, n-i].  flags/counts may be NULL. */
int oracle_par_filter_u32(uint64_t n, const uint32_t* in, uint32_t* out,
                          uint32_t* flags, uint32_t op, uint32_t a, uint32_t b,
                          uint32_t* counts);

/* The filter over records: predicate on the 32-bit key column, every column
 * moved; counts = [kept, rest]. */
int oracle_filter_soa(uint64_t n, uint32_t ncols, const void* const* cols,
                      const uint32_t* col_bytes, uint32_t key_col,
                      void* const* out_cols, uint32_t op, uint32_t a, uint32_t b,
                      uint32_t* counts);

/* Multi-column discriminator z ? D(x) : D(y) (MdiscrHost_tuples.cu.h:27,46):
 * class = cols[sel_col] != 0 ? D(cols[a_col]) : D(cols[b_col]), columns of
 * 1- or 4-byte values. */
int oracle_seq_mdiscr_soa_select(uint64_t n, uint32_t ncols, const void* const* cols,
                                 const uint32_t* col_bytes, uint32_t sel_col,
                                 uint32_t a_col, uint32_t b_col, void* const* out_cols,
                                 uint32_t kind, uint32_t m, uint32_t* counts);

/* Segmented partition: segment s is [seg_off[s], seg_off[s+1]), each
 * partitioned independently in place of its input range.  seg_counts is
 * nseg*m (may be NULL). */
int oracle_seq_mdiscr_segm(uint64_t n, const uint32_t* in,
                           const uint64_t* seg_off, uint64_t nseg,
                           uint32_t* out, uint32_t kind, uint32_t m,
                           uint32_t* seg_counts);

/* Reference flag-array form of the segmented partition
 * (MdiscrSegmHost.cu.h:36-42, validated per segment as in
 * MdiscrSegmTest.cu:21-38): segment_sizes[i] holds the segment length at
 * each segment start and 0 elsewhere; new_sizes is the per-segment sizes
 * array.  Returns -1 if the flags do not tile [0,n). */
int oracle_segm_flags(uint64_t n, const uint32_t* in,
                      const int32_t* segment_sizes, uint32_t* out,
                      int32_t* new_sizes, uint32_t kind, uint32_t m);

/* counts -> N-length sizes array (HelpersCommon.h:295-311). */
int oracle_sizes_from_counts(const uint32_t* counts, uint32_t m, uint64_t n,
                             uint32_t* sizes);

/* Restatement of OpenmpMdiscr.h:7-83: parallel map to one-hot m-tuples,
 * serial inclusive tuple scan, parallel scatter + flags.  nthreads <= 0 uses
 * the OpenMP default.  Needs n*(4 + 8m) bytes of scratch; returns -2 when
 * that allocation fails. */
int oracle_openmp_mdiscr_u32(uint64_t n, const uint32_t* in, uint32_t* out,
                             int32_t* sizes, uint32_t kind, uint32_t m,
                             int nthreads);

/* Wall-clock seconds (steady clock) for timing the CPU baseline. */
double oracle_now(void);

#ifdef __cplusplus
}
#endif

#endif /* MDISCR_ORACLE_H */
