/*
 * mdiscr.h -- C ABI of the MI355X (gfx950) m-discriminator.
 *
 * An m-discriminator is a multi-way stable partition: a discriminator maps
 * each element to a class in [0,m); the output holds every class-0 element,
 * then every class-1 element and so on, each class in input order, and the m
 * class counts come back beside it.
 *
 * The entry points replace the reference's header-only C++ templates
 * (asbjornvg/mdiscriminator, cuda_code/):
 *   mdiscr<DISCR>(n, num_hwd_thds, d_in, d_out) -> counts
 *       MdiscrHost_optimized_clean.cu.h:9-15
 *   mdiscr<DISCR>(n, num_hwd_thds, d_in1..3, d_out1..3) -> counts (SoA)
 *       MdiscrHost_tuples.cu.h:262-272
 *   mdiscrSegm<DISCR>(n, in, segment_sizes, out, new_sizes)
 *       MdiscrSegmHost.cu.h:36-42
 *   tupleToSizeArray(counts, sizes, n)
 *       HelpersCommon.h:295-311
 * The compile-time discriminator policy (Mod4, Mod8, Mod<N>:
 * HelpersCommon.h:587-596, 840-849, 1200-1210) becomes the runtime
 * mdiscr_discr value.
 *
 * Conventions
 *   - Every pointer named d_* is device memory on the context's device.
 *   - Calls are stream-ordered on `stream` (a hipStream_t; NULL = the null
 *     stream) unless documented as synchronous; mdiscr_sync waits.
 *   - Inputs are never written.  Outputs must not overlap inputs.
 *   - Errors are returned as mdiscr_status, never by exiting the process.
 *   - A context is not thread-safe: use one per host thread (the serving
 *     loop owns its own).  The context owns all temporary device memory, so
 *     steady-state calls do not allocate.
 */
#ifndef MDISCR_H
#define MDISCR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MDISCR_OK = 0,
    MDISCR_EINVAL = 1,      /* bad argument (null pointer, m out of range, ...) */
    MDISCR_ENOMEM = 2,      /* device or host allocation failed / over budget */
    MDISCR_EDEVICE = 3,     /* a HIP runtime call failed */
    MDISCR_EBUSY = 4,       /* serving queue full */
    MDISCR_EUNSUPPORTED = 5 /* valid request outside what the library handles */
} mdiscr_status;

/* Discriminator kinds: the class of a 32-bit key x. */
typedef enum {
    MDISCR_DISCR_MOD = 0,  /* x % m           (Mod<N>, HelpersCommon.h:1207-1209) */
    MDISCR_DISCR_MASK = 1  /* x & (m-1), m a power of two (Mod4/Mod8) */
} mdiscr_discr_kind;

typedef struct {
    uint32_t kind; /* mdiscr_discr_kind */
    uint32_t m;    /* number of classes, 1 <= m <= MDISCR_MAX_CLASSES */
} mdiscr_discr;

#define MDISCR_MAX_CLASSES 64u
/* Largest n accepted: element indices and counts are 32-bit, as in the
 * reference (unsigned int throughout), with one tile of headroom. */
#define MDISCR_MAX_ELEMS 0xFFFF0000ull

typedef struct mdiscr_ctx mdiscr_ctx;

/* ---- context --------------------------------------------------------- */

/* Creates a context on HIP device `device` (the current device if < 0). */
mdiscr_status mdiscr_ctx_create(mdiscr_ctx** out, int device);
mdiscr_status mdiscr_ctx_destroy(mdiscr_ctx* ctx);
/* Waits for all work the context issued on `stream`. */
mdiscr_status mdiscr_sync(mdiscr_ctx* ctx, void* stream);
/* Human-readable text for a status code (static storage). */
const char* mdiscr_status_str(mdiscr_status s);
/* Text of the last failure on this context (static per context). */
const char* mdiscr_last_error(const mdiscr_ctx* ctx);

/* ---- plain partition (MdiscrHost_optimized_clean.cu.h:9-151) ---------- */

/* Stable partition of d_in[0,n) into d_out[0,n).  d_counts (device, m
 * entries, may be NULL) receives the class counts, stream-ordered. */
mdiscr_status mdiscr_partition_u32(mdiscr_ctx* ctx, const uint32_t* d_in,
                                   uint32_t* d_out, uint64_t n,
                                   mdiscr_discr d, uint32_t* d_counts,
                                   void* stream);

/* Where mdiscr_partition_u32_ex writes counts and offsets. */
#define MDISCR_OUT_HOST 1u /* host memory; the call returns once the device work is done */

/* The plain partition with its class offsets beside the counts.  counts (m
 * entries) and offsets (m+1 entries, offsets[k] = counts[0] + ... +
 * counts[k-1], offsets[m] = n: the exclusive scan of SeqMdiscr.h:34-38) may
 * each be NULL.  Without flags they are device memory, written
 * stream-ordered like d_counts above.  With MDISCR_OUT_HOST they are host
 * memory and the call waits for `stream`, as the reference returns its counts
 * by value after a device-to-host copy (MdiscrHost_optimized_clean.cu.h:87-92,
 * 149). */
mdiscr_status mdiscr_partition_u32_ex(mdiscr_ctx* ctx, const uint32_t* d_in,
                                      uint32_t* d_out, uint64_t n,
                                      mdiscr_discr d, uint32_t* counts,
                                      uint64_t* offsets, uint32_t flags,
                                      void* stream);

/* Bytes of the fixed device scratch of mdiscr_partition_u32_noalloc (one per
 * GPU, held until the process ends). */
#define MDISCR_NOALLOC_SCRATCH_BYTES (8u << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops and harnesses that forbid hipMalloc): the plain
 * partition of mdiscr_partition_u32_ex (device counts and offsets, either
 * may be NULL) with its temporaries in a fixed static device buffer of
 * MDISCR_NOALLOC_SCRATCH_BYTES per GPU instead of the context workspace.  The
 * call allocates and frees nothing (its kernels need (2*ld + 1) * m * 4 bytes,
 * ld = ceil(n / 4096) rounded up to a multiple of 4: n up to 2^29 at m = 8,
 * about 2^26 at m = 64); larger n and m are refused with MDISCR_ENOMEM.  Every
 * call on one GPU shares that GPU's buffer: calls that overlap on one GPU (on
 * different streams or host threads) race and must be serialised by the
 * caller.  Streams under graph capture are refused like everywhere else. */
mdiscr_status mdiscr_partition_u32_noalloc(mdiscr_ctx* ctx, const uint32_t* d_in,
                                           uint32_t* d_out, uint64_t n,
                                           mdiscr_discr d, uint32_t* d_counts,
                                           uint64_t* d_offsets, void* stream);

/* Reference-shaped synchronous call: like mdiscr<DISCR>(n, num_hwd_thds,
 * d_in, d_out), the class counts are returned in host memory h_counts (m
 * entries) after the device work has finished.  num_hwd_thds is accepted
 * for signature parity and ignored (the tile geometry is chosen here). */
mdiscr_status mdiscr_u32(mdiscr_ctx* ctx, uint64_t n, uint32_t num_hwd_thds,
                         const uint32_t* d_in, uint32_t* d_out,
                         mdiscr_discr d, uint32_t* h_counts);

/* ---- struct-of-arrays partition (MdiscrHost_tuples.cu.h:262-429) ------ */

/* Partitions ncols parallel columns by the class of the key column
 * d_cols[key_col]; every column moves to the same index as raw bits.
 * d_cols / d_out_cols / col_bytes are HOST arrays of ncols entries holding
 * device pointers / element sizes in {1,2,4,8}.  ncols <= 8.  The key column
 * holds 32-bit keys (class = D(key), like ModTuple<N>) or 1-byte values
 * (class = D(value): a caller-computed class array with m > value). */
mdiscr_status mdiscr_partition_soa(mdiscr_ctx* ctx, uint32_t ncols,
                                   const void* const* d_cols,
                                   const uint32_t* col_bytes, uint32_t key_col,
                                   void* const* d_out_cols, uint64_t n,
                                   mdiscr_discr d, uint32_t* d_counts,
                                   void* stream);

/* ---- array-valued elements (MdiscrHost_arrays.cu.h:250-470) ------------ */

/* Elements are len consecutive uint32 words (d_in[i*len .. i*len+len)); the
 * class is D(word 0) (ModArray<N,LEN>, MdiscrHost_arrays.cu.h:27-36) and each
 * whole element moves.  1 <= len <= 8; n counts elements. */
mdiscr_status mdiscr_partition_arrays(mdiscr_ctx* ctx, const uint32_t* d_in,
                                      uint32_t* d_out, uint64_t n, uint32_t len,
                                      mdiscr_discr d, uint32_t* d_counts,
                                      void* stream);

/* ---- general discriminators ------------------------------------------ */

/* Two-way filter predicates on unsigned 32-bit keys. */
typedef enum {
    MDISCR_PRED_LT = 0,     /* x <  a */
    MDISCR_PRED_LE = 1,     /* x <= a */
    MDISCR_PRED_GT = 2,     /* x >  a */
    MDISCR_PRED_GE = 3,     /* x >= a */
    MDISCR_PRED_EQ = 4,     /* x == a */
    MDISCR_PRED_NE = 5,     /* x != a */
    MDISCR_PRED_BITS = 6,   /* (x & a) == b */
    MDISCR_PRED_RANGE = 7   /* a <= x < b */
} mdiscr_pred_op;

typedef struct {
    uint32_t op;  /* mdiscr_pred_op */
    uint32_t a;
    uint32_t b;   /* BITS and RANGE only */
} mdiscr_pred;

/* parFilter cond arr (ParFilter.hs:8-21; no CUDA counterpart): d_out holds
 * the elements of d_in where the predicate holds, in input order, then the
 * rest, in input order.  d_counts (device, 2 entries, may be NULL) receives
 * [kept, rest]; the reference's flag array (write [0,i] [i,n-i]) is
 * mdiscr_sizes_from_counts(ctx, d_counts, 2, n, ...). */
mdiscr_status mdiscr_filter_u32(mdiscr_ctx* ctx, const uint32_t* d_in,
                                uint32_t* d_out, uint64_t n, mdiscr_pred pred,
                                uint32_t* d_counts, void* stream);

/* The filter over struct-of-arrays records: the predicate reads the 4-byte
 * key column and every column moves, as in mdiscr_partition_soa. */
mdiscr_status mdiscr_filter_soa(mdiscr_ctx* ctx, uint32_t ncols,
                                const void* const* d_cols,
                                const uint32_t* col_bytes, uint32_t key_col,
                                void* const* d_out_cols, uint64_t n,
                                mdiscr_pred pred, uint32_t* d_counts,
                                void* stream);

/* Multi-column discriminator: the alternative body of ModTuple4/ModTuple<N>
 * ::apply (MdiscrHost_tuples.cu.h:27,46, z ? x % N : y % N):
 * class(i) = cols[sel_col][i] != 0 ? D(cols[a_col][i]) : D(cols[b_col][i]).
 * The three columns hold 1- or 4-byte values and may coincide; every column
 * moves, as in mdiscr_partition_soa. */
mdiscr_status mdiscr_partition_soa_select(mdiscr_ctx* ctx, uint32_t ncols,
                                          const void* const* d_cols,
                                          const uint32_t* col_bytes,
                                          uint32_t sel_col, uint32_t a_col,
                                          uint32_t b_col,
                                          void* const* d_out_cols, uint64_t n,
                                          mdiscr_discr d, uint32_t* d_counts,
                                          void* stream);

/* ---- segmented partition (MdiscrSegmHost.cu.h:36-230) ---------------- */

/* Each segment s = [d_seg_off[s], d_seg_off[s+1]) of d_in is partitioned
 * independently into the same range of d_out.  d_seg_off is device memory
 * with nseg+1 entries, d_seg_off[0] = 0, d_seg_off[nseg] = n, non-decreasing.
 * h_seg_off (optional, host copy of the same offsets) lets the call plan the
 * work without reading them back.  d_seg_counts (device, nseg*m, may be
 * NULL) receives the per-segment class counts. */
mdiscr_status mdiscr_partition_segmented(mdiscr_ctx* ctx, const uint32_t* d_in,
                                         uint32_t* d_out, uint64_t n,
                                         const uint64_t* d_seg_off,
                                         const uint64_t* h_seg_off,
                                         uint64_t nseg, mdiscr_discr d,
                                         uint32_t* d_seg_counts, void* stream);

/* Reference flag-array form, mdiscrSegm(n, in, segment_sizes, out,
 * new_sizes): d_segment_sizes[i] is the segment length at each segment start
 * and 0 elsewhere; d_new_sizes receives, per segment, the sizes array of
 * that segment (class count at each class start, 0 elsewhere).  Synchronous
 * on `stream`. */
mdiscr_status mdiscr_segm_flags(mdiscr_ctx* ctx, uint64_t n,
                                const uint32_t* d_in,
                                const int32_t* d_segment_sizes,
                                uint32_t* d_out, int32_t* d_new_sizes,
                                mdiscr_discr d, void* stream);

/* ---- batched independent requests ----------------------------------------- */

/* One partition request: its own buffers, length and discriminator. */
typedef struct {
    const uint32_t* d_in;
    uint32_t* d_out;
    uint64_t n;
    mdiscr_discr d;
    uint32_t* d_counts; /* m entries, may be NULL */
} mdiscr_request;

/* Runs nreq independent requests as ONE segmented launch (each request is a
 * segment with its own input, output and m).  No reference counterpart: it
 * is the batched form that mdiscrSegm (MdiscrSegmHost.cu.h:36-42) provides
 * for segments of one array, generalised to separate buffers.  reqs is a
 * HOST array; the call is stream-ordered. */
mdiscr_status mdiscr_partition_batch(mdiscr_ctx* ctx, const mdiscr_request* reqs,
                                     uint32_t nreq, void* stream);

/* ---- serving ----------------------------------------------------------------
 * No reference counterpart (the reference has no batching or serving).  A
 * server owns nstreams contexts and streams and one scheduler thread.
 * Requests arrive with their input already in device memory; each is
 * admitted once its output buffer fits the device-memory budget (FIFO, so a
 * large request is not overtaken), and every admitted request that is not yet
 * on the GPU joins the next batch: a single mdiscr_partition_batch launch on
 * a free stream, while earlier batches still run on the others.  Outputs stay
 * resident until released.  Thread-safe: any host thread may submit, wait and
 * release. */
typedef struct mdiscr_server mdiscr_server;

typedef struct {
    int device;
    uint64_t budget_bytes;       /* device memory for request outputs */
    uint32_t nstreams;           /* batches in flight at once (1..16) */
    uint32_t max_batch_requests; /* requests per launch */
    uint64_t max_batch_elems;    /* elements per launch */
    uint32_t max_pending;        /* submit returns MDISCR_EBUSY beyond this */
    uint64_t reserve_bytes;      /* output arena mapped at creation (<= budget) */
} mdiscr_server_config;

typedef struct {
    uint64_t submitted, completed, batches, elems;
    double p50_s, p99_s, mean_s, max_s; /* submit -> results ready */
    uint64_t peak_bytes;                /* most output memory held at once */
    double mean_batch_requests;
} mdiscr_server_stats;

mdiscr_status mdiscr_server_create(mdiscr_server** out, const mdiscr_server_config* cfg);
mdiscr_status mdiscr_server_destroy(mdiscr_server* srv);
/* Queues a request; *ticket identifies it.  d_in must stay valid until the
 * request completes. */
mdiscr_status mdiscr_server_submit(mdiscr_server* srv, const uint32_t* d_in, uint64_t n,
                                   mdiscr_discr d, uint64_t* ticket);
/* *done = 1 once the request's output and counts are ready. */
mdiscr_status mdiscr_server_poll(mdiscr_server* srv, uint64_t ticket, int* done);
/* Blocks until the request completes; returns its device output (valid until
 * release), its class counts (h_counts, m entries, may be NULL) and its
 * latency from submit to completion. */
mdiscr_status mdiscr_server_wait(mdiscr_server* srv, uint64_t ticket, uint32_t** d_out,
                                 uint32_t* h_counts, double* latency_s);
/* Returns the request's output buffer to the budget. */
mdiscr_status mdiscr_server_release(mdiscr_server* srv, uint64_t ticket);
/* Withdraws a request that is not yet on the GPU (queued for budget or
 * admitted and waiting for a slot): its output block returns to the budget,
 * the ticket becomes unknown, and a thread blocked in mdiscr_server_wait on it
 * returns MDISCR_EINVAL.  MDISCR_EBUSY once the request has been launched or
 * has finished (wait for it, then release it).  No reference counterpart: the
 * reference has no request queue. */
mdiscr_status mdiscr_server_cancel(mdiscr_server* srv, uint64_t ticket);
mdiscr_status mdiscr_server_stats_get(mdiscr_server* srv, mdiscr_server_stats* out);

/* ---- output format (HelpersCommon.h:295-311) --------------------------- */

/* d_sizes[0,n) = 0 except d_sizes[off[k]] = counts[k] for non-empty k, from
 * device counts d_counts[0,m). */
mdiscr_status mdiscr_sizes_from_counts(mdiscr_ctx* ctx,
                                       const uint32_t* d_counts, uint32_t m,
                                       uint64_t n, uint32_t* d_sizes,
                                       void* stream);

/* ---- kernel timing (measurement support) ------------------------------ */

/* When enabled, every kernel the context launches is bracketed by HIP
 * events on its stream and its duration accumulated per kernel name. */
mdiscr_status mdiscr_timing_enable(mdiscr_ctx* ctx, int enable);
/* Reads (synchronising the recorded events) and resets the accumulated
 * time of kernel `name`: total milliseconds and launch count. */
mdiscr_status mdiscr_timing_read(mdiscr_ctx* ctx, const char* name,
                                 double* total_ms, uint64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* MDISCR_H */
