/*
 * mdiscr_segbench.c -- C host driver of the C ABI: the batched segmented
 * workload against one plain launch per segment.
 *
 *   mdiscr_segbench [nseg] [median] [sigma] [m] [reps]
 *
 * Segment lengths are round(exp(ln(median) + sigma*z)) clipped to [1, 4Mi]
 * with z from Box-Muller on a SplitMix64 stream; values are SplitMix64 >> 33
 * (uniform in [0, 2^31)); class = x & (m-1).  Prints one JSON line: the
 * batched call (mdiscr_partition_segmented on a stream, graph-replayed after
 * its first call) and the loop of mdiscr_partition_u32 calls, both timed with
 * HIP events, and whether their outputs and class counts agree bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "mdiscr.h"

#define CHECK_HIP(x)                                                        \
    do {                                                                    \
        hipError_t e_ = (x);                                                \
        if (e_ != hipSuccess) {                                             \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));         \
            return 1;                                                       \
        }                                                                   \
    } while (0)
#define CHECK_MD(ctx, x)                                                    \
    do {                                                                    \
        mdiscr_status s_ = (x);                                             \
        if (s_ != MDISCR_OK) {                                              \
            fprintf(stderr, "%s: %s (%s)\n", #x, mdiscr_status_str(s_),     \
                    mdiscr_last_error(ctx));                                \
            return 1;                                                       \
        }                                                                   \
    } while (0)

static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t sm64(uint64_t seed, uint64_t i) { return mix(seed + (i + 1) * 0x9E3779B97F4A7C15ull); }

int main(int argc, char** argv) {
    const uint64_t nseg = argc > 1 ? strtoull(argv[1], NULL, 10) : 10000;
    const double median = argc > 2 ? atof(argv[2]) : 4096.0;
    const double sigma = argc > 3 ? atof(argv[3]) : 1.5;
    const uint32_t m = argc > 4 ? (uint32_t)atoi(argv[4]) : 4;
    const int reps = argc > 5 ? atoi(argv[5]) : 5;
    const uint64_t seed = 0x5EED0000ull + 4;

    uint64_t* off = (uint64_t*)malloc((nseg + 1) * sizeof(uint64_t));
    off[0] = 0;
    for (uint64_t s = 0; s < nseg; s++) {
        const double u1 = ((double)(sm64(seed, 2 * s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(sm64(seed, 2 * s + 1) >> 11) * (1.0 / 9007199254740992.0);
        const double z = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
        double L = rint(exp(log(median) + sigma * z));
        if (L < 1) L = 1;
        if (L > 4194304.0) L = 4194304.0;
        off[s + 1] = off[s] + (uint64_t)L;
    }
    const uint64_t n = off[nseg];
    uint32_t* h_in = (uint32_t*)malloc(n * 4);
    for (uint64_t i = 0; i < n; i++) h_in[i] = (uint32_t)(sm64(seed, i) >> 33);

    mdiscr_ctx* ctx = NULL;
    if (mdiscr_ctx_create(&ctx, 0) != MDISCR_OK) {
        fprintf(stderr, "no device\n");
        return 1;
    }
    hipStream_t st;
    CHECK_HIP(hipStreamCreate(&st));
    uint32_t *d_in, *d_out_b, *d_out_p, *d_cnt_b, *d_cnt_p;
    uint64_t* d_off;
    CHECK_HIP(hipMalloc((void**)&d_in, n * 4));
    CHECK_HIP(hipMalloc((void**)&d_out_b, n * 4));
    CHECK_HIP(hipMalloc((void**)&d_out_p, n * 4));
    CHECK_HIP(hipMalloc((void**)&d_cnt_b, nseg * m * 4));
    CHECK_HIP(hipMalloc((void**)&d_cnt_p, nseg * m * 4));
    CHECK_HIP(hipMalloc((void**)&d_off, (nseg + 1) * 8));
    CHECK_HIP(hipMemcpy(d_in, h_in, n * 4, hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_off, off, (nseg + 1) * 8, hipMemcpyHostToDevice));
    const mdiscr_discr d = {MDISCR_DISCR_MASK, m};

    hipEvent_t a, b;
    CHECK_HIP(hipEventCreate(&a));
    CHECK_HIP(hipEventCreate(&b));
    /* batched: first call tunes and captures, the second replays */
    for (int w = 0; w < 2; w++)
        CHECK_MD(ctx, mdiscr_partition_segmented(ctx, d_in, d_out_b, n, d_off, off, nseg, d,
                                                 d_cnt_b, st));
    CHECK_HIP(hipStreamSynchronize(st));
    CHECK_HIP(hipEventRecord(a, st));
    for (int r = 0; r < reps; r++)
        CHECK_MD(ctx, mdiscr_partition_segmented(ctx, d_in, d_out_b, n, d_off, off, nseg, d,
                                                 d_cnt_b, st));
    CHECK_HIP(hipEventRecord(b, st));
    CHECK_HIP(hipEventSynchronize(b));
    float ms_b = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms_b, a, b));
    ms_b /= reps;

    /* one plain launch per segment */
    const int preps = reps > 2 ? 2 : reps;
    for (uint64_t s = 0; s < nseg; s++)
        CHECK_MD(ctx, mdiscr_partition_u32(ctx, d_in + off[s], d_out_p + off[s], off[s + 1] - off[s],
                                           d, d_cnt_p + s * m, st));
    CHECK_HIP(hipStreamSynchronize(st));
    CHECK_HIP(hipEventRecord(a, st));
    for (int r = 0; r < preps; r++)
        for (uint64_t s = 0; s < nseg; s++)
            CHECK_MD(ctx, mdiscr_partition_u32(ctx, d_in + off[s], d_out_p + off[s],
                                               off[s + 1] - off[s], d, d_cnt_p + s * m, st));
    CHECK_HIP(hipEventRecord(b, st));
    CHECK_HIP(hipEventSynchronize(b));
    float ms_p = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms_p, a, b));
    ms_p /= preps;

    uint32_t* h_ob = (uint32_t*)malloc(n * 4);
    uint32_t* h_op = (uint32_t*)malloc(n * 4);
    uint32_t* h_cb = (uint32_t*)malloc(nseg * m * 4);
    uint32_t* h_cp = (uint32_t*)malloc(nseg * m * 4);
    CHECK_HIP(hipMemcpy(h_ob, d_out_b, n * 4, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(h_op, d_out_p, n * 4, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(h_cb, d_cnt_b, nseg * m * 4, hipMemcpyDeviceToHost));
    CHECK_HIP(hipMemcpy(h_cp, d_cnt_p, nseg * m * 4, hipMemcpyDeviceToHost));
    const int same = memcmp(h_ob, h_op, n * 4) == 0 && memcmp(h_cb, h_cp, nseg * m * 4) == 0;

    printf("{\"nseg\": %llu, \"n\": %llu, \"m\": %u, \"median\": %g, \"sigma\": %g, "
           "\"batched_ms\": %.4f, \"batched_gelem_s\": %.2f, \"per_segment_ms\": %.4f, "
           "\"per_segment_gelem_s\": %.2f, \"speedup\": %.2f, \"identical\": %s}\n",
           (unsigned long long)nseg, (unsigned long long)n, m, median, sigma, ms_b,
           n / (ms_b * 1e6), ms_p, n / (ms_p * 1e6), ms_p / ms_b, same ? "true" : "false");
    mdiscr_ctx_destroy(ctx);
    return same ? 0 : 3;
}
