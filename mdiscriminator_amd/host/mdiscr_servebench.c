/*
 * mdiscr_servebench.c -- C host driver: the serving workload.
 *
 *   mdiscr_servebench [nreq] [load] [budget_gib] [nstreams] [reserve_gib]
 *
 * nreq partition requests (default 50000) with sizes log-uniform over
 * [1Ki, 16Mi] uint32 values (uniform in [0, 2^31)), m uniform in {2,4,8,16},
 * class = x & (m-1).  Inputs are views into one resident 64Mi-value pool.
 * Three capacities are measured on all requests submitted at once: the plain
 * per-request path (one mdiscr_partition_u32 call + stream sync each), the
 * plain path on nstreams streams (one context each, requests dealt
 * round-robin, no host wait per request) and the continuous-batching server
 * (mdiscr_server_*).  Poisson arrival schedules at `load` (default 0.8)
 * times the plain and the server capacity are then served by all three, and
 * one JSON line reports p50/p99 latency (arrival -> result ready) and the
 * aggregate Gelem/s of each, plus spot checks that served outputs equal the
 * plain path's bit for bit.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime_api.h>

#include "mdiscr.h"

#define POOL (64ull << 20)

static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t sm64(uint64_t seed, uint64_t i) { return mix(seed + (i + 1) * 0x9E3779B97F4A7C15ull); }
static double unit(uint64_t seed, uint64_t i) { return (double)(sm64(seed, i) >> 11) * (1.0 / 9007199254740992.0); }
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
static void sleep_until(double t) {
    for (;;) {
        const double d = t - now_s();
        if (d <= 0) return;
        if (d > 2e-4) {
            struct timespec ts = {0, (long)((d - 1e-4) * 1e9)};
            nanosleep(&ts, NULL);
        }
    }
}
static int cmp_d(const void* a, const void* b) {
    const double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}
static double pct(double* v, int n, double p) {
    qsort(v, n, sizeof(double), cmp_d);
    int i = (int)(p * (n - 1) + 0.5);
    return v[i < n ? i : n - 1];
}

typedef struct {
    uint64_t n, off;
    uint32_t m;
    double t_arrive;
    volatile uint64_t ticket;
} Req;

typedef struct {
    mdiscr_server* srv;
    Req* r;
    int nreq;
    const uint32_t* pool;
    double t0;
} Client;

typedef struct {
    double p50_ms, p99_ms, gelem_s, mean_batch;
    uint64_t batches;
    double peak_gib;
    int identical;
} Result;

static void* submitter(void* arg) {
    Client* c = (Client*)arg;
    for (int i = 0; i < c->nreq; i++) {
        sleep_until(c->t0 + c->r[i].t_arrive);
        const mdiscr_discr d = {MDISCR_DISCR_MASK, c->r[i].m};
        uint64_t t = 0;
        mdiscr_status st;
        while ((st = mdiscr_server_submit(c->srv, c->pool + c->r[i].off, c->r[i].n, d, &t)) ==
               MDISCR_EBUSY)
            sleep_until(now_s() + 1e-4);
        if (st != MDISCR_OK) {
            fprintf(stderr, "submit %d: %s\n", i, mdiscr_status_str(st));
            exit(1);
        }
        __atomic_store_n(&c->r[i].ticket, t, __ATOMIC_RELEASE);
    }
    return NULL;
}

static uint64_t g_total;

static Result run_server(Req* r, int nreq, const uint32_t* pool, const mdiscr_server_config* cfg,
                         mdiscr_ctx* ctx, hipStream_t st, uint32_t* d_plain) {
    Result res = {0};
    mdiscr_server* srv = NULL;
    if (mdiscr_server_create(&srv, cfg) != MDISCR_OK) exit(1);
    for (int i = 0; i < nreq; i++) r[i].ticket = 0;
    Client c = {srv, r, nreq, pool, now_s() + 0.02};
    pthread_t th;
    pthread_create(&th, NULL, submitter, &c);
    double* lat = (double*)malloc(nreq * sizeof(double));
    uint32_t* h_a = (uint32_t*)malloc((16ull << 20) * 4);
    uint32_t* h_b = (uint32_t*)malloc((16ull << 20) * 4);
    res.identical = 1;
    double t_last = 0;
    const int every = nreq / 32 ? nreq / 32 : 1;
    // Spot-checked outputs stay held (at most 32 x 64 MB of the budget) and
    // are compared after the run, so that the check's host copies do not stall
    // the client's releases, and admission with them, while latency is measured.
    int nchk = 0;
    int* chk_i = (int*)malloc((nreq / every + 1) * sizeof(int));
    uint64_t* chk_t = (uint64_t*)malloc((nreq / every + 1) * sizeof(uint64_t));
    uint32_t** chk_out = (uint32_t**)malloc((nreq / every + 1) * sizeof(uint32_t*));
    for (int i = 0; i < nreq; i++) {
        uint64_t t;
        while ((t = __atomic_load_n(&r[i].ticket, __ATOMIC_ACQUIRE)) == 0) sleep_until(now_s() + 2e-5);
        uint32_t* d_out = NULL;
        double l = 0;
        if (mdiscr_server_wait(srv, t, &d_out, NULL, &l) != MDISCR_OK) exit(1);
        lat[i] = l;
        t_last = now_s();
        if (i % every == 0) {
            chk_i[nchk] = i;
            chk_t[nchk] = t;
            chk_out[nchk++] = d_out;
        } else {
            mdiscr_server_release(srv, t);
        }
    }
    pthread_join(th, NULL);
    for (int k = 0; k < nchk; k++) {
        const int i = chk_i[k];
        const mdiscr_discr d = {MDISCR_DISCR_MASK, r[i].m};
        mdiscr_partition_u32(ctx, pool + r[i].off, d_plain, r[i].n, d, NULL, st);
        hipStreamSynchronize(st);
        hipMemcpy(h_a, chk_out[k], r[i].n * 4, hipMemcpyDeviceToHost);
        hipMemcpy(h_b, d_plain, r[i].n * 4, hipMemcpyDeviceToHost);
        if (memcmp(h_a, h_b, r[i].n * 4)) res.identical = 0;
        mdiscr_server_release(srv, chk_t[k]);
    }
    free(chk_i);
    free(chk_t);
    free(chk_out);
    mdiscr_server_stats ss;
    mdiscr_server_stats_get(srv, &ss);
    mdiscr_server_destroy(srv);
    res.p50_ms = pct(lat, nreq, 0.5) * 1e3;
    res.p99_ms = pct(lat, nreq, 0.99) * 1e3;
    res.gelem_s = g_total / (t_last - c.t0) / 1e9;
    res.batches = ss.batches;
    res.mean_batch = ss.mean_batch_requests;
    res.peak_gib = ss.peak_bytes / (double)(1ull << 30);
    free(lat);
    free(h_a);
    free(h_b);
    return res;
}

static Result run_plain(Req* r, int nreq, const uint32_t* pool, mdiscr_ctx* ctx, hipStream_t st,
                        uint32_t* d_plain) {
    Result res = {0};
    const double t0 = now_s() + 0.02;
    double* lat = (double*)malloc(nreq * sizeof(double));
    double t_free = t0;
    for (int i = 0; i < nreq; i++) {
        sleep_until(t0 + r[i].t_arrive);
        const mdiscr_discr d = {MDISCR_DISCR_MASK, r[i].m};
        mdiscr_partition_u32(ctx, pool + r[i].off, d_plain, r[i].n, d, NULL, st);
        hipStreamSynchronize(st);
        t_free = now_s();
        lat[i] = t_free - (t0 + r[i].t_arrive);
    }
    res.p50_ms = pct(lat, nreq, 0.5) * 1e3;
    res.p99_ms = pct(lat, nreq, 0.99) * 1e3;
    res.gelem_s = g_total / (t_free - t0) / 1e9;
    res.identical = 1;
    free(lat);
    return res;
}

/* The plain path on nstreams streams: one context, stream and output buffer
 * per stream, requests dealt round-robin as they arrive, no host wait per
 * request.  A second thread records each request's completion by polling
 * the event recorded after it. */
typedef struct {
    hipEvent_t* ev;
    volatile int submitted;
    int nreq;
    double* done_t;
} Poller;

static void* poll_events(void* arg) {
    Poller* p = (Poller*)arg;
    int next = 0;
    while (next < p->nreq) {
        const int sub = __atomic_load_n(&p->submitted, __ATOMIC_ACQUIRE);
        for (int j = next; j < sub; j++) {
            if (p->done_t[j] > 0) continue;
            if (hipEventQuery(p->ev[j]) == hipSuccess) p->done_t[j] = now_s();
        }
        while (next < sub && p->done_t[next] > 0) next++;
        if (next < p->nreq) sleep_until(now_s() + 5e-6);
    }
    return NULL;
}

static Result run_plain_streams(Req* r, int nreq, const uint32_t* pool, int k, mdiscr_ctx** ctxs,
                                hipStream_t* sts, uint32_t** outs, hipEvent_t* ev) {
    Result res = {0};
    double* done_t = (double*)calloc(nreq, sizeof(double));
    Poller pl = {ev, 0, nreq, done_t};
    pthread_t th;
    pthread_create(&th, NULL, poll_events, &pl);
    const double t0 = now_s() + 0.02;
    for (int i = 0; i < nreq; i++) {
        sleep_until(t0 + r[i].t_arrive);
        const int s = i % k;
        const mdiscr_discr d = {MDISCR_DISCR_MASK, r[i].m};
        if (mdiscr_partition_u32(ctxs[s], pool + r[i].off, outs[s], r[i].n, d, NULL, sts[s]) !=
                MDISCR_OK ||
            hipEventRecord(ev[i], sts[s]) != hipSuccess)
            exit(1);
        __atomic_store_n(&pl.submitted, i + 1, __ATOMIC_RELEASE);
    }
    pthread_join(th, NULL);
    double* lat = (double*)malloc(nreq * sizeof(double));
    double t_last = t0;
    for (int i = 0; i < nreq; i++) {
        lat[i] = done_t[i] - (t0 + r[i].t_arrive);
        if (done_t[i] > t_last) t_last = done_t[i];
    }
    res.p50_ms = pct(lat, nreq, 0.5) * 1e3;
    res.p99_ms = pct(lat, nreq, 0.99) * 1e3;
    res.gelem_s = g_total / (t_last - t0) / 1e9;
    res.identical = 1;
    free(lat);
    free(done_t);
    return res;
}

static void schedule(Req* r, int nreq, double lam, uint64_t seed) {
    double t = 0;
    for (int i = 0; i < nreq; i++) {
        if (lam > 0) t += -log(1.0 - unit(seed, (uint64_t)i)) / lam;
        r[i].t_arrive = t;
    }
}

static void print_result(const char* name, Result x, int last) {
    printf("\"%s\": {\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"gelem_s\": %.2f", name, x.p50_ms,
           x.p99_ms, x.gelem_s);
    if (x.batches)
        printf(", \"batches\": %llu, \"mean_batch_requests\": %.2f, \"peak_gib\": %.3f",
               (unsigned long long)x.batches, x.mean_batch, x.peak_gib);
    printf("}%s", last ? "" : ", ");
}

int main(int argc, char** argv) {
    const int nreq = argc > 1 ? atoi(argv[1]) : 50000;
    const double load = argc > 2 ? atof(argv[2]) : 0.8;
    const double budget_gib = argc > 3 ? atof(argv[3]) : 64.0;
    const int nstreams = argc > 4 ? atoi(argv[4]) : 4;
    const double reserve_gib = argc > 5 ? atof(argv[5]) : 16.0;
    const uint64_t seed = 0x5EED0000ull + 5;

    uint32_t* h_pool = (uint32_t*)malloc(POOL * 4);
    for (uint64_t i = 0; i < POOL; i++) h_pool[i] = (uint32_t)(sm64(seed, i) >> 33);
    Req* r = (Req*)calloc(nreq, sizeof(Req));
    const double lo = log(1024.0), hi = log(16777216.0);
    for (int i = 0; i < nreq; i++) {
        static const uint32_t ms[4] = {2, 4, 8, 16};
        r[i].n = (uint64_t)llround(exp(lo + unit(seed + 1, 3 * (uint64_t)i) * (hi - lo)));
        r[i].m = ms[sm64(seed + 1, 3 * (uint64_t)i + 1) % 4];
        r[i].off = sm64(seed + 1, 3 * (uint64_t)i + 2) % (POOL - r[i].n + 1);
        g_total += r[i].n;
    }
    mdiscr_ctx* ctx = NULL;
    if (mdiscr_ctx_create(&ctx, 0) != MDISCR_OK) return 1;
    hipStream_t st;
    uint32_t *d_pool, *d_plain;
    if (hipStreamCreate(&st) != hipSuccess || hipMalloc((void**)&d_pool, POOL * 4) != hipSuccess ||
        hipMalloc((void**)&d_plain, (16ull << 20) * 4) != hipSuccess)
        return 1;
    hipMemcpy(d_pool, h_pool, POOL * 4, hipMemcpyHostToDevice);
    mdiscr_server_config cfg = {0, (uint64_t)(budget_gib * (1ull << 30)), (uint32_t)nstreams,
                                1024, 1ull << 28, 0, (uint64_t)(reserve_gib * (1ull << 30))};

    /* the plain path on nstreams streams (one context each, no host wait) */
    mdiscr_ctx** pctx = (mdiscr_ctx**)calloc(nstreams, sizeof(mdiscr_ctx*));
    hipStream_t* psts = (hipStream_t*)calloc(nstreams, sizeof(hipStream_t));
    uint32_t** pouts = (uint32_t**)calloc(nstreams, sizeof(uint32_t*));
    for (int k = 0; k < nstreams; k++)
        if (mdiscr_ctx_create(&pctx[k], 0) != MDISCR_OK ||
            hipStreamCreateWithFlags(&psts[k], hipStreamNonBlocking) != hipSuccess ||
            hipMalloc((void**)&pouts[k], (16ull << 20) * 4) != hipSuccess)
            return 1;
    hipEvent_t* pev = (hipEvent_t*)calloc(nreq, sizeof(hipEvent_t));
    for (int i = 0; i < nreq; i++)
        if (hipEventCreateWithFlags(&pev[i], hipEventDisableTiming) != hipSuccess) return 1;

    /* capacities: everything submitted at t = 0 */
    schedule(r, nreq, 0, 0);
    (void)run_plain(r, nreq < 2000 ? nreq : 2000, d_pool, ctx, st, d_plain);  /* warm-up */
    (void)run_plain_streams(r, nreq < 2000 ? nreq : 2000, d_pool, nstreams, pctx, psts, pouts, pev);
    const Result cap_plain = run_plain(r, nreq, d_pool, ctx, st, d_plain);
    const Result cap_streams = run_plain_streams(r, nreq, d_pool, nstreams, pctx, psts, pouts, pev);
    const Result cap_srv = run_server(r, nreq, d_pool, &cfg, ctx, st, d_plain);
    const double mean_n = (double)g_total / nreq;
    fprintf(stderr, "capacity: plain %.1f, plain on %d streams %.1f, server %.1f Gelem/s\n",
            cap_plain.gelem_s, nstreams, cap_streams.gelem_s, cap_srv.gelem_s);

    printf("{\"nreq\": %d, \"total_elems\": %llu, \"load\": %.2f, \"budget_gib\": %.1f, "
           "\"nstreams\": %d, \"capacity_gelem_s\": {\"plain\": %.2f, \"plain_streams\": %.2f, "
           "\"server\": %.2f}, ",
           nreq, (unsigned long long)g_total, load, budget_gib, nstreams, cap_plain.gelem_s,
           cap_streams.gelem_s, cap_srv.gelem_s);
    int identical = cap_srv.identical;
    const char* names[2] = {"at_plain_capacity", "at_server_capacity"};
    const double caps[2] = {cap_plain.gelem_s, cap_srv.gelem_s};
    for (int k = 0; k < 2; k++) {
        const double lam = load * caps[k] * 1e9 / mean_n;
        schedule(r, nreq, lam, seed + 2 + k);
        const Result s = run_server(r, nreq, d_pool, &cfg, ctx, st, d_plain);
        const Result p = run_plain(r, nreq, d_pool, ctx, st, d_plain);
        const Result q = run_plain_streams(r, nreq, d_pool, nstreams, pctx, psts, pouts, pev);
        identical &= s.identical;
        printf("\"%s\": {\"arrival_rate_req_s\": %.1f, \"offered_gelem_s\": %.2f, ", names[k], lam,
               load * caps[k]);
        print_result("server", s, 0);
        print_result("plain", p, 0);
        print_result("plain_streams", q, 1);
        printf("}, ");
    }
    printf("\"identical_spot_checks\": %s}\n", identical ? "true" : "false");
    for (int i = 0; i < nreq; i++) hipEventDestroy(pev[i]);
    for (int k = 0; k < nstreams; k++) {
        mdiscr_ctx_destroy(pctx[k]);
        hipStreamDestroy(psts[k]);
        hipFree(pouts[k]);
    }
    mdiscr_ctx_destroy(ctx);
    return identical ? 0 : 3;
}
