// segmented.hip -- C ABI of the batched segmented partition.
//
// mdiscr_partition_segmented: segments of one array given by offsets.
// mdiscr_segm_flags: the reference's flag-array interface, mdiscrSegm(n, in,
//   segment_sizes, out, new_sizes) (MdiscrSegmHost.cu.h:36-42).
// Both run descriptor -> plan -> count -> scan -> partition.  An offsets call
// on a non-null stream is captured once into a HIP graph (keyed by its
// arguments and the context workspace) and replayed afterwards.  The (block size, elements per thread, staged or
// direct scatter) configuration is picked by timing every candidate on the
// first call of each (discriminator kind, class bits, log2 size) bucket;
// choices persist in the file named by MDISCR_TUNE_CACHE when set.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>

#include "api_internal.hpp"
#include "segmented_launch.hpp"

using namespace mdiscr;

// Batches of one-tile requests skip the count and scan launches (1).
#ifndef MDISCR_BATCH_SKIP_ROWS
#define MDISCR_BATCH_SKIP_ROWS 1
#endif

namespace {

template <bool MASK>
bool launch_any(uint32_t nbits, int cfg, const SegArgs& a, hipStream_t s) {
    switch (nbits) {
        case 0: return launch_nbits<0, MASK>(cfg, a, s);
        case 1: return launch_nbits<1, MASK>(cfg, a, s);
        case 2: return launch_nbits<2, MASK>(cfg, a, s);
        case 3: return launch_nbits<3, MASK>(cfg, a, s);
        case 4: return launch_nbits<4, MASK>(cfg, a, s);
        case 5: return launch_nbits<5, MASK>(cfg, a, s);
        case 6: return launch_nbits<6, MASK>(cfg, a, s);
    }
    return false;
}

bool config_available(int cfg) { return cfg >= 0 && cfg < kNumConfigs; }

// ---- per-context state: tuning table and captured graphs ------------------

// A captured graph holds every pointer it was captured with, the context
// workspace's included: the key carries the workspace base, so that a grown
// (reallocated) workspace is captured anew instead of replayed freed.
struct GraphKey {
    std::tuple<const void*, const void*, const void*, const void*, uint64_t, uint64_t, uint32_t,
               uint32_t, int, void*, const void*>
        k;
    bool operator<(const GraphKey& o) const { return k < o.k; }
};

struct SegState {
    std::map<std::string, int> tuned;  // bucket -> config
    bool tune_loaded = false;
    std::map<GraphKey, hipGraphExec_t> graphs;
    void* flag_ws = nullptr;
    size_t flag_ws_bytes = 0;
    void* cnt_ws = nullptr;
    size_t cnt_ws_bytes = 0;
    void* host_desc = nullptr;       // pinned descriptor staging (batches)
    void* host_desc_dev = nullptr;   // its device-side address
    size_t host_desc_bytes = 0;
    hipEvent_t staged_evt = nullptr;  // last upload out of host_desc
};

SegState* seg_state(mdiscr_ctx* ctx) {
    if (!ctx->extra) ctx->extra = new SegState();
    return static_cast<SegState*>(ctx->extra);
}

// The tuning bucket of a call: element type, discriminator kind, class count
// and the sizes rounded down to powers of two.
std::string bucket_key(bool mask, uint32_t m, uint64_t n, uint64_t nseg) {
    int lg = 0;
    while ((1ull << (lg + 1)) <= n) ++lg;
    int ls = 0;
    while ((1ull << (ls + 1)) <= nseg) ++ls;
    char buf[96];
    snprintf(buf, sizeof(buf), "u32 %s m%u n%d s%d", mask ? "mask" : "mod", m, lg, ls);
    return buf;
}

void tune_load(SegState* st) {
    if (st->tune_loaded) return;
    st->tune_loaded = true;
    const char* path = getenv("MDISCR_TUNE_CACHE");
    if (!path) return;
    FILE* f = fopen(path, "r");
    if (!f) return;
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        char* eq = strrchr(line, '=');
        if (!eq) continue;
        *eq = 0;
        const int cfg = atoi(eq + 1);
        if (cfg >= 0 && cfg < kNumConfigs) st->tuned[line] = cfg;
    }
    fclose(f);
}

void tune_save(SegState* st) {
    const char* path = getenv("MDISCR_TUNE_CACHE");
    if (!path) return;
    FILE* f = fopen(path, "w");
    if (!f) return;
    for (auto& kv : st->tuned)
        fprintf(f, "%s=%d %s\n", kv.first.c_str(), kv.second, kConfigs[kv.second].name);
    fclose(f);
}

struct SegWorkspace {
    SegDesc* desc;
    SegItem* items;
    SegLong* longs;
    SegPlan* plan;
    uint32_t* counts;
    uint32_t* prefix;
    uint32_t ld;
    uint32_t max_rows;
};

// Workspace for any candidate configuration (sized by the smallest tile).
mdiscr_status seg_workspace(mdiscr_ctx* ctx, uint64_t n, uint64_t nseg, uint32_t m,
                            SegWorkspace& w) {
    const uint64_t max_items = 2 * nseg + n / kMinTile + 2;
    const uint64_t max_rows = nseg + n / kMinTile + 2;
    const uint64_t ld = (max_rows + 1 + 3) & ~3ull;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_desc = take(nseg * sizeof(SegDesc));
    const size_t o_items = take(max_items * sizeof(SegItem));
    const size_t o_longs = take(nseg * sizeof(SegLong));
    const size_t o_plan = take(sizeof(SegPlan));
    const size_t o_counts = take((size_t)m * ld * sizeof(uint32_t));
    const size_t o_prefix = take((size_t)m * ld * sizeof(uint32_t));
    void* base = nullptr;
    mdiscr_status s = ctx_workspace(ctx, off, &base);
    if (s) return s;
    char* b = static_cast<char*>(base);
    w.desc = reinterpret_cast<SegDesc*>(b + o_desc);
    w.items = reinterpret_cast<SegItem*>(b + o_items);
    w.longs = reinterpret_cast<SegLong*>(b + o_longs);
    w.plan = reinterpret_cast<SegPlan*>(b + o_plan);
    w.counts = reinterpret_cast<uint32_t*>(b + o_counts);
    w.prefix = reinterpret_cast<uint32_t*>(b + o_prefix);
    w.ld = (uint32_t)ld;
    w.max_rows = (uint32_t)max_rows;
    return MDISCR_OK;
}

uint32_t grid_cap(uint64_t max_items) {
    return (uint32_t)(max_items > 0xFFFFFFFFull ? 0xFFFFFFFFull : (max_items ? max_items : 1));
}

mdiscr_status run_segmented(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                            const uint64_t* d_seg_off, uint64_t nseg, mdiscr_discr d,
                            uint32_t* d_seg_counts, hipStream_t s, int cfg, bool allow_graph) {
    const bool mask = d.kind == MDISCR_DISCR_MASK || (d.m & (d.m - 1)) == 0;
    const uint32_t nbits = class_bits(d.m);
    SegWorkspace w;
    mdiscr_status st = seg_workspace(ctx, n, nseg, d.m, w);
    if (st) return st;
    SegArgs a;
    a.desc = w.desc;
    a.items = w.items;
    a.longs = w.longs;
    a.plan = w.plan;
    a.counts = w.counts;
    a.prefix = w.prefix;
    a.ld = w.ld;
    a.max_rows = w.max_rows;
    a.nseg = (uint32_t)nseg;
    a.mrows = d.m;
    a.grid_cap = grid_cap(2 * nseg + n / kMinTile + 2);
    a.item_cap = a.grid_cap;
    a.num_cu = (uint32_t)ctx->num_cu;
    a.os = OffSrc{d_seg_off, d_in, d_out, d_seg_counts, d.m, mdiscr_divmagic_make(d.m), n};

    auto issue = [&](hipStream_t q) -> bool {
        return mask ? launch_any<true>(nbits, cfg, a, q) : launch_any<false>(nbits, cfg, a, q);
    };

    SegState* ss = seg_state(ctx);
    if (allow_graph && s != nullptr && !ctx->timing) {
        GraphKey key{std::make_tuple((const void*)d_in, (const void*)d_out, (const void*)d_seg_off,
                                     (const void*)d_seg_counts, n, nseg, d.m, d.kind, cfg,
                                     (void*)s, (const void*)w.desc)};
        auto it = ss->graphs.find(key);
        if (it == ss->graphs.end()) {
            if (ss->graphs.size() >= 32) {
                for (auto& kv : ss->graphs) (void)hipGraphExecDestroy(kv.second);
                ss->graphs.clear();
            }
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return ctx_hip(ctx, e, "hipStreamBeginCapture");
            const bool ok = issue(s);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(s, &graph);
            if (!ok) {
                if (graph) (void)hipGraphDestroy(graph);
                return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "segmented config %d", cfg);
            }
            if (e != hipSuccess) return ctx_hip(ctx, e, "hipStreamEndCapture");
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return ctx_hip(ctx, e, "hipGraphInstantiate");
            it = ss->graphs.emplace(key, exec).first;
        }
        return ctx_hip(ctx, hipGraphLaunch(it->second, s), "hipGraphLaunch");
    }
    TimedLaunch tl;
    timing_begin(ctx, s, tl);
    const bool ok = issue(s);
    timing_end(ctx, s, tl, "segmented");
    if (!ok) return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "segmented config %d", cfg);
    return ctx_hip(ctx, hipGetLastError(), "segmented launch");
}

// MDISCR_SEG_CONFIG: the configuration every segmented and batch call uses
// where it is available, or -1.  Read on every call, so a process can change it.
int forced_config() {
    const char* f = getenv("MDISCR_SEG_CONFIG");
    if (!f) return -1;
    const int c = atoi(f);
    return c >= 0 && c < kNumConfigs ? c : -1;
}

int pick_config(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                const uint64_t* d_seg_off, uint64_t nseg, mdiscr_discr d, hipStream_t s,
                mdiscr_status* err) {
    *err = MDISCR_OK;
    const bool mask = d.kind == MDISCR_DISCR_MASK || (d.m & (d.m - 1)) == 0;
    if (const int c = forced_config(); c >= 0 && config_available(c)) return c;
    const char* at = getenv("MDISCR_AUTOTUNE");
    if (at && at[0] == '0') return 0;
    SegState* ss = seg_state(ctx);
    tune_load(ss);
    const std::string key = bucket_key(mask, d.m, n, nseg);
    auto it = ss->tuned.find(key);
    if (it != ss->tuned.end()) return it->second;
    // Time every candidate on this call's own data (all of them write the
    // same, exact result), best of 3 after one warm-up.
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 0;
    int best = 0;
    float best_ms = 1e30f;
    const bool saved_timing = ctx->timing;
    ctx->timing = false;
    for (int c = 0; c < kNumConfigs; ++c) {
        if (!config_available(c)) continue;
        float mn = 1e30f;
        for (int rep = 0; rep < 4; ++rep) {
            mdiscr_status st = ctx_hip(ctx, hipEventRecord(a, s), "hipEventRecord");
            if (!st)
                st = run_segmented(ctx, d_in, d_out, n, d_seg_off, nseg, d, nullptr, s, c, false);
            if (!st) st = ctx_hip(ctx, hipEventRecord(b, s), "hipEventRecord");
            if (!st) st = ctx_hip(ctx, hipEventSynchronize(b), "hipEventSynchronize");
            float ms = 0.f;
            if (!st) st = ctx_hip(ctx, hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
            if (st) {
                // A failed trial must not be remembered as the fastest choice.
                *err = st;
                ctx->timing = saved_timing;
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
                return 0;
            }
            if (rep > 0 && ms < mn) mn = ms;
        }
        if (mn < best_ms) {
            best_ms = mn;
            best = c;
        }
    }
    ctx->timing = saved_timing;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ss->tuned[key] = best;
    tune_save(ss);
    return best;
}

// ---- reference flag-array form --------------------------------------------

// Positions i with flags[i] != 0 become segment starts: per-block counts, a
// scan over blocks, then a scatter of the start positions.
constexpr int kFlagBlock = 1024;

__global__ void k_flag_count(const int32_t* __restrict__ flags, uint32_t n,
                             uint32_t* __restrict__ bc) {
    __shared__ uint32_t tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * kFlagBlock + threadIdx.x;
    const uint64_t bal = ballot64(i < n && flags[i] != 0);
    if (lane_id() == 0) atomicAdd(&tot, (uint32_t)__popcll(bal));
    __syncthreads();
    if (threadIdx.x == 0) bc[blockIdx.x] = tot;
}

__global__ void k_flag_scatter(const int32_t* __restrict__ flags, uint32_t n,
                               const uint32_t* __restrict__ bpre, uint64_t* __restrict__ off,
                               uint32_t* __restrict__ bad) {
    __shared__ uint32_t wcnt[kFlagBlock / kWave];
    const uint32_t i = blockIdx.x * kFlagBlock + threadIdx.x;
    const int32_t f = i < n ? flags[i] : 0;
    const uint64_t bal = ballot64(f != 0);
    const uint32_t wave = threadIdx.x / kWave, lane = lane_id();
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t pre = bpre[blockIdx.x];
    for (uint32_t w = 0; w < wave; ++w) pre += wcnt[w];
    if (f != 0) {
        off[pre + mbcnt64(bal)] = i;
        if (f < 0 || (uint64_t)i + (uint32_t)f > n) atomicAdd(bad, 1u);
    }
    if (i == 0 && n > 0 && f == 0) atomicAdd(bad, 1u);  // the first segment must start at 0
}

// Each segment's length must equal the flag at its start.
__global__ void k_flag_check(const int32_t* __restrict__ flags, uint64_t* __restrict__ off,
                             uint32_t nseg, uint32_t n, uint32_t* __restrict__ bad) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nseg) return;
    if (s == nseg) {
        off[nseg] = n;
        return;
    }
    const uint64_t end = (s + 1 < nseg) ? off[s + 1] : n;
    if ((uint64_t)flags[off[s]] != end - off[s]) atomicAdd(bad, 1u);
}

// new_sizes: class count at each class start of each segment (after a zero fill).
__global__ void k_seg_new_sizes(const uint64_t* __restrict__ off, uint32_t nseg,
                                const uint32_t* __restrict__ seg_counts, uint32_t m,
                                int32_t* __restrict__ new_sizes) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    uint64_t pos = off[s];
    for (uint32_t c = 0; c < m; ++c) {
        const uint32_t k = seg_counts[(size_t)s * m + c];
        if (k) new_sizes[pos] = (int32_t)k;
        pos += k;
    }
}

// Pinned descriptor staging for at least nreq batch requests (mapped, so
// the planner reads it directly).  The caller has waited for the last upload
// out of the old buffer.
mdiscr_status stage_capacity(mdiscr_ctx* ctx, SegState* ss, uint32_t nreq) {
    const size_t dbytes = (size_t)nreq * sizeof(SegDesc);
    if (dbytes <= ss->host_desc_bytes) return MDISCR_OK;
    if (ss->host_desc) (void)hipHostFree(ss->host_desc);
    ss->host_desc = nullptr;
    ss->host_desc_dev = nullptr;
    ss->host_desc_bytes = 0;
    hipError_t e = hipHostMalloc(&ss->host_desc, dbytes * 2, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&ss->host_desc_dev, ss->host_desc, 0);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return ctx_fail(ctx, MDISCR_ENOMEM, "pinned staging: %s", hipGetErrorString(e));
    }
    ss->host_desc_bytes = dbytes * 2;
    return MDISCR_OK;
}

}  // namespace

mdiscr_status ctx_reserve_batch(mdiscr_ctx* ctx, uint64_t max_elems, uint32_t max_reqs) {
    mdiscr_status st = ctx_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (st) return st;
    SegWorkspace w;
    st = seg_workspace(ctx, max_elems, max_reqs, MDISCR_MAX_CLASSES, w);
    if (st) return st;
    return stage_capacity(ctx, seg_state(ctx), max_reqs);
}

void ctx_release_extra(mdiscr_ctx* ctx) {
    if (ctx->small) (void)hipFree(ctx->small);
    ctx->small = nullptr;
    if (ctx->extra) {
        SegState* ss = static_cast<SegState*>(ctx->extra);
        for (auto& kv : ss->graphs) (void)hipGraphExecDestroy(kv.second);
        if (ss->flag_ws) (void)hipFree(ss->flag_ws);
        if (ss->cnt_ws) (void)hipFree(ss->cnt_ws);
        if (ss->host_desc) (void)hipHostFree(ss->host_desc);
        if (ss->staged_evt) (void)hipEventDestroy(ss->staged_evt);
        delete ss;
        ctx->extra = nullptr;
    }
}

extern "C" {

mdiscr_status mdiscr_partition_segmented(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out,
                                         uint64_t n, const uint64_t* d_seg_off,
                                         const uint64_t* h_seg_off, uint64_t nseg,
                                         mdiscr_discr d, uint32_t* d_seg_counts, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    if (n > MDISCR_MAX_ELEMS || nseg > 0xFFFFFFF0ull)
        return ctx_fail(ctx, MDISCR_EINVAL, "n or nseg too large");
    if (!d_seg_off) return ctx_fail(ctx, MDISCR_EINVAL, "null segment offsets");
    if (n && (!d_in || !d_out)) return ctx_fail(ctx, MDISCR_EINVAL, "null data pointer");
    if (n && ranges_overlap(d_in, d_out, n * 4))
        return ctx_fail(ctx, MDISCR_EINVAL, "output overlaps input");
    if (h_seg_off) {  // optional host-side validation of the offsets
        if (h_seg_off[0] != 0 || h_seg_off[nseg] != n)
            return ctx_fail(ctx, MDISCR_EINVAL, "offsets must run from 0 to n");
        for (uint64_t s = 0; s < nseg; ++s)
            if (h_seg_off[s + 1] < h_seg_off[s])
                return ctx_fail(ctx, MDISCR_EINVAL, "offsets must not decrease");
    }
    hipStream_t s = (hipStream_t)stream;
    st = ctx_begin(ctx, s);
    if (st) return st;
    if (nseg == 0) return MDISCR_OK;
    const int cfg = pick_config(ctx, d_in, d_out, n, d_seg_off, nseg, d, s, &st);
    if (st) return st;
    return run_segmented(ctx, d_in, d_out, n, d_seg_off, nseg, d, d_seg_counts, s, cfg, true);
}

mdiscr_status mdiscr_segm_flags(mdiscr_ctx* ctx, uint64_t n, const uint32_t* d_in,
                                const int32_t* d_segment_sizes, uint32_t* d_out,
                                int32_t* d_new_sizes, mdiscr_discr d, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    if (n > MDISCR_MAX_ELEMS) return ctx_fail(ctx, MDISCR_EINVAL, "n too large");
    if (n && (!d_in || !d_out || !d_segment_sizes))
        return ctx_fail(ctx, MDISCR_EINVAL, "null pointer");
    if (n && ranges_overlap(d_in, d_out, n * 4))
        return ctx_fail(ctx, MDISCR_EINVAL, "output overlaps input");
    hipStream_t s = (hipStream_t)stream;
    st = ctx_begin(ctx, s);
    if (st) return st;
    if (n == 0) return MDISCR_OK;
    SegState* ss = seg_state(ctx);
    // flag workspace: bad-flag counter | block counts (nb+1) | offsets (n+1)
    const uint32_t nb = (uint32_t)((n + kFlagBlock - 1) / kFlagBlock);
    const uint32_t nbl = (nb + 1 + 3) & ~3u;
    const size_t o_bc = 256;
    const size_t o_off = o_bc + (((size_t)nbl * 4 + 255) & ~(size_t)255);
    const size_t need = o_off + (n + 1) * 8;
    st = ctx_grow(ctx, &ss->flag_ws, &ss->flag_ws_bytes, need);
    if (st) return st;
    char* b = static_cast<char*>(ss->flag_ws);
    uint32_t* bad = reinterpret_cast<uint32_t*>(b);
    uint32_t* bc = reinterpret_cast<uint32_t*>(b + o_bc);
    uint64_t* off = reinterpret_cast<uint64_t*>(b + o_off);
    hipLaunchKernelGGL(k_fill_u32<>, dim3(1), dim3(64), 0, s, bad, 0u, 1u);
    hipLaunchKernelGGL(k_fill_u32<>, dim3((nbl + 255) / 256), dim3(256), 0, s, bc, 0u, nbl);
    hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(kFlagBlock), 0, s, d_segment_sizes,
                       (uint32_t)n, bc);
    // exclusive scan of the block counts; entry nb becomes the segment count
    launch_tile_scan(bc, bc, nb + 1, nbl, 1u, nullptr, nullptr, nullptr, s);
    hipLaunchKernelGGL(k_flag_scatter, dim3(nb), dim3(kFlagBlock), 0, s, d_segment_sizes,
                       (uint32_t)n, bc, off, bad);
    uint32_t nseg = 0;
    st = ctx_hip(ctx, hipMemcpyAsync(&nseg, bc + nb, 4, hipMemcpyDeviceToHost, s), "nseg readback");
    if (st) return st;
    st = ctx_hip(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
    if (st) return st;
    hipLaunchKernelGGL(k_flag_check, dim3((nseg + 1 + 255) / 256), dim3(256), 0, s, d_segment_sizes,
                       off, nseg, (uint32_t)n, bad);
    uint32_t nbad = 0;
    st = ctx_hip(ctx, hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s), "flag check");
    if (st) return st;
    st = ctx_hip(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
    if (st) return st;
    if (nbad) return ctx_fail(ctx, MDISCR_EINVAL, "segment_sizes do not tile [0,n)");
    st = ctx_grow(ctx, &ss->cnt_ws, &ss->cnt_ws_bytes, (size_t)nseg * d.m * 4 + 4);
    if (st) return st;
    uint32_t* seg_counts = static_cast<uint32_t*>(ss->cnt_ws);
    const int cfg = pick_config(ctx, d_in, d_out, n, off, nseg, d, s, &st);
    if (!st) st = run_segmented(ctx, d_in, d_out, n, off, nseg, d, seg_counts, s, cfg, false);
    if (!st && d_new_sizes) {
        hipLaunchKernelGGL(k_fill_u32<>, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                           (uint32_t*)d_new_sizes, 0u, (uint32_t)n);
        hipLaunchKernelGGL(k_seg_new_sizes, dim3((nseg + 255) / 256), dim3(256), 0, s, off, nseg,
                           seg_counts, d.m, d_new_sizes);
        st = ctx_hip(ctx, hipGetLastError(), "new_sizes launch");
    }
    if (st) return st;
    return ctx_hip(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
}

mdiscr_status mdiscr_partition_batch(mdiscr_ctx* ctx, const mdiscr_request* reqs, uint32_t nreq,
                                     void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    if (nreq && !reqs) return ctx_fail(ctx, MDISCR_EINVAL, "null request array");
    uint64_t total = 0;
    uint32_t mmax = 1;
    bool all_pow2 = true;
    for (uint32_t i = 0; i < nreq; ++i) {
        const mdiscr_request& r = reqs[i];
        mdiscr_status st = check_discr(ctx, r.d);
        if (st) return st;
        if (r.n > 0xFFFFFFFFull) return ctx_fail(ctx, MDISCR_EINVAL, "request %u: n too large", i);
        if (r.n && (!r.d_in || !r.d_out))
            return ctx_fail(ctx, MDISCR_EINVAL, "request %u: null data pointer", i);
        if (r.n && ranges_overlap(r.d_in, r.d_out, r.n * 4))
            return ctx_fail(ctx, MDISCR_EINVAL, "request %u: output overlaps input", i);
        total += r.n;
        if (r.d.m > mmax) mmax = r.d.m;
        if (r.d.kind != MDISCR_DISCR_MASK && (r.d.m & (r.d.m - 1))) all_pow2 = false;
    }
    if (total > MDISCR_MAX_ELEMS) return ctx_fail(ctx, MDISCR_EINVAL, "batch too large");
    hipStream_t s = (hipStream_t)stream;
    mdiscr_status st = ctx_begin(ctx, s);
    if (st || nreq == 0) return st;
    SegState* ss = seg_state(ctx);
    // Host descriptors in pinned staging; the previous batch's upload must
    // have left the staging buffer before it is rewritten.
    if (ss->staged_evt) {
        st = ctx_hip(ctx, hipEventSynchronize(ss->staged_evt), "hipEventSynchronize");
        if (st) return st;
    } else {
        st = ctx_hip(ctx, hipEventCreateWithFlags(&ss->staged_evt, hipEventDisableTiming),
                     "hipEventCreate");
        if (st) return st;
    }
    st = stage_capacity(ctx, ss, nreq);
    if (st) return st;
    SegDesc* hd = static_cast<SegDesc*>(ss->host_desc);
    for (uint32_t i = 0; i < nreq; ++i) {
        hd[i].in = reqs[i].d_in;
        hd[i].out = reqs[i].d_out;
        hd[i].counts = reqs[i].d_counts;
        hd[i].len = (uint32_t)reqs[i].n;
        hd[i].m = reqs[i].d.m;
        hd[i].g = mdiscr_divmagic_make(reqs[i].d.m);
    }
    SegWorkspace w;
    st = seg_workspace(ctx, total, nreq, mmax, w);
    if (st) return st;
    const uint32_t nbits = class_bits(mmax);
    int cfg = 0;
    auto it = ss->tuned.find(bucket_key(all_pow2, mmax, total, nreq));
    if (const int c = forced_config(); c >= 0 && config_available(c))
        cfg = c;
    else if (it != ss->tuned.end() && config_available(it->second))
        cfg = it->second;
    SegArgs a;
    a.staged = static_cast<const SegDesc*>(ss->host_desc_dev);  // the planner copies them
    hipError_t staged_err = hipSuccess;
    a.staged_read = ss->staged_evt;
    a.staged_err = &staged_err;
    a.desc = w.desc;
    a.items = w.items;
    a.longs = w.longs;
    a.plan = w.plan;
    a.counts = w.counts;
    a.prefix = w.prefix;
    a.ld = w.ld;
    a.max_rows = w.max_rows;
    a.nseg = nreq;
    a.mrows = mmax;
    {   // as many workgroups as the queue will hold, up to the resident capacity
        const uint64_t tile = (uint64_t)kConfigs[cfg].nw * kWave * kConfigs[cfg].r;
        const uint64_t wt = (uint64_t)kWave * kConfigs[cfg].r;
        uint64_t est = 0, nmax = 0;
        for (uint32_t i = 0; i < nreq; ++i) {
            est += reqs[i].n <= wt ? 1 : (reqs[i].n + tile - 1) / tile;
            nmax = reqs[i].n > nmax ? reqs[i].n : nmax;
        }
        a.grid_cap = grid_cap(est);
        // Requests of at most one tile each: every one is written from its
        // tile's LDS, with no count rows to count or scan (batches of 1 to 64
        // requests of 4096 keys: 0.031-0.032 -> 0.023-0.024 ms per call).
        a.rows = !MDISCR_BATCH_SKIP_ROWS || nmax > tile;
        a.item_cap = grid_cap(2 * (uint64_t)nreq + total / kMinTile + 2);
        a.num_cu = (uint32_t)ctx->num_cu;
    }
    TimedLaunch tl;
    timing_begin(ctx, s, tl);
    const bool ok = all_pow2 ? launch_any<true>(nbits, cfg, a, s) : launch_any<false>(nbits, cfg, a, s);
    timing_end(ctx, s, tl, "batch");
    if (!ok) return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "batch config");
    st = ctx_hip(ctx, hipGetLastError(), "batch launch");
    if (!st) st = ctx_hip(ctx, staged_err, "hipEventRecord");
    // After a failure the staging event may not cover this batch's planner:
    // drain the stream so that the next batch can restage safely.
    if (st) (void)hipStreamSynchronize(s);
    return st;
}

}  // extern "C"
