// mdiscr_api.hip -- C ABI of the gfx950 m-discriminator (include/mdiscr.h).
//
// The context owns one device workspace that grows to the largest request
// and is reused, so steady-state calls make no device allocation (the
// reference cudaMallocs and frees its temporaries on every call,
// MdiscrHost_optimized_clean.cu.h:43-45,145-147).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mdiscr.h"
#include "api_internal.hpp"
#include "plain_kernels.hpp"

using namespace mdiscr;

// ---------------------------------------------------------------------------
// context

mdiscr_status ctx_fail(mdiscr_ctx* ctx, mdiscr_status s, const char* fmt, ...) {
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
        va_end(ap);
    }
    return s;
}

mdiscr_status ctx_hip(mdiscr_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return MDISCR_OK;
    return ctx_fail(ctx, MDISCR_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

mdiscr_status ctx_grow(mdiscr_ctx* ctx, void** buf, size_t* have, size_t need) {
    need = (need + 255) & ~(size_t)255;
    if (need <= *have) return MDISCR_OK;
    // At least double, so that the retired buffers stay below the live one.
    const size_t bytes = std::max(need, *have * 2);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return ctx_fail(ctx, MDISCR_ENOMEM, "%zu-byte device buffer: %s", bytes,
                        hipGetErrorString(e));
    }
    // Queued work may still use the old buffer: it is freed with the context,
    // never here (hipFree would also wait for the whole device, stalling every
    // other stream).
    if (*buf) ctx->retired.push_back(*buf);
    *buf = p;
    *have = bytes;
    return MDISCR_OK;
}

mdiscr_status ctx_workspace(mdiscr_ctx* ctx, size_t bytes, void** out) {
    mdiscr_status st = ctx_grow(ctx, &ctx->ws, &ctx->ws_bytes, bytes);
    if (st) return st;
    *out = ctx->ws;
    return MDISCR_OK;
}

mdiscr_status ctx_begin(mdiscr_ctx* ctx, hipStream_t s) {
    mdiscr_status st = ctx_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (st) return st;
    // A call under stream capture would record its kernels into the caller's
    // graph with the context workspace's address baked in, and could grow
    // (reallocate) that workspace inside the capture.  Neither is supported:
    // the call is refused before it touches the device.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const hipError_t e = hipStreamIsCapturing(s, &cs);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "stream capture state unknown: %s",
                        hipGetErrorString(e));
    }
    if (cs != hipStreamCaptureStatusNone)
        return ctx_fail(ctx, MDISCR_EUNSUPPORTED,
                        "the stream is being captured into a graph; calls under capture are "
                        "not supported");
    return MDISCR_OK;
}

void timing_begin(mdiscr_ctx* ctx, hipStream_t s, TimedLaunch& t) {
    t.on = ctx->timing;
    if (!t.on) return;
    t.a = t.b = nullptr;
    // Timing is a diagnostic: a launch whose events cannot be made or
    // recorded runs untimed rather than failing.
    if (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess ||
        hipEventRecord(t.a, s) != hipSuccess) {
        (void)hipGetLastError();
        if (t.a) (void)hipEventDestroy(t.a);
        if (t.b) (void)hipEventDestroy(t.b);
        t.on = false;
    }
}

void timing_end(mdiscr_ctx* ctx, hipStream_t s, TimedLaunch& t, const char* name) {
    if (!t.on) return;
    if (hipEventRecord(t.b, s) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
        return;
    }
    t.name = name;
    ctx->pending.push_back(t);
}

// ---------------------------------------------------------------------------
// Fixed device scratch of the no-allocation entry point: per GPU one copy of
// `scratch`, resolved once per device at context creation.  Host code uses
// only the resolved addresses (the array's name is a host-side stand-in).

static __device__ __align__(256) uint32_t scratch[MDISCR_NOALLOC_SCRATCH_BYTES / 4];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), resolved on first use.
mdiscr_status scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return MDISCR_EUNSUPPORTED;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return MDISCR_EDEVICE;
        }
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return MDISCR_OK;
}
}  // namespace

extern "C" {

mdiscr_status mdiscr_ctx_create(mdiscr_ctx** out, int device) {
    if (!out) return MDISCR_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return MDISCR_EDEVICE;
    }
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) return MDISCR_EDEVICE;
    }
    if (device >= ndev) return MDISCR_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MDISCR_EDEVICE;
    mdiscr_ctx* ctx = new (std::nothrow) mdiscr_ctx();
    if (!ctx) return MDISCR_ENOMEM;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
    // Set-up of the no-allocation entry point's scratch on this device.
    const mdiscr_status st = scratch_address(device, &ctx->scratch);
    if (st != MDISCR_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return MDISCR_OK;
}

mdiscr_status mdiscr_ctx_destroy(mdiscr_ctx* ctx) {
    if (!ctx) return MDISCR_EINVAL;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    for (auto& t : ctx->pending) {
        hipEventDestroy(t.a);
        hipEventDestroy(t.b);
    }
    if (ctx->ws) hipFree(ctx->ws);
    for (void* p : ctx->retired) hipFree(p);
    ctx_release_extra(ctx);
    delete ctx;
    return MDISCR_OK;
}

mdiscr_status mdiscr_sync(mdiscr_ctx* ctx, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    return ctx_hip(ctx, hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

const char* mdiscr_status_str(mdiscr_status s) {
    switch (s) {
        case MDISCR_OK: return "ok";
        case MDISCR_EINVAL: return "invalid argument";
        case MDISCR_ENOMEM: return "out of memory";
        case MDISCR_EDEVICE: return "device error";
        case MDISCR_EBUSY: return "busy";
        case MDISCR_EUNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

const char* mdiscr_last_error(const mdiscr_ctx* ctx) {
    return ctx ? ctx->err : "no context";
}

mdiscr_status mdiscr_timing_enable(mdiscr_ctx* ctx, int enable) {
    if (!ctx) return MDISCR_EINVAL;
    ctx->timing = enable != 0;
    return MDISCR_OK;
}

mdiscr_status mdiscr_timing_read(mdiscr_ctx* ctx, const char* name, double* total_ms,
                                 uint64_t* launches) {
    if (!ctx || !name || !total_ms || !launches) return MDISCR_EINVAL;
    // Fold every finished record into the per-name totals, then report one.
    size_t done = 0;
    hipError_t e = hipSuccess;
    for (; done < ctx->pending.size(); ++done) {
        const TimedLaunch& t = ctx->pending[done];
        e = hipEventSynchronize(t.b);
        if (e != hipSuccess) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) != hipSuccess) (void)hipGetLastError();
        auto& acc = ctx->timing_acc[t.name];
        acc.first += ms;
        acc.second += 1;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    // Records already folded (and their events destroyed) leave the list even
    // when a later one fails, so a retry never touches a destroyed event.
    ctx->pending.erase(ctx->pending.begin(), ctx->pending.begin() + done);
    if (e != hipSuccess) return ctx_hip(ctx, e, "hipEventSynchronize");
    auto it = ctx->timing_acc.find(name);
    if (it == ctx->timing_acc.end()) {
        *total_ms = 0.0;
        *launches = 0;
    } else {
        *total_ms = it->second.first;
        *launches = it->second.second;
        ctx->timing_acc.erase(it);
    }
    return MDISCR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// plain partition

namespace {

#ifndef MDISCR_PLAIN_NW
#define MDISCR_PLAIN_NW 4
#endif
#ifndef MDISCR_PLAIN_R
#define MDISCR_PLAIN_R 16
#endif
// Tiles per count workgroup.  One tile each reads fastest (2^28 keys, m=8,
// interleaved A/B: count 0.1515 / 0.1523 / 0.1572 / 0.1629 ms for 1 / 2 / 4 /
// 8 tiles; whole call 0.538 vs 0.547 ms for 1 vs 8).
#ifndef MDISCR_COUNT_CT
#define MDISCR_COUNT_CT 1
#endif
#ifndef MDISCR_PLAIN_PREFETCH
#define MDISCR_PLAIN_PREFETCH 0
#endif
#ifndef MDISCR_PLAIN_WB
#define MDISCR_PLAIN_WB 8
#endif
#ifndef MDISCR_PLAIN_MINW
#define MDISCR_PLAIN_MINW 1
#endif
// The ballot-ranked path's tile up to 16 classes (an input that is not
// 16-byte aligned): kNW waves x kR rounds, one tile per partition workgroup.
constexpr int kNW = MDISCR_PLAIN_NW;  // waves per workgroup
constexpr int kR = MDISCR_PLAIN_R;    // rounds (elements per lane per tile)
constexpr int kCCT = MDISCR_COUNT_CT; // tiles per count workgroup
constexpr bool kPrefetch = MDISCR_PLAIN_PREFETCH != 0;
constexpr uint32_t kTile = Geom<kNW, kR, 1>::kTile;  // the smallest plain tile

// The quad path (16-byte loads, packed-counter ranks) serves up to 16
// classes on a 16-byte aligned input, in tiles of its own size; the
// ballot-ranked kernel serves the rest.  MDISCR_PLAIN_QUAD = 2 also stages
// class runs aligned to the output for 16-byte stores, 1 keeps 4-byte
// stores, 0 keeps the ballot-ranked kernel everywhere.
#ifndef MDISCR_PLAIN_QUAD
#define MDISCR_PLAIN_QUAD 1
#endif
#ifndef MDISCR_QUAD_NW
#define MDISCR_QUAD_NW 4
#endif
#ifndef MDISCR_QUAD_R
#define MDISCR_QUAD_R 32
#endif
// Waves of the quad path's count workgroup: one tile each, 16 rounds of
// 16-byte loads per lane (more rounds per lane read slower).
#ifndef MDISCR_QUAD_COUNT_NW
#define MDISCR_QUAD_COUNT_NW 8
#endif
// Quad partition workgroups resident per CU (0: as many as fit).  The kernel
// needs few registers, but more resident tiles interleave more read and
// write streams at once, which HBM serves more slowly: the launch pads its
// LDS allocation to hold the count at this many.  2^28 keys, m = 8, 8192-key
// tiles, interleaved A/B: 0.360 ms at 2 per CU against 0.377 ms for 4096-key
// tiles at 4 per CU and 0.439 / 0.377 / 0.384 / 0.402 ms at 2 / 4 / 5 / 6.
#ifndef MDISCR_PLAIN_WG_PER_CU
#define MDISCR_PLAIN_WG_PER_CU 2
#endif
constexpr bool kQuadOut16 = MDISCR_PLAIN_QUAD == 2;
constexpr int kQNW = MDISCR_QUAD_NW;
constexpr int kQR = MDISCR_QUAD_R;
constexpr uint32_t kQTile = (uint32_t)kQNW * kWave * kQR;
constexpr int kQCNW = MDISCR_QUAD_COUNT_NW;
constexpr int kQCR = (int)(kQTile / (kQCNW * kWave));
static_assert(kQCR * kQCNW * kWave == (int)kQTile, "count and partition tiles agree");
static_assert(kQTile >= kTile, "the workspace is sized by the smaller tile");

// The count and tile scan of a call in tiles of CNW * 64 * CR elements;
// counts, prefix and totals are laid out in the workspace as above.
template <int CNW, int CR, int NBITS, class Discr>
void launch_count_scan(mdiscr_ctx* ctx, const uint32_t* in, uint32_t n, Discr dis, uint32_t m,
                       uint32_t* user_counts, hipStream_t s, uint32_t* counts, uint32_t* prefix,
                       uint32_t* totals, uint32_t T, uint32_t ld) {
    TimedLaunch tl;
    const uint32_t Kc = (T + kCCT - 1) / kCCT;
    timing_begin(ctx, s, tl);
    if (((uintptr_t)in & 15u) == 0)
        hipLaunchKernelGGL((k_chunk_count<CNW, CR, kCCT, NBITS, true, Discr>), dim3(Kc),
                           dim3(CNW * kWave), 0, s, in, n, dis, counts, ld, m);
    else
        hipLaunchKernelGGL((k_chunk_count<CNW, CR, kCCT, NBITS, false, Discr>), dim3(Kc),
                           dim3(CNW * kWave), 0, s, in, n, dis, counts, ld, m);
    timing_end(ctx, s, tl, "chunk_count");
    timing_begin(ctx, s, tl);
    launch_tile_scan(counts, prefix, T, ld, m, totals, user_counts, nullptr, s);
    timing_end(ctx, s, tl, "tile_scan");
}

// The quad path: count, scan and partition in kQTile tiles, or the
// partition alone for an input of one tile (it counts that tile itself).
template <int NBITS, class Discr>
mdiscr_status launch_plain_quad(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                                Discr dis, uint32_t m, uint32_t* user_counts, uint64_t* offsets,
                                hipStream_t s, uint32_t* ws) {
    const uint32_t T = (uint32_t)(((uint64_t)n + kQTile - 1) / kQTile);
    const uint32_t ld = (T + 3u) & ~3u;
    uint32_t* counts = ws;
    uint32_t* prefix = ws + (size_t)m * ld;
    uint32_t* totals = ws + (size_t)2 * m * ld;
    TimedLaunch tl;
    if (T == 1) {
        timing_begin(ctx, s, tl);
        hipLaunchKernelGGL((k_partition_quad<kQNW, kQR, NBITS, MDISCR_PLAIN_WB, kQuadOut16,
                                             MDISCR_PLAIN_MINW, Discr, true>),
                           dim3(1), dim3(kQNW * kWave), 0, s, in, out, n, dis, nullptr, nullptr,
                           0u, m, user_counts, offsets);
        timing_end(ctx, s, tl, "partition");
        return ctx_hip(ctx, hipGetLastError(), "plain partition launch");
    }
    launch_count_scan<kQCNW, kQCR, NBITS>(ctx, in, n, dis, m, user_counts, s, counts, prefix,
                                          totals, T, ld);
    constexpr auto kern = k_partition_quad<kQNW, kQR, NBITS, MDISCR_PLAIN_WB, kQuadOut16,
                                           MDISCR_PLAIN_MINW, Discr>;
    static const size_t pad = lds_pad_for(kern, MDISCR_PLAIN_WG_PER_CU);
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL(kern, dim3(T), dim3(kQNW * kWave), pad, s, in, out, n, dis, prefix, totals,
                       ld, m, nullptr, offsets);
    timing_end(ctx, s, tl, "partition");
    return ctx_hip(ctx, hipGetLastError(), "plain partition launch");
}

// The two-digit quad path (k_partition_lsd) on a 16-byte aligned input of
// more than one of its tiles: NW waves x R keys per lane, WGCU workgroups
// resident per CU (0: as many as fit, which is 4 by its 35 KB of LDS).
// Unlike the one-pass quad kernel it wants every workgroup that fits: its
// five barriers per tile leave a lone pair of workgroups idle.  2^28 keys,
// partition ms at m = 64 and m = 32: 0.409 and 0.397 at 4 per CU, 0.544 and
// 0.529 at 2 (3 per CU measured as 2); 8-wave tiles of 16 rounds 0.418 and
// 0.400; 4096-key tiles 0.438 and 0.404; the ballot kernel 0.505 and 0.442.
// MDISCR_PLAIN_LSD = 0 keeps the ballot-ranked kernel.
#ifndef MDISCR_PLAIN_LSD
#define MDISCR_PLAIN_LSD 1
#endif
#ifndef MDISCR_LSD_NW
#define MDISCR_LSD_NW 4
#endif
// Under x % m the two-digit kernel carries each key's class from its first
// pass to the second and to the write-out as a byte in LDS (x % m once per
// key instead of three times), in tiles of MDISCR_LSD_CLS_R keys per lane.
#ifndef MDISCR_LSD_CLS
#define MDISCR_LSD_CLS 1
#endif
#ifndef MDISCR_LSD_CLS_R
#define MDISCR_LSD_CLS_R 24
#endif
// ... and from 17 classes in tiles of 16 keys per lane (4096 keys): 2^28 keys,
// partition x % 33 0.463 -> 0.432 ms, x % 48 0.471 -> 0.442 ms; at 9..16
// classes the 6144-key count is faster (x % 13: call 0.601 against 0.605 ms).
#ifndef MDISCR_LSD_CLS_R_WIDE
#define MDISCR_LSD_CLS_R_WIDE 16
#endif
#ifndef MDISCR_LSD_R
#define MDISCR_LSD_R 32
#endif
#ifndef MDISCR_LSD_WG_PER_CU
#define MDISCR_LSD_WG_PER_CU 0
#endif

// Where the two-digit path serves: 17..64 classes, 16 classes under the mask
// (2^28 keys: 0.412 -> 0.386 ms), and 9..16 under the modulus when it carries
// the classes (m = 9 / 13: 0.437 / 0.439 ms one-pass against 0.398 / 0.406;
// evaluating x % m in every pass instead, it lost: 0.463 / 0.463 against
// 0.466 / 0.470).
#ifndef MDISCR_LSD_MOD16
#define MDISCR_LSD_MOD16 1
#endif
template <int NBITS, class Discr>
constexpr bool use_lsd() {
    return MDISCR_PLAIN_LSD != 0 &&
           (NBITS >= 5 ||
            (NBITS == 4 && (std::is_same<Discr, DiscrMask>::value ||
                            (MDISCR_LSD_MOD16 && MDISCR_LSD_CLS && CostlyDiscr<Discr>::value))));
}
// The two-digit kernel's geometry for a discriminator.
template <class Discr, int NBITS>
struct LsdGeom {
    static constexpr bool CLS = MDISCR_LSD_CLS && CostlyDiscr<Discr>::value;
    static constexpr int NW = MDISCR_LSD_NW;
    static constexpr int R =
        CLS ? (NBITS >= 5 ? MDISCR_LSD_CLS_R_WIDE : MDISCR_LSD_CLS_R) : MDISCR_LSD_R;
    static constexpr uint32_t TILE = (uint32_t)NW * kWave * R;
    static_assert(TILE >= kTile, "the workspace is sized by the smaller tile");
};

template <int NBITS, class Discr>
mdiscr_status launch_plain_lsd(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                               Discr dis, uint32_t m, uint32_t* user_counts, uint64_t* offsets,
                               hipStream_t s, uint32_t* ws) {
    using G = LsdGeom<Discr, NBITS>;
    constexpr int NW = G::NW, R = G::R;
    constexpr uint32_t TILE = G::TILE;
    constexpr int CNW = 8;
    constexpr int CR = (int)(TILE / (CNW * kWave));
    static_assert(CR * CNW * kWave == (int)TILE, "count and partition tiles agree");
    const uint32_t T = (uint32_t)(((uint64_t)n + TILE - 1) / TILE);
    const uint32_t ld = (T + 3u) & ~3u;
    uint32_t* counts = ws;
    uint32_t* prefix = ws + (size_t)m * ld;
    uint32_t* totals = ws + (size_t)2 * m * ld;
    launch_count_scan<CNW, CR, NBITS>(ctx, in, n, dis, m, user_counts, s, counts, prefix, totals,
                                      T, ld);
    constexpr auto kern =
        k_partition_lsd<NW, R, NBITS, MDISCR_PLAIN_WB, MDISCR_PLAIN_MINW, Discr, G::CLS>;
    static const size_t pad = lds_pad_for(kern, MDISCR_LSD_WG_PER_CU);
    TimedLaunch tl;
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL(kern, dim3(T), dim3(NW * kWave), pad, s, in, out, n, dis, prefix, totals,
                       ld, m, offsets);
    timing_end(ctx, s, tl, "partition");
    return ctx_hip(ctx, hipGetLastError(), "plain partition launch");
}

// Geometry of the ballot-ranked path from MDISCR_BIG_MIN_M classes: NW waves
// x R rounds per tile, CNW count waves per tile, WGCU partition workgroups
// per CU (0: as many as fit).  With many classes a 4096-key tile's class runs
// are short (64 keys at m = 64); 8192-key tiles of 8 waves, two per CU, write
// longer runs.  2^28 keys, interleaved A/B, partition ms for 4096-key tiles
// (4 waves, no cap) against this geometry: m = 64 0.638 / 0.531, m = 33
// 0.536 / 0.577, m = 32 0.483 / 0.484.
#ifndef MDISCR_BIG_MIN_M
#define MDISCR_BIG_MIN_M 48
#endif
#ifndef MDISCR_BIG_NW
#define MDISCR_BIG_NW 8
#endif
#ifndef MDISCR_BIG_R
#define MDISCR_BIG_R 16
#endif
#ifndef MDISCR_BIG_COUNT_NW
#define MDISCR_BIG_COUNT_NW 8
#endif
#ifndef MDISCR_BIG_WG_PER_CU
#define MDISCR_BIG_WG_PER_CU 2
#endif

// The ballot-ranked path in tiles of NW * 64 * R elements (count in the same
// tiles with CNW waves), or the partition alone for an input of one tile.
template <int NW, int R, int CNW, int WGCU, int NBITS, class Discr>
mdiscr_status launch_plain_ballot(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                                  Discr dis, uint32_t m, uint32_t* user_counts,
                                  uint64_t* offsets, hipStream_t s, uint32_t* ws) {
    constexpr uint32_t TILE = (uint32_t)NW * kWave * R;
    constexpr int CR = (int)(TILE / (CNW * kWave));
    static_assert(CR * CNW * kWave == (int)TILE, "count and partition tiles agree");
    static_assert(TILE >= kTile, "the workspace is sized by the smallest tile");
    const uint32_t T = (uint32_t)(((uint64_t)n + TILE - 1) / TILE);
    const uint32_t ld = (T + 3u) & ~3u;     // 16-byte aligned class rows
    uint32_t* counts = ws;                      // m * ld
    uint32_t* prefix = ws + (size_t)m * ld;     // m * ld
    uint32_t* totals = ws + (size_t)2 * m * ld; // m
    TimedLaunch tl;
    if (T == 1) {
        // One tile: the partition counts it itself (3 launches -> 1; a call on
        // one tile or less is bound by launch latency, not bandwidth).
        timing_begin(ctx, s, tl);
        hipLaunchKernelGGL((k_partition<NW, R, 1, NBITS, kPrefetch, MDISCR_PLAIN_WB,
                                        MDISCR_PLAIN_MINW, Discr, true>),
                           dim3(1), dim3(NW * kWave), 0, s, in, out, n, dis, nullptr, nullptr,
                           0u, m, user_counts, offsets);
        timing_end(ctx, s, tl, "partition");
        return ctx_hip(ctx, hipGetLastError(), "plain partition launch");
    }
    launch_count_scan<CNW, CR, NBITS>(ctx, in, n, dis, m, user_counts, s, counts, prefix, totals,
                                      T, ld);
    constexpr auto kern = k_partition<NW, R, 1, NBITS, kPrefetch, MDISCR_PLAIN_WB,
                                      MDISCR_PLAIN_MINW, Discr>;
    static const size_t pad = lds_pad_for(kern, WGCU);
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL(kern, dim3(T), dim3(NW * kWave), pad, s, in, out, n, dis, prefix, totals,
                       ld, m, nullptr, offsets);
    timing_end(ctx, s, tl, "partition");
    return ctx_hip(ctx, hipGetLastError(), "plain partition launch");
}

template <int NBITS, class Discr>
mdiscr_status launch_plain(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                           Discr dis, uint32_t m, uint32_t* user_counts, uint64_t* offsets,
                           hipStream_t s, uint32_t* ws) {
    if constexpr (use_lsd<NBITS, Discr>()) {
        if (((uintptr_t)in & 15u) == 0 && n > LsdGeom<Discr, NBITS>::TILE)
            return launch_plain_lsd<NBITS>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
    }
    if constexpr (NBITS <= 4 && MDISCR_PLAIN_QUAD != 0) {
        if (((uintptr_t)in & 15u) == 0)
            return launch_plain_quad<NBITS>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
    }
    if constexpr (NBITS == 6) {
        if (m >= MDISCR_BIG_MIN_M)
            return launch_plain_ballot<MDISCR_BIG_NW, MDISCR_BIG_R, MDISCR_BIG_COUNT_NW,
                                       MDISCR_BIG_WG_PER_CU, NBITS>(ctx, in, out, n, dis, m,
                                                                    user_counts, offsets, s, ws);
    }
    return launch_plain_ballot<kNW, kR, kNW, 0, NBITS>(ctx, in, out, n, dis, m, user_counts,
                                                       offsets, s, ws);
}

template <class Discr>
mdiscr_status dispatch_plain(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                             Discr dis, uint32_t m, uint32_t* user_counts, uint64_t* offsets,
                             hipStream_t s, uint32_t* ws) {
    switch (class_bits(m)) {
        case 0: return launch_plain<0>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
        case 1: return launch_plain<1>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
        case 2: return launch_plain<2>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
        case 3: return launch_plain<3>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
        case 4: return launch_plain<4>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
        case 5: return launch_plain<5>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
        case 6: return launch_plain<6>(ctx, in, out, n, dis, m, user_counts, offsets, s, ws);
    }
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "m=%u", m);
}

// Argument, device and stream checks of one plain-path call, and its n = 0
// case (zero counts and offsets).  *done is set when nothing is left to launch.
mdiscr_status plain_check(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                          uint32_t m, uint32_t* d_counts, uint64_t* d_offsets, hipStream_t s,
                          bool* done) {
    *done = true;
    if (n > MDISCR_MAX_ELEMS) return ctx_fail(ctx, MDISCR_EINVAL, "n=%llu too large", (unsigned long long)n);
    if (n && (!d_in || !d_out)) return ctx_fail(ctx, MDISCR_EINVAL, "null data pointer");
    if (n && ranges_overlap(d_in, d_out, n * 4))
        return ctx_fail(ctx, MDISCR_EINVAL, "output overlaps input");
    mdiscr_status st = ctx_begin(ctx, s);
    if (st) return st;
    if (n == 0) {
        if (d_counts)
            hipLaunchKernelGGL(k_fill_u32<>, dim3(1), dim3(64), 0, s, d_counts, 0u, m);
        if (d_offsets)  // m+1 zero 64-bit words
            hipLaunchKernelGGL(k_fill_u32<>, dim3(1), dim3(64), 0, s,
                               reinterpret_cast<uint32_t*>(d_offsets), 0u, 2 * (m + 1));
        return ctx_hip(ctx, hipGetLastError(), "fill launch");
    }
    *done = false;
    return MDISCR_OK;
}

// Workspace bytes of a plain-path call: per-tile class counts and their
// prefix (class-major rows of ld entries), then the m class totals.
size_t plain_ws_bytes(uint64_t n, uint32_t m) {
    const uint64_t T = (n + kTile - 1) / kTile;
    const uint64_t ld = (T + 3) & ~3ull;
    return (2 * ld + 1) * m * sizeof(uint32_t);
}

// The launches of one checked plain-path call with its workspace.
mdiscr_status plain_run(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                        mdiscr_discr d, uint32_t* d_counts, uint64_t* d_offsets, hipStream_t s,
                        uint32_t* ws) {
    if (d.kind == MDISCR_DISCR_MASK || (d.m & (d.m - 1)) == 0)
        return dispatch_plain(ctx, d_in, d_out, (uint32_t)n, DiscrMask{d.m - 1}, d.m, d_counts,
                              d_offsets, s, ws);
    return dispatch_plain(ctx, d_in, d_out, (uint32_t)n, DiscrMod{mdiscr_divmagic_make(d.m)}, d.m,
                          d_counts, d_offsets, s, ws);
}

// Checks, workspace and launches of a stream-ordered plain call.
mdiscr_status plain_call(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint64_t n,
                         mdiscr_discr d, uint32_t* d_counts, uint64_t* d_offsets, hipStream_t s) {
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    bool done = true;
    st = plain_check(ctx, d_in, d_out, n, d.m, d_counts, d_offsets, s, &done);
    if (st || done) return st;
    void* w = nullptr;
    st = ctx_workspace(ctx, plain_ws_bytes(n, d.m), &w);
    if (st) return st;
    return plain_run(ctx, d_in, d_out, n, d, d_counts, d_offsets, s, static_cast<uint32_t*>(w));
}

}  // namespace

bool pred_window(mdiscr_pred p, uint32_t w[4]) {
    // {mask, lo, hi, inv}; "never" is the always-true window inverted.
    uint32_t mask = ~0u, lo = 0, hi = 0, inv = 0;
    switch (p.op) {
        case MDISCR_PRED_LT:  // x < a  <=>  x <= a-1 (never for a = 0)
        case MDISCR_PRED_GE:  // not (x < a)
            if (p.a == 0) { mask = 0; hi = ~0u; inv = 1; }
            else hi = p.a - 1;
            if (p.op == MDISCR_PRED_GE) inv ^= 1u;
            break;
        case MDISCR_PRED_LE:
        case MDISCR_PRED_GT:  // not (x <= a)
            hi = p.a;
            inv = p.op == MDISCR_PRED_GT;
            break;
        case MDISCR_PRED_EQ:
        case MDISCR_PRED_NE:
            lo = p.a;
            inv = p.op == MDISCR_PRED_NE;
            break;
        case MDISCR_PRED_BITS:  // (x & a) - b == 0
            mask = p.a;
            lo = p.b;
            break;
        case MDISCR_PRED_RANGE:  // x - a <= b - a - 1 (never for b <= a)
            if (p.b <= p.a) { mask = 0; hi = ~0u; inv = 1; }
            else { lo = p.a; hi = p.b - p.a - 1; }
            break;
        default:
            return false;
    }
    w[0] = mask;
    w[1] = lo;
    w[2] = hi;
    w[3] = inv ^ 1u;
    return true;
}

mdiscr_status check_discr(mdiscr_ctx* ctx, mdiscr_discr d) {
    if (d.m == 0 || d.m > MDISCR_MAX_CLASSES)
        return ctx_fail(ctx, MDISCR_EINVAL, "m=%u outside [1,%u]", d.m, MDISCR_MAX_CLASSES);
    if (d.kind == MDISCR_DISCR_MASK) {
        if (d.m & (d.m - 1))
            return ctx_fail(ctx, MDISCR_EINVAL, "MASK discriminator needs a power-of-two m");
        return MDISCR_OK;
    }
    if (d.kind != MDISCR_DISCR_MOD)
        return ctx_fail(ctx, MDISCR_EINVAL, "unknown discriminator kind %u", d.kind);
    return MDISCR_OK;
}

uint32_t class_bits(uint32_t m) {
    uint32_t b = 0;
    while ((1u << b) < m) ++b;
    return b;
}

extern "C" {

mdiscr_status mdiscr_partition_u32(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out,
                                   uint64_t n, mdiscr_discr d, uint32_t* d_counts,
                                   void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    return plain_call(ctx, d_in, d_out, n, d, d_counts, nullptr, (hipStream_t)stream);
}

mdiscr_status mdiscr_partition_u32_ex(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out,
                                      uint64_t n, mdiscr_discr d, uint32_t* counts,
                                      uint64_t* offsets, uint32_t flags, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    if (flags & ~(uint32_t)MDISCR_OUT_HOST)
        return ctx_fail(ctx, MDISCR_EINVAL, "unknown flags 0x%x", flags);
    hipStream_t s = (hipStream_t)stream;
    if (!(flags & MDISCR_OUT_HOST)) return plain_call(ctx, d_in, d_out, n, d, counts, offsets, s);
    // Host destinations: the kernels write device staging in the context's
    // small buffer (counts at byte 0, offsets at byte 512), copied back after.
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    st = ctx_begin(ctx, s);
    if (st) return st;
    char* stage = nullptr;
    st = ctx_small_buffer(ctx, 512 + (MDISCR_MAX_CLASSES + 1) * sizeof(uint64_t), (void**)&stage);
    if (st) return st;
    uint32_t* dc = counts ? reinterpret_cast<uint32_t*>(stage) : nullptr;
    uint64_t* dof = offsets ? reinterpret_cast<uint64_t*>(stage + 512) : nullptr;
    st = plain_call(ctx, d_in, d_out, n, d, dc, dof, s);
    if (st) return st;
    if (counts)
        st = ctx_hip(ctx, hipMemcpyAsync(counts, dc, d.m * sizeof(uint32_t),
                                         hipMemcpyDeviceToHost, s), "counts readback");
    if (!st && offsets)
        st = ctx_hip(ctx, hipMemcpyAsync(offsets, dof, (d.m + 1) * sizeof(uint64_t),
                                         hipMemcpyDeviceToHost, s), "offsets readback");
    if (st) return st;
    return ctx_hip(ctx, hipStreamSynchronize(s), "hipStreamSynchronize");
}

mdiscr_status mdiscr_filter_u32(mdiscr_ctx* ctx, const uint32_t* d_in, uint32_t* d_out,
                                uint64_t n, mdiscr_pred pred, uint32_t* d_counts, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    uint32_t w[4];
    if (!pred_window(pred, w)) return ctx_fail(ctx, MDISCR_EINVAL, "unknown predicate %u", pred.op);
    hipStream_t s = (hipStream_t)stream;
    bool done = true;
    mdiscr_status st = plain_check(ctx, d_in, d_out, n, 2, d_counts, nullptr, s, &done);
    if (st || done) return st;
    void* ws = nullptr;
    st = ctx_workspace(ctx, plain_ws_bytes(n, 2), &ws);
    if (st) return st;
    return launch_plain<1>(ctx, d_in, d_out, (uint32_t)n, DiscrPred{w[0], w[1], w[2], w[3]}, 2u,
                           d_counts, nullptr, s, static_cast<uint32_t*>(ws));
}

mdiscr_status mdiscr_u32(mdiscr_ctx* ctx, uint64_t n, uint32_t num_hwd_thds,
                         const uint32_t* d_in, uint32_t* d_out, mdiscr_discr d,
                         uint32_t* h_counts) {
    (void)num_hwd_thds;
    if (!ctx || !h_counts) return ctx_fail(ctx, MDISCR_EINVAL, "null argument");
    return mdiscr_partition_u32_ex(ctx, d_in, d_out, n, d, h_counts, nullptr, MDISCR_OUT_HOST,
                                   nullptr);
}

mdiscr_status mdiscr_partition_u32_noalloc(mdiscr_ctx* ctx, const uint32_t* d_in,
                                           uint32_t* d_out, uint64_t n, mdiscr_discr d,
                                           uint32_t* d_counts, uint64_t* d_offsets,
                                           void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    if (n <= MDISCR_MAX_ELEMS && plain_ws_bytes(n, d.m) > MDISCR_NOALLOC_SCRATCH_BYTES)
        return ctx_fail(ctx, MDISCR_ENOMEM,
                        "n=%llu with m=%u needs %zu bytes of temporaries; the no-allocation "
                        "scratch holds %u",
                        (unsigned long long)n, d.m, plain_ws_bytes(n, d.m),
                        (unsigned)MDISCR_NOALLOC_SCRATCH_BYTES);
    hipStream_t s = (hipStream_t)stream;
    bool done = true;
    st = plain_check(ctx, d_in, d_out, n, d.m, d_counts, d_offsets, s, &done);
    if (st || done) return st;
    // No timing records: they would create events.
    const bool timing = ctx->timing;
    ctx->timing = false;
    st = plain_run(ctx, d_in, d_out, n, d, d_counts, d_offsets, s,
                   static_cast<uint32_t*>(ctx->scratch));
    ctx->timing = timing;
    return st;
}

mdiscr_status mdiscr_sizes_from_counts(mdiscr_ctx* ctx, const uint32_t* d_counts, uint32_t m,
                                       uint64_t n, uint32_t* d_sizes, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    if (!d_counts || m == 0 || m > MDISCR_MAX_CLASSES || (n && !d_sizes) || n > MDISCR_MAX_ELEMS)
        return ctx_fail(ctx, MDISCR_EINVAL, "bad sizes_from_counts arguments");
    hipStream_t s = (hipStream_t)stream;
    mdiscr_status st = ctx_begin(ctx, s);
    if (st || n == 0) return st;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192);
    TimedLaunch tl;
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL(k_sizes_from_counts<>, dim3(blocks), dim3(256), 0, s, d_counts, m, d_sizes,
                       (uint32_t)n);
    timing_end(ctx, s, tl, "sizes_from_counts");
    return ctx_hip(ctx, hipGetLastError(), "sizes launch");
}

}  // extern "C"

mdiscr_status ctx_small_buffer(mdiscr_ctx* ctx, size_t bytes, void** out) {
    constexpr size_t kSmall = 64 * 1024;
    if (bytes > kSmall) return ctx_fail(ctx, MDISCR_EINVAL, "small buffer request too large");
    if (!ctx->small) {
        hipError_t e = hipMalloc(&ctx->small, kSmall);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return ctx_fail(ctx, MDISCR_ENOMEM, "small buffer: %s", hipGetErrorString(e));
        }
    }
    *out = ctx->small;
    return MDISCR_OK;
}
