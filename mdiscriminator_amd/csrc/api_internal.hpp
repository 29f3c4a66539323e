// api_internal.hpp -- context state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mdiscr.h"

struct TimedLaunch {
    bool on = false;
    hipEvent_t a = nullptr, b = nullptr;
    const char* name = nullptr;
};

struct mdiscr_ctx {
    int device = 0;
    int num_cu = 256;
    void* ws = nullptr;        // grow-only workspace
    size_t ws_bytes = 0;
    // Buffers replaced by a larger one.  Work queued before the growth (or a
    // graph captured with them) may still read them, so they are freed only
    // when the context is destroyed; growth at least doubles, which bounds
    // them by the live buffer's size.
    std::vector<void*> retired;
    void* small = nullptr;     // 64 KiB of device scratch for count readbacks
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::map<std::string, std::pair<double, uint64_t>> timing_acc;
    void* extra = nullptr;     // segmented-path planning buffers
    void* scratch = nullptr;   // this device's no-allocation scratch (device address)
    char err[256] = {0};
};

mdiscr_status ctx_fail(mdiscr_ctx* ctx, mdiscr_status s, const char* fmt, ...);
mdiscr_status ctx_hip(mdiscr_ctx* ctx, hipError_t e, const char* what);
mdiscr_status ctx_workspace(mdiscr_ctx* ctx, size_t bytes, void** out);
// Grow-only device buffer: a larger request allocates anew and retires the
// old buffer to ctx->retired (no device-wide wait, no free while in use).
mdiscr_status ctx_grow(mdiscr_ctx* ctx, void** buf, size_t* have, size_t need);
// Device and stream checks every launching entry point makes first: the
// context's device is made current, and a stream that is being captured into
// a graph is refused with MDISCR_EUNSUPPORTED before anything is allocated,
// synchronised or launched.
mdiscr_status ctx_begin(mdiscr_ctx* ctx, hipStream_t s);
mdiscr_status ctx_small_buffer(mdiscr_ctx* ctx, size_t bytes, void** out);
void ctx_release_extra(mdiscr_ctx* ctx);
// Sizes the batch path's workspace and pinned descriptor staging for batches
// of up to max_reqs requests and max_elems elements (any m), so that such
// batches never grow a buffer.
mdiscr_status ctx_reserve_batch(mdiscr_ctx* ctx, uint64_t max_elems, uint32_t max_reqs);
void timing_begin(mdiscr_ctx* ctx, hipStream_t s, TimedLaunch& t);
void timing_end(mdiscr_ctx* ctx, hipStream_t s, TimedLaunch& t, const char* name);
mdiscr_status check_discr(mdiscr_ctx* ctx, mdiscr_discr d);
uint32_t class_bits(uint32_t m);

// A two-way filter predicate normalised to one unsigned window test:
// holds(x) = ((x & w[0]) - w[1] <= w[2]) != inv, with w[3] = inv ^ 1 so that
// class(x) = window ^ w[3] (0 where the predicate holds).  False for an
// unknown op.
bool pred_window(mdiscr_pred p, uint32_t w[4]);

// Dynamic LDS bytes that make `kernel` (static LDS as compiled) fit at most
// `per_cu` workgroups per CU: the CU's LDS divided by per_cu, less the static part.
template <typename K>
inline size_t lds_pad_for(K kernel, int per_cu) {
    if (per_cu <= 0) return 0;
    hipFuncAttributes fa{};
    int dev = 0, lds_cu = 0;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kernel)) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
            hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    // The padded size is kept to a KiB multiple.
    const size_t target = ((size_t)lds_cu / (size_t)per_cu) & ~(size_t)1023;
    return target > fa.sharedSizeBytes ? target - fa.sharedSizeBytes : 0;
}

inline bool ranges_overlap(const void* a, const void* b, uint64_t bytes) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return x < y + bytes && y < x + bytes;
}

// [a, a+na) and [b, b+nb) share a byte.
inline bool extents_overlap(const void* a, uint64_t na, const void* b, uint64_t nb) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return na && nb && x < y + nb && y < x + na;
}
