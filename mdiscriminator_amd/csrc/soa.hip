// soa.hip -- C ABI of the struct-of-arrays partition (mdiscr_partition_soa).
//
// Reference: mdiscr<DISCR>(n, num_hwd_thds, d_in1, d_in2, d_in3, d_out1,
// d_out2, d_out3) (MdiscrHost_tuples.cu.h:262-429), generalised to 1..8
// columns of 1/2/4/8-byte elements with any column as the 32-bit key.
#include <hip/hip_runtime.h>

#include "api_internal.hpp"
#include "soa_kernels.hpp"

using namespace mdiscr;

namespace {

constexpr int kNW = 8;
constexpr int kR = 8;
constexpr int kCT = 1;    // tiles per partition workgroup
constexpr int kCCT = 8;   // tiles per count workgroup
using SoaGeom = Geom<kNW, kR, kCT>;
constexpr uint32_t kTile = SoaGeom::kTile;

template <int NBITS, class Discr>
mdiscr_status launch_soa(mdiscr_ctx* ctx, const SoaCols& cols, uint32_t n, Discr dis,
                         uint32_t m, uint32_t* user_counts, hipStream_t s, uint32_t* ws) {
    const uint32_t T = (uint32_t)(((uint64_t)n + kTile - 1) / kTile);
    const uint32_t ld = (T + 3u) & ~3u;
    const uint32_t Kc = (T + kCCT - 1) / kCCT;
    const uint32_t Kp = (T + kCT - 1) / kCT;
    uint32_t* counts = ws;
    uint32_t* totals = ws + (size_t)m * ld;
    const uint32_t* key = static_cast<const uint32_t*>(cols.in[cols.key]);
    const bool vec = ((uintptr_t)key & 15u) == 0;
    TimedLaunch tl;
    timing_begin(ctx, s, tl);
    if (vec)
        hipLaunchKernelGGL((k_chunk_count<kNW, kR, kCCT, NBITS, true, Discr>), dim3(Kc),
                           dim3(kNW * kWave), 0, s, key, n, dis, counts, ld, m);
    else
        hipLaunchKernelGGL((k_chunk_count<kNW, kR, kCCT, NBITS, false, Discr>), dim3(Kc),
                           dim3(kNW * kWave), 0, s, key, n, dis, counts, ld, m);
    timing_end(ctx, s, tl, "soa_count");
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL((k_tile_scan<1024, 64>), dim3(m), dim3(1024), 0, s, counts, T, ld, totals,
                       user_counts);
    timing_end(ctx, s, tl, "soa_scan");
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL((k_partition_soa<kNW, kR, kCT, NBITS, Discr>), dim3(Kp), dim3(kNW * kWave),
                       0, s, cols, n, dis, counts, totals, ld, m);
    timing_end(ctx, s, tl, "soa_partition");
    return ctx_hip(ctx, hipGetLastError(), "SoA partition launch");
}

template <class Discr>
mdiscr_status dispatch_soa(mdiscr_ctx* ctx, const SoaCols& cols, uint32_t n, Discr dis,
                           uint32_t m, uint32_t* user_counts, hipStream_t s, uint32_t* ws) {
    switch (class_bits(m)) {
        case 0: return launch_soa<0>(ctx, cols, n, dis, m, user_counts, s, ws);
        case 1: return launch_soa<1>(ctx, cols, n, dis, m, user_counts, s, ws);
        case 2: return launch_soa<2>(ctx, cols, n, dis, m, user_counts, s, ws);
        case 3: return launch_soa<3>(ctx, cols, n, dis, m, user_counts, s, ws);
        case 4: return launch_soa<4>(ctx, cols, n, dis, m, user_counts, s, ws);
        case 5: return launch_soa<5>(ctx, cols, n, dis, m, user_counts, s, ws);
        case 6: return launch_soa<6>(ctx, cols, n, dis, m, user_counts, s, ws);
    }
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "m=%u", m);
}

}  // namespace

extern "C" mdiscr_status mdiscr_partition_soa(mdiscr_ctx* ctx, uint32_t ncols,
                                              const void* const* d_cols,
                                              const uint32_t* col_bytes, uint32_t key_col,
                                              void* const* d_out_cols, uint64_t n,
                                              mdiscr_discr d, uint32_t* d_counts,
                                              void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    if (ncols == 0 || ncols > (uint32_t)kMaxCols || !d_cols || !col_bytes || !d_out_cols)
        return ctx_fail(ctx, MDISCR_EINVAL, "need 1..%d columns", kMaxCols);
    if (key_col >= ncols || col_bytes[key_col] != 4)
        return ctx_fail(ctx, MDISCR_EINVAL, "key column must be a 32-bit column");
    if (n > MDISCR_MAX_ELEMS) return ctx_fail(ctx, MDISCR_EINVAL, "n too large");
    SoaCols cols = {};
    cols.ncols = ncols;
    cols.key = key_col;
    for (uint32_t j = 0; j < ncols; ++j) {
        const uint32_t b = col_bytes[j];
        if (b != 1 && b != 2 && b != 4 && b != 8)
            return ctx_fail(ctx, MDISCR_EINVAL, "column %u: element size %u", j, b);
        if (n && (!d_cols[j] || !d_out_cols[j]))
            return ctx_fail(ctx, MDISCR_EINVAL, "column %u: null pointer", j);
        if (((uintptr_t)d_cols[j] | (uintptr_t)d_out_cols[j]) % b)
            return ctx_fail(ctx, MDISCR_EINVAL, "column %u: misaligned pointer", j);
        cols.in[j] = d_cols[j];
        cols.out[j] = d_out_cols[j];
        cols.bytes[j] = b;
    }
    for (uint32_t j = 0; j < ncols; ++j) {
        const uint64_t ej = n * cols.bytes[j];
        for (uint32_t i = 0; i < ncols; ++i) {
            const uint64_t ei = n * cols.bytes[i];
            if (extents_overlap(cols.out[j], ej, cols.in[i], ei) ||
                (i != j && extents_overlap(cols.out[j], ej, cols.out[i], ei)))
                return ctx_fail(ctx, MDISCR_EINVAL, "output column %u overlaps a column", j);
        }
    }
    hipStream_t s = (hipStream_t)stream;
    st = ctx_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    if (st) return st;
    if (n == 0) {
        if (d_counts) {
            hipLaunchKernelGGL(k_fill_u32<>, dim3(1), dim3(64), 0, s, d_counts, 0u, d.m);
            return ctx_hip(ctx, hipGetLastError(), "fill launch");
        }
        return MDISCR_OK;
    }
    const uint64_t T = (n + kTile - 1) / kTile;
    const uint64_t ld = (T + 3) & ~3ull;
    void* ws = nullptr;
    st = ctx_workspace(ctx, (ld + 1) * d.m * sizeof(uint32_t), &ws);
    if (st) return st;
    if (d.kind == MDISCR_DISCR_MASK || (d.m & (d.m - 1)) == 0)
        return dispatch_soa(ctx, cols, (uint32_t)n, DiscrMask{d.m - 1}, d.m, d_counts, s,
                            (uint32_t*)ws);
    return dispatch_soa(ctx, cols, (uint32_t)n, DiscrMod{mdiscr_divmagic_make(d.m)}, d.m,
                        d_counts, s, (uint32_t*)ws);
}
