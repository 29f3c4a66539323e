// soa.hip -- C ABI of the struct-of-arrays partition (mdiscr_partition_soa).
//
// Reference: mdiscr<DISCR>(n, num_hwd_thds, d_in1, d_in2, d_in3, d_out1,
// d_out2, d_out3) (MdiscrHost_tuples.cu.h:262-429), generalised to 1..8
// columns of 1/2/4/8-byte elements with any column as the 32-bit key.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "api_internal.hpp"
#include "soa_kernels.hpp"
#include "soa_quad_kernels.hpp"

using namespace mdiscr;

namespace {

// Tile geometry by staging width and class bits (measured on 2^26
// three-column records, interleaved A/B, with write-back column stores).
// 8-byte staging (a column wider than 4 bytes): 8 waves x 8 rounds = 4096
// elements.  4-byte staging: 8 waves x 8 rounds = 4096 elements up to 32
// classes (against 12 rounds: m=8 0.297 -> 0.292 ms, m=16 0.295 -> 0.289,
// m=32 0.306 -> 0.301); 8 waves x 12 rounds = 6144 elements for 64 classes
// (0.320 ms; 16 waves x 12: 0.406, 16 x 8: 0.417, 8 x 8: 0.340, 4 x 12:
// 0.330).  The 16-wave tile of 12288 won only while the stores were
// non-temporal, when the partial lines at the ends of shorter class runs cost
// HBM bytes; with write-back stores the L2 joins them, and two workgroups
// per CU overlap each other's load, stage and write phases.
constexpr int kCT = 1;    // tiles per partition workgroup

#ifndef MDISCR_SOA_R
#define MDISCR_SOA_R 8
#endif
#ifndef MDISCR_SOA64_R
#define MDISCR_SOA64_R 12
#endif
#ifndef MDISCR_SOA64_NW
#define MDISCR_SOA64_NW 8
#endif
#ifndef MDISCR_SOA_NW
#define MDISCR_SOA_NW 8
#endif
#ifndef MDISCR_SOA_WG_PER_CU
#define MDISCR_SOA_WG_PER_CU 0
#endif
#ifndef MDISCR_SOA64_WG_PER_CU
#define MDISCR_SOA64_WG_PER_CU 0
#endif
// The select class source (three gathered columns, typically 4 + 1 + 1 bytes)
// at 8..64 classes: 4 waves x 12 rounds = 3072 elements (m=64: 0.321 ms with
// the 6144-element tile of 4-byte keys, 0.268 ms with this one).  Measured on 2^26
// (uint, uchar, bool) records, m=8, interleaved A/B: count 0.0658 -> 0.0598 ms,
// partition 0.2862 -> 0.2669 ms against 8 waves (4 x 16: 0.2918 ms).  The same
// change slows 4-byte key columns (m=16 partition 0.323 -> 0.347 ms).  With the
// destinations looked up in LDS, whole select calls: 4 x 12 0.320 ms, 4 x 8
// 0.329, 4 x 16 0.330, 8 x 8 0.337, 2 x 16 0.341.
#ifndef MDISCR_SEL_NW
#define MDISCR_SEL_NW 4
#endif
// The select geometry also below 8 classes (else the SoA one: 8 x 8).
#ifndef MDISCR_SEL_ALL_M
#define MDISCR_SEL_ALL_M 1
#endif
#ifndef MDISCR_SEL_R
#define MDISCR_SEL_R 12
#endif
template <typename ST, int NBITS, bool SEL = false>
struct SoaGeom {
    static constexpr bool kLarge = sizeof(ST) == 4 && NBITS >= 6;
    static constexpr bool kMid = sizeof(ST) == 4 && NBITS >= 3 && NBITS < 6;
    static constexpr bool kSel = SEL && sizeof(ST) == 4 && (MDISCR_SEL_ALL_M || NBITS >= 3);
    static constexpr int NW =
        kSel ? MDISCR_SEL_NW : (kLarge ? MDISCR_SOA64_NW : MDISCR_SOA_NW);
    static constexpr int R =
        kSel ? MDISCR_SEL_R : (kLarge ? MDISCR_SOA64_R : (kMid ? MDISCR_SOA_R : 8));
    static constexpr uint32_t TILE = NW * kWave * R;
    // Tiles per count workgroup (2^26 keys, interleaved A/B, count ms for 1 /
    // 2 / 8 tiles: m=16 0.042 / 0.045 / 0.051, select m=8 0.066 / 0.069 /
    // 0.079, m=64 0.050 / 0.046 / 0.048).
    static constexpr int CCT = kLarge ? 2 : 1;
    // Partition workgroups resident per CU (0: as many as fit).
    static constexpr int WGCU = kLarge ? MDISCR_SOA64_WG_PER_CU : MDISCR_SOA_WG_PER_CU;
};

// The tile launch_soa uses for these columns, classes and class source (the
// workspace size).
uint32_t soa_tile(bool wide, uint32_t nbits, bool sel) {
    if (wide) return SoaGeom<uint64_t, 0>::TILE;
    if (nbits >= 6) return sel ? SoaGeom<uint32_t, 6, true>::TILE : SoaGeom<uint32_t, 6>::TILE;
    if (nbits >= 3) return sel ? SoaGeom<uint32_t, 3, true>::TILE : SoaGeom<uint32_t, 3>::TILE;
    return sel ? SoaGeom<uint32_t, 0, true>::TILE : SoaGeom<uint32_t, 0>::TILE;
}
static_assert(SoaGeom<uint32_t, 5>::TILE == SoaGeom<uint32_t, 3>::TILE &&
                  SoaGeom<uint32_t, 5, true>::TILE == SoaGeom<uint32_t, 3, true>::TILE &&
                  SoaGeom<uint32_t, 2>::TILE == SoaGeom<uint32_t, 0>::TILE &&
                  SoaGeom<uint32_t, 2, true>::TILE == SoaGeom<uint32_t, 0, true>::TILE &&
                  SoaGeom<uint64_t, 6>::TILE == SoaGeom<uint64_t, 0>::TILE &&
                  SoaGeom<uint64_t, 3, true>::TILE == SoaGeom<uint64_t, 0>::TILE,
              "soa_tile covers every geometry");

template <int NBITS, typename ST, class Src, class Discr>
mdiscr_status launch_soa(mdiscr_ctx* ctx, const SoaCols& cols, Src src, uint32_t n, Discr dis,
                         uint32_t m, uint32_t* user_counts, hipStream_t s, uint32_t* ws) {
    using G = SoaGeom<ST, NBITS, IsSelect<Src>::value>;
    constexpr int NW = G::NW, R = G::R;
    constexpr int kCCT = G::CCT;
    constexpr uint32_t TILE = G::TILE;
    const uint32_t T = (uint32_t)(((uint64_t)n + TILE - 1) / TILE);
    const uint32_t ld = (T + 3u) & ~3u;
    const uint32_t Kc = (T + kCCT - 1) / kCCT;
    const uint32_t Kp = (T + kCT - 1) / kCT;
    uint32_t* counts = ws;
    uint32_t* prefix = ws + (size_t)m * ld;
    uint32_t* totals = ws + (size_t)2 * m * ld;
    TimedLaunch tl;
    if (T == 1) {
        // One tile: the partition counts it itself (one launch instead of three).
        timing_begin(ctx, s, tl);
        hipLaunchKernelGGL((k_partition_soa<NW, R, kCT, NBITS, Src, ST, Discr, true>), dim3(1),
                           dim3(NW * kWave), 0, s, cols, src, n, dis, nullptr, nullptr, 0u, m,
                           user_counts);
        timing_end(ctx, s, tl, "soa_partition");
        return ctx_hip(ctx, hipGetLastError(), "SoA partition launch");
    }
    timing_begin(ctx, s, tl);
    bool vec = false;
    if constexpr (std::is_same_v<Src, KeySrc<uint32_t>>)
        vec = src.stride == 1 && ((uintptr_t)src.key & 15u) == 0;
    if (vec) {
        if constexpr (std::is_same_v<Src, KeySrc<uint32_t>>)
            hipLaunchKernelGGL((k_chunk_count<NW, R, kCCT, NBITS, true, Discr>), dim3(Kc),
                               dim3(NW * kWave), 0, s, src.key, n, dis, counts, ld, m);
    } else {
        hipLaunchKernelGGL((k_src_count<NW, R, kCCT, NBITS, Src, Discr>), dim3(Kc),
                           dim3(NW * kWave), 0, s, src, n, dis, counts, ld, m);
    }
    timing_end(ctx, s, tl, "soa_count");
    timing_begin(ctx, s, tl);
    launch_tile_scan(counts, prefix, T, ld, m, totals, user_counts, nullptr, s);
    timing_end(ctx, s, tl, "soa_scan");
    constexpr auto kern = k_partition_soa<NW, R, kCT, NBITS, Src, ST, Discr>;
    static const size_t pad = lds_pad_for(kern, G::WGCU);
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL(kern, dim3(Kp), dim3(NW * kWave), pad, s, cols, src, n, dis, prefix, totals,
                       ld, m, nullptr);
    timing_end(ctx, s, tl, "soa_partition");
    return ctx_hip(ctx, hipGetLastError(), "SoA partition launch");
}

// The quad-layout partition (up to 16 classes, every column of 1, 2 or 4
// bytes and aligned to 4 elements): NW waves x 64 lanes x Q quads per tile.
// 2^26 records, partition ms against the ballot-ranked kernel (interleaved
// A/B): (u32, u8, u8) keyed on the u32 at m=4 0.191 -> 0.165, m=16 0.207 ->
// 0.185; select z ? x & 7 : y & 7 over (u32, u8, u8) 0.221 -> 0.149; a 1-byte
// class array with a u32 payload, m=16 0.171 -> 0.149.  Tiles of 8 x 2, 4 x 2
// and 4 x 3 quads per lane are within 3% of each other.
#ifndef MDISCR_SOA_QUAD
#define MDISCR_SOA_QUAD 1
#endif
#ifndef MDISCR_SOAQ_NW
#define MDISCR_SOAQ_NW 4
#endif
#ifndef MDISCR_SOAQ_Q
#define MDISCR_SOAQ_Q 2
#endif
#ifndef MDISCR_SOAQ_WG_PER_CU
#define MDISCR_SOAQ_WG_PER_CU 0
#endif
constexpr uint32_t kQuadTile = MDISCR_SOAQ_NW * kWave * 4 * MDISCR_SOAQ_Q;

#ifndef MDISCR_SOAQ_ALL
#define MDISCR_SOAQ_ALL 0
#endif
// Whether the quad-layout partition takes these columns: it can when every
// column holds 1-, 2- or 4-byte elements aligned to 4 elements, and it is
// faster when some column is narrower than 4 bytes or the class comes from a
// select.  For 4-byte columns alone the ballot-ranked kernel is 2-3% faster
// (2^26 (key, f32, index) records, m=8 / 16).
bool quad_columns(const SoaCols& cols, bool select) {
    bool narrow = false;
    for (uint32_t j = 0; j < cols.ncols; ++j) {
        const uint32_t b = cols.bytes[j];
        if (b > 4) return false;
        if ((((uintptr_t)cols.in[j] | (uintptr_t)cols.out[j]) % (4u * b)) != 0) return false;
        narrow |= b < 4;
    }
    return narrow || select || MDISCR_SOAQ_ALL;
}

template <int NBITS, class Src, class Discr>
mdiscr_status launch_soa_quad(mdiscr_ctx* ctx, const SoaCols& cols, Src src, uint32_t n,
                              Discr dis, uint32_t m, uint32_t* user_counts, hipStream_t s,
                              uint32_t* ws) {
    constexpr int NW = MDISCR_SOAQ_NW, Q = MDISCR_SOAQ_Q;
    constexpr uint32_t TILE = kQuadTile;
    const uint32_t T = (uint32_t)(((uint64_t)n + TILE - 1) / TILE);
    const uint32_t ld = (T + 3u) & ~3u;
    uint32_t* counts = ws;
    uint32_t* prefix = ws + (size_t)m * ld;
    uint32_t* totals = ws + (size_t)2 * m * ld;
    TimedLaunch tl;
    if (T == 1) {
        timing_begin(ctx, s, tl);
        hipLaunchKernelGGL((k_partition_soa_quad<NW, Q, NBITS, Src, Discr, true>), dim3(1),
                           dim3(NW * kWave), 0, s, cols, src, n, dis, nullptr, nullptr, 0u, m,
                           user_counts);
        timing_end(ctx, s, tl, "soa_partition");
        return ctx_hip(ctx, hipGetLastError(), "SoA partition launch");
    }
    timing_begin(ctx, s, tl);
    if constexpr (std::is_same_v<Src, KeySrc<uint32_t>>) {
        // 16-byte aligned keys (quad_columns): 16-byte count loads.
        hipLaunchKernelGGL((k_chunk_count<NW, 4 * Q, 1, NBITS, true, Discr>), dim3(T),
                           dim3(NW * kWave), 0, s, src.key, n, dis, counts, ld, m);
    } else {
        hipLaunchKernelGGL((k_src_count<NW, 4 * Q, 1, NBITS, Src, Discr>), dim3(T),
                           dim3(NW * kWave), 0, s, src, n, dis, counts, ld, m);
    }
    timing_end(ctx, s, tl, "soa_count");
    timing_begin(ctx, s, tl);
    launch_tile_scan(counts, prefix, T, ld, m, totals, user_counts, nullptr, s);
    timing_end(ctx, s, tl, "soa_scan");
    constexpr auto kern = k_partition_soa_quad<NW, Q, NBITS, Src, Discr>;
    static const size_t pad = lds_pad_for(kern, MDISCR_SOAQ_WG_PER_CU);
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL(kern, dim3(T), dim3(NW * kWave), pad, s, cols, src, n, dis, prefix, totals,
                       ld, m, nullptr);
    timing_end(ctx, s, tl, "soa_partition");
    return ctx_hip(ctx, hipGetLastError(), "SoA partition launch");
}

template <typename ST, class Src, class Discr>
mdiscr_status dispatch_bits(mdiscr_ctx* ctx, const SoaCols& cols, Src src, uint32_t n, Discr dis,
                            uint32_t m, uint32_t* user_counts, hipStream_t s, uint32_t* ws) {
    if constexpr (sizeof(ST) == 4 && MDISCR_SOA_QUAD) {
        if (quad_columns(cols, IsSelect<Src>::value)) {
            switch (class_bits(m)) {
                case 0: return launch_soa_quad<0>(ctx, cols, src, n, dis, m, user_counts, s, ws);
                case 1: return launch_soa_quad<1>(ctx, cols, src, n, dis, m, user_counts, s, ws);
                case 2: return launch_soa_quad<2>(ctx, cols, src, n, dis, m, user_counts, s, ws);
                case 3: return launch_soa_quad<3>(ctx, cols, src, n, dis, m, user_counts, s, ws);
                case 4: return launch_soa_quad<4>(ctx, cols, src, n, dis, m, user_counts, s, ws);
                default: break;
            }
        }
    }
    switch (class_bits(m)) {
        case 0: return launch_soa<0, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
        case 1: return launch_soa<1, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
        case 2: return launch_soa<2, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
        case 3: return launch_soa<3, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
        case 4: return launch_soa<4, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
        case 5: return launch_soa<5, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
        case 6: return launch_soa<6, ST>(ctx, cols, src, n, dis, m, user_counts, s, ws);
    }
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "m=%u", m);
}

template <typename ST, class Src>
mdiscr_status dispatch_discr(mdiscr_ctx* ctx, const SoaCols& cols, Src src, uint32_t n,
                             mdiscr_discr d, uint32_t* user_counts, hipStream_t s, uint32_t* ws) {
    if (d.kind == MDISCR_DISCR_MASK || (d.m & (d.m - 1)) == 0)
        return dispatch_bits<ST>(ctx, cols, src, n, DiscrMask{d.m - 1}, d.m, user_counts, s, ws);
    return dispatch_bits<ST>(ctx, cols, src, n, DiscrMod{mdiscr_divmagic_make(d.m)}, d.m,
                             user_counts, s, ws);
}

// What feeds the discriminator of one SoA call.
enum class ClassKind { Key, Pred, Select };
struct ClassSpec {
    ClassKind kind = ClassKind::Key;
    mdiscr_discr d{};                // Key, Select
    uint32_t pw[4] = {0, 0, 0, 0};   // Pred: the normalised window (m = 2)
    uint32_t sel = 0, a = 0, b = 0;  // Select: column indices
    uint32_t m() const { return kind == ClassKind::Pred ? 2u : d.m; }
};

template <typename ST>
mdiscr_status run_staged(mdiscr_ctx* ctx, const SoaCols& cols, uint32_t n, const ClassSpec& cs,
                         uint32_t* d_counts, hipStream_t s, uint32_t* w) {
    switch (cs.kind) {
        case ClassKind::Key:
            if (cols.bytes[cols.key] == 1)
                return dispatch_discr<ST>(
                    ctx, cols, KeySrc<uint8_t>{static_cast<const uint8_t*>(cols.in[cols.key]), 1u},
                    n, cs.d, d_counts, s, w);
            return dispatch_discr<ST>(
                ctx, cols, KeySrc<uint32_t>{static_cast<const uint32_t*>(cols.in[cols.key]), 1u},
                n, cs.d, d_counts, s, w);
        case ClassKind::Pred: {
            const KeySrc<uint32_t> ks{static_cast<const uint32_t*>(cols.in[cols.key]), 1u};
            const DiscrPred dp{cs.pw[0], cs.pw[1], cs.pw[2], cs.pw[3]};
            if constexpr (sizeof(ST) == 4 && MDISCR_SOA_QUAD) {
                if (quad_columns(cols, false))
                    return launch_soa_quad<1>(ctx, cols, ks, n, dp, 2u, d_counts, s, w);
            }
            return launch_soa<1, ST>(ctx, cols, ks, n, dp, 2u, d_counts, s, w);
        }
        case ClassKind::Select:
            return dispatch_discr<ST>(
                ctx, cols,
                SelectSrc{cols.in[cs.sel], cols.in[cs.a], cols.in[cs.b], cols.bytes[cs.sel] == 4u,
                          cols.bytes[cs.a] == 4u, cols.bytes[cs.b] == 4u, cs.sel, cs.a, cs.b},
                n, cs.d, d_counts, s, w);
    }
    return ctx_fail(ctx, MDISCR_EINVAL, "class kind");
}

// Validated columns -> launch (shared by the SoA, array, filter and select
// entry points).
mdiscr_status run_soa(mdiscr_ctx* ctx, SoaCols& cols, uint64_t n, const ClassSpec& cs,
                      uint32_t* d_counts, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    const uint32_t m = cs.m();
    mdiscr_status st = ctx_begin(ctx, s);
    if (st) return st;
    if (n == 0) {
        if (d_counts) {
            hipLaunchKernelGGL(k_fill_u32<>, dim3(1), dim3(64), 0, s, d_counts, 0u, m);
            return ctx_hip(ctx, hipGetLastError(), "fill launch");
        }
        return MDISCR_OK;
    }
    bool wide = false;
    for (uint32_t j = 0; j < cols.ncols; ++j) wide |= cols.bytes[j] > 4;
    // sized by the smallest tile the call may use
    const uint64_t tile = std::min<uint64_t>(
        soa_tile(wide, class_bits(m), cs.kind == ClassKind::Select), kQuadTile);
    const uint64_t T = (n + tile - 1) / tile;
    const uint64_t ld = (T + 3) & ~3ull;
    void* ws = nullptr;
    st = ctx_workspace(ctx, (2 * ld + 1) * m * sizeof(uint32_t), &ws);  // counts, prefix, totals
    if (st) return st;
    uint32_t* w = static_cast<uint32_t*>(ws);
    return wide ? run_staged<uint64_t>(ctx, cols, (uint32_t)n, cs, d_counts, s, w)
                : run_staged<uint32_t>(ctx, cols, (uint32_t)n, cs, d_counts, s, w);
}

// Input and output columns must not overlap, nor output columns each other.
mdiscr_status check_overlaps(mdiscr_ctx* ctx, const SoaCols& cols, uint64_t n) {
    for (uint32_t j = 0; j < cols.ncols; ++j) {
        const uint64_t ej = n * cols.bytes[j];
        for (uint32_t i = 0; i < cols.ncols; ++i) {
            const uint64_t ei = n * cols.bytes[i];
            if (extents_overlap(cols.out[j], ej, cols.in[i], ei))
                return ctx_fail(ctx, MDISCR_EINVAL, "output column %u overlaps an input", j);
            if (i != j && extents_overlap(cols.out[j], ej, cols.out[i], ei))
                return ctx_fail(ctx, MDISCR_EINVAL, "output columns %u and %u overlap", j, i);
        }
    }
    return MDISCR_OK;
}

}  // namespace

namespace {

// Contiguous caller columns -> SoaCols, with the checks every SoA entry point
// shares (count, element sizes, null and misaligned pointers, overlaps).
mdiscr_status build_cols(mdiscr_ctx* ctx, uint32_t ncols, const void* const* d_cols,
                         const uint32_t* col_bytes, void* const* d_out_cols, uint64_t n,
                         SoaCols& cols) {
    if (ncols == 0 || ncols > (uint32_t)kMaxCols || !d_cols || !col_bytes || !d_out_cols)
        return ctx_fail(ctx, MDISCR_EINVAL, "need 1..%d columns", kMaxCols);
    if (n > MDISCR_MAX_ELEMS) return ctx_fail(ctx, MDISCR_EINVAL, "n too large");
    cols = SoaCols{};
    cols.ncols = ncols;
    cols.key = (uint32_t)kMaxCols;  // no column is written from registers
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
    return check_overlaps(ctx, cols, n);
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
    SoaCols cols;
    st = build_cols(ctx, ncols, d_cols, col_bytes, d_out_cols, n, cols);
    if (st) return st;
    if (key_col >= ncols || (cols.bytes[key_col] != 4 && cols.bytes[key_col] != 1))
        return ctx_fail(ctx, MDISCR_EINVAL, "key column must hold 4-byte keys or 1-byte classes");
    cols.key = key_col;
    ClassSpec cs;
    cs.d = d;
    return run_soa(ctx, cols, n, cs, d_counts, stream);
}

extern "C" mdiscr_status mdiscr_filter_soa(mdiscr_ctx* ctx, uint32_t ncols,
                                           const void* const* d_cols, const uint32_t* col_bytes,
                                           uint32_t key_col, void* const* d_out_cols, uint64_t n,
                                           mdiscr_pred pred, uint32_t* d_counts, void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    ClassSpec cs;
    cs.kind = ClassKind::Pred;
    if (!pred_window(pred, cs.pw)) return ctx_fail(ctx, MDISCR_EINVAL, "unknown predicate %u", pred.op);
    SoaCols cols;
    mdiscr_status st = build_cols(ctx, ncols, d_cols, col_bytes, d_out_cols, n, cols);
    if (st) return st;
    if (key_col >= ncols || cols.bytes[key_col] != 4)
        return ctx_fail(ctx, MDISCR_EINVAL, "the predicate reads a 4-byte key column");
    cols.key = key_col;
    return run_soa(ctx, cols, n, cs, d_counts, stream);
}

extern "C" mdiscr_status mdiscr_partition_soa_select(mdiscr_ctx* ctx, uint32_t ncols,
                                                     const void* const* d_cols,
                                                     const uint32_t* col_bytes, uint32_t sel_col,
                                                     uint32_t a_col, uint32_t b_col,
                                                     void* const* d_out_cols, uint64_t n,
                                                     mdiscr_discr d, uint32_t* d_counts,
                                                     void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    SoaCols cols;
    st = build_cols(ctx, ncols, d_cols, col_bytes, d_out_cols, n, cols);
    if (st) return st;
    const uint32_t which[3] = {sel_col, a_col, b_col};
    for (uint32_t c : which)
        if (c >= ncols || (cols.bytes[c] != 1 && cols.bytes[c] != 4))
            return ctx_fail(ctx, MDISCR_EINVAL,
                            "select column %u must exist and hold 1- or 4-byte values", c);
    ClassSpec cs;
    cs.kind = ClassKind::Select;
    cs.d = d;
    cs.sel = sel_col;
    cs.a = a_col;
    cs.b = b_col;
    return run_soa(ctx, cols, n, cs, d_counts, stream);
}
