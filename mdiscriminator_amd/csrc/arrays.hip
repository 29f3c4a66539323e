// arrays.hip -- array-valued elements: n elements of LEN consecutive uint32
// words, classified by word 0 and moved whole (mdiscr_partition_arrays).
//
// Reference: ModArray4<LEN> / ModArray<N,LEN> and mdiscr<DISCR,LEN>
// (MdiscrHost_arrays.cu.h:14-36, 250-470).  The reference transposes the
// LEN-word rows into LEN columns and back (matTransposeSpecial(Reverse),
// MdiscrKernels_arrays.cu.h:87-105) around a column-wise partition.  Here a
// tile's rows are read flat and written as whole rows, with no transposes:
//
//   count     word 0 of every element (the key column at stride LEN), per tile
//   scan      per-class exclusive scan over tiles (shared with the plain path)
//   partition per tile, rows of 2, 4 or 8 words with vector-aligned buffers
//             (k_partition_rows_vec): whole rows loaded in (round, lane) order,
//             ranked by their word 0, staged at their LDS slot and written out
//             whole.  Other rows (k_partition_rows): (1) flat coalesced loads
//             of the tile's TILE*LEN words, plus the keys (word 0, the same
//             cache lines); (2) in-wave ballot rank of the keys, each
//             element's LDS slot recorded in slot_of; (3) every loaded word is
//             staged at slot*LEN + word in LDS, so LDS holds the tile's rows
//             grouped by class; (4) LDS word w goes to
//             out[(dpos[class] + w/LEN)*LEN + w%LEN]: contiguous runs.
//
// HBM bytes per element: 4*LEN (count: the strided key reads fetch the whole
// rows once LEN*4 < the cache line) + 8*LEN (partition read + write).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "api_internal.hpp"
#include "soa_kernels.hpp"

using namespace mdiscr;

namespace {

// Output stores of the row kernels (1 = non-temporal hint).  Write-back, so
// that partial lines at class-run ends merge in L2 (2^25 rows, m=16: LEN=2
// 0.099 -> 0.089 ms, LEN=8 0.597 -> 0.469).
#ifndef MDISCR_ARR_NT_STORE
#define MDISCR_ARR_NT_STORE 0
#endif
#ifndef MDISCR_ROWS_NW
#define MDISCR_ROWS_NW 4
#endif
#ifndef MDISCR_ROWS_COUNT_CT
#define MDISCR_ROWS_COUNT_CT 8
#endif
#ifndef MDISCR_ROWS_WPT
#define MDISCR_ROWS_WPT 32
#endif
#ifndef MDISCR_ROWS_NT_LOAD
#define MDISCR_ROWS_NT_LOAD 0
#endif
// Words per thread for 2-word rows: half the general figure (2^25 rows, m=16,
// interleaved A/B with write-back stores: call 0.145 -> 0.132 ms, the count in
// the same tiles 0.054 -> 0.046; 3-, 4- and 8-word rows were 1-11% slower
// with fewer words per thread).
#ifndef MDISCR_ROWS_WPT_LEN2
#define MDISCR_ROWS_WPT_LEN2 16
#endif
constexpr int kNW = MDISCR_ROWS_NW;  // waves per workgroup
// Rounds per lane: about MDISCR_ROWS_WPT words per thread, so a tile's rows
// take NW * 64 * WPT * 4 bytes of LDS (32 KiB at the defaults).
constexpr int rounds_for(int len) {
    const int wpt = len == 2 ? MDISCR_ROWS_WPT_LEN2 : MDISCR_ROWS_WPT;
    return wpt / len < 2 ? 2 : wpt / len;
}

template <int NW, int R, int LEN, int NBITS, class Discr, bool SINGLE = false>
__global__ __launch_bounds__(NW* kWave) void k_partition_rows(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n, Discr dis,
    const uint32_t* __restrict__ prefix, const uint32_t* __restrict__ totals, uint32_t ld,
    uint32_t m, uint32_t* __restrict__ user_counts) {
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = NW * kWave * R;  // elements
    constexpr int THREADS = NW * kWave;
    constexpr int WPT = R * LEN;           // words per thread
    __shared__ uint32_t stage[TILE * LEN];
    __shared__ uint16_t slot_of[TILE];     // tile-local element -> LDS slot
    __shared__ uint8_t cls[TILE];          // LDS slot -> class
    __shared__ uint32_t cursor[NW][MC];
    __shared__ uint32_t dpos[MC];

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tbase = t * (uint32_t)TILE;
    const uint32_t wbase = tbase + wave * (uint32_t)(kWave * R);
    const bool full = (uint64_t)tbase + TILE <= n;
    const uint32_t valid = full ? (uint32_t)TILE : n - tbase;
    const uint32_t vwords = valid * (uint32_t)LEN;
    const uint32_t* __restrict__ tin = in + (size_t)tbase * LEN;

    uint32_t pre = 0, tot = 0;
    if (!SINGLE && wave == 0 && lane < m) {
        pre = prefix[(size_t)lane * ld + t];
        tot = totals[lane];
    }

    // 1. The tile's words, flat: thread-word k covers tile word k*THREADS + tid.
    uint32_t w[WPT];
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const uint32_t q = k * THREADS + threadIdx.x;
        if (full || q < vwords)
            w[k] = MDISCR_ROWS_NT_LOAD ? __builtin_nontemporal_load(tin + q) : tin[q];
        else
            w[k] = 0u;
    }
    // The keys: word 0 of elements wbase + r*64 + lane (lines fetched above).
    uint32_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t idx = wbase + r * kWave + lane;
        v[r] = (full || idx < n) ? in[(size_t)idx * LEN] : 0u;
    }
    for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
    __builtin_amdgcn_wave_barrier();

    // 2. In-wave stable rank, then tile offsets (wave 0, lane = class).
    uint32_t slot[R];
    if (full) rank_rounds<R, NBITS, true>(v, slot, cursor[wave], wbase, n, m, lane, dis);
    else rank_rounds<R, NBITS, false>(v, slot, cursor[wave], wbase, n, m, lane, dis);
    __syncthreads();
    if (wave == 0) {
        uint32_t running;
        if constexpr (SINGLE) {
            // One tile is all the rows: its class counts are the totals.
            // Rows past n were ranked as class m-1: not counted.
            uint32_t ct = 0;
            if (lane < (uint32_t)MC)
#pragma unroll
                for (int w = 0; w < NW; ++w) ct += cursor[w][lane];
            running = wave_excl_scan(ct, lane);
            if (user_counts && lane < m)
                user_counts[lane] = ct - (lane == m - 1u ? (uint32_t)TILE - n : 0u);
        } else {
            running = wave_excl_scan(tot, lane) + pre;
        }
        tile_offsets<NW, MC>(cursor, dpos, running, lane);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t c = slot[r] >> 16;
        const uint32_t s = cursor[wave][c] + (slot[r] & 0xFFFFu);
        cls[s] = (uint8_t)c;
        slot_of[wave * (kWave * R) + r * kWave + lane] = (uint16_t)s;
    }
    __syncthreads();

    // 3. Rows into LDS, grouped by class.
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const uint32_t q = k * THREADS + threadIdx.x;
        const uint32_t e = q / (uint32_t)LEN;
        if (full || q < vwords) stage[slot_of[e] * (uint32_t)LEN + (q - e * (uint32_t)LEN)] = w[k];
    }
    __syncthreads();

    // 4. Contiguous class runs of whole rows.
#pragma unroll
    for (int k = 0; k < WPT; ++k) {
        const uint32_t q = k * THREADS + threadIdx.x;
        const uint32_t s = q / (uint32_t)LEN;
        if (full || q < vwords) {
            const size_t dst = (size_t)(dpos[cls[s]] + s) * LEN + (q - s * (uint32_t)LEN);
            st_out<MDISCR_ARR_NT_STORE>(stage[q], out + dst);
        }
    }
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Rows of 2, 4 or 8 words whose input and output are aligned to the row's
// vector (8 or 16 bytes): moved as whole rows.  Round r of wave w holds rows
// wbase + r*64 + lane, loaded as one vector per lane (two for LEN 8), so each
// key (word 0) arrives with its row and needs no second load.  A row is staged
// at its slot with vector LDS accesses, and slot s goes out whole to output
// row dpos[class] + s.  Against k_partition_rows this takes one class and
// offset lookup per row instead of per word.  Same tile as the count pass.
template <int NW, int R, int LEN, int NBITS, class Discr, bool SINGLE = false>
__global__ __launch_bounds__(NW* kWave) void k_partition_rows_vec(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n, Discr dis,
    const uint32_t* __restrict__ prefix, const uint32_t* __restrict__ totals, uint32_t ld,
    uint32_t m, uint32_t* __restrict__ user_counts) {
    static_assert(LEN == 2 || LEN == 4 || LEN == 8, "rows of whole vectors");
    using V = typename std::conditional<LEN == 2, u32x2, u32x4>::type;
    constexpr int VPR = LEN == 8 ? 2 : 1;  // vectors per row
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = NW * kWave * R;   // rows
    constexpr int THREADS = NW * kWave;
    __shared__ V stage[TILE * VPR];
    __shared__ uint8_t cls[TILE];          // LDS slot -> class
    __shared__ uint32_t cursor[NW][MC];
    __shared__ uint32_t dpos[MC];

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tbase = t * (uint32_t)TILE;
    const uint32_t wbase = tbase + wave * (uint32_t)(kWave * R);
    const bool full = (uint64_t)tbase + TILE <= n;
    const uint32_t valid = full ? (uint32_t)TILE : n - tbase;

    uint32_t pre = 0, tot = 0;
    if (!SINGLE && wave == 0 && lane < m) {
        pre = prefix[(size_t)lane * ld + t];
        tot = totals[lane];
    }

    // 1. Whole rows in (round, lane) order; word 0 is the key.
    const V* __restrict__ lp = reinterpret_cast<const V*>(in) + ((size_t)wbase + lane) * VPR;
    V row[R][VPR];
    uint32_t v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool ok = full || wbase + r * kWave + lane < n;
#pragma unroll
        for (int j = 0; j < VPR; ++j)
            row[r][j] = ok ? __builtin_nontemporal_load(lp + (r * kWave) * VPR + j) : (V)(0u);
        v[r] = row[r][0].x;
    }
    for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
    __builtin_amdgcn_wave_barrier();

    // 2. In-wave stable rank, then tile offsets (wave 0, lane = class).
    uint32_t slot[R];
    if (full) rank_rounds<R, NBITS, true>(v, slot, cursor[wave], wbase, n, m, lane, dis);
    else rank_rounds<R, NBITS, false>(v, slot, cursor[wave], wbase, n, m, lane, dis);
    __syncthreads();
    if (wave == 0) {
        uint32_t running;
        if constexpr (SINGLE) {
            // One tile is all the rows: its class counts are the totals.
            // Rows past n were ranked as class m-1: not counted.
            uint32_t ct = 0;
            if (lane < (uint32_t)MC)
#pragma unroll
                for (int w = 0; w < NW; ++w) ct += cursor[w][lane];
            running = wave_excl_scan(ct, lane);
            if (user_counts && lane < m)
                user_counts[lane] = ct - (lane == m - 1u ? (uint32_t)TILE - n : 0u);
        } else {
            running = wave_excl_scan(tot, lane) + pre;
        }
        tile_offsets<NW, MC>(cursor, dpos, running, lane);
    }
    __syncthreads();

    // 3. Rows into LDS at their slots, grouped by class.
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t c = slot[r] >> 16;
        const uint32_t sl = cursor[wave][c] + (slot[r] & 0xFFFFu);
        cls[sl] = (uint8_t)c;
#pragma unroll
        for (int j = 0; j < VPR; ++j) stage[sl * VPR + j] = row[r][j];
    }
    __syncthreads();

    // 4. Slot s to output row dpos[class] + s: contiguous class runs.
    V* __restrict__ ov = reinterpret_cast<V*>(out);
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint32_t sl = k * THREADS + threadIdx.x;
        if (full || sl < valid) {
            const size_t d = (size_t)(dpos[cls[sl]] + sl) * VPR;
#pragma unroll
            for (int j = 0; j < VPR; ++j) st_out<MDISCR_ARR_NT_STORE>(stage[sl * VPR + j], ov + d + j);
        }
    }
}

template <int LEN, int NBITS, class Discr>
mdiscr_status launch_rows(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                          Discr dis, uint32_t m, uint32_t* user_counts, hipStream_t s,
                          uint32_t* ws) {
    constexpr int R = rounds_for(LEN);
    constexpr uint32_t TILE = kNW * kWave * R;
    constexpr int kCCT = MDISCR_ROWS_COUNT_CT;  // tiles per count workgroup
    const uint32_t T = (uint32_t)(((uint64_t)n + TILE - 1) / TILE);
    const uint32_t ld = (T + 3u) & ~3u;
    uint32_t* counts = ws;
    uint32_t* prefix = ws + (size_t)m * ld;
    uint32_t* totals = ws + (size_t)2 * m * ld;
    TimedLaunch tl;
    constexpr uintptr_t kVecAlign = LEN == 2 ? 8 : 16;
    if (T == 1) {
        // One tile: the partition counts it itself (one launch instead of three).
        timing_begin(ctx, s, tl);
        bool done = false;
        if constexpr (LEN == 2 || LEN == 4 || LEN == 8) {
            if ((((uintptr_t)in | (uintptr_t)out) & (kVecAlign - 1)) == 0) {
                hipLaunchKernelGGL((k_partition_rows_vec<kNW, R, LEN, NBITS, Discr, true>), dim3(1),
                                   dim3(kNW * kWave), 0, s, in, out, n, dis, nullptr, nullptr,
                                   0u, m, user_counts);
                done = true;
            }
        }
        if (!done)
            hipLaunchKernelGGL((k_partition_rows<kNW, R, LEN, NBITS, Discr, true>), dim3(1),
                               dim3(kNW * kWave), 0, s, in, out, n, dis, nullptr, nullptr, 0u, m,
                               user_counts);
        timing_end(ctx, s, tl, "rows_partition");
        return ctx_hip(ctx, hipGetLastError(), "array partition launch");
    }
    timing_begin(ctx, s, tl);
    hipLaunchKernelGGL((k_src_count<kNW, R, kCCT, NBITS, KeySrc<uint32_t>, Discr>),
                       dim3((T + kCCT - 1) / kCCT), dim3(kNW * kWave), 0, s,
                       KeySrc<uint32_t>{in, (uint32_t)LEN}, n, dis, counts, ld, m);
    timing_end(ctx, s, tl, "rows_count");
    timing_begin(ctx, s, tl);
    launch_tile_scan(counts, prefix, T, ld, m, totals, user_counts, nullptr, s);
    timing_end(ctx, s, tl, "rows_scan");
    timing_begin(ctx, s, tl);
    if constexpr (LEN == 2 || LEN == 4 || LEN == 8) {
        if ((((uintptr_t)in | (uintptr_t)out) & (kVecAlign - 1)) == 0) {
            hipLaunchKernelGGL((k_partition_rows_vec<kNW, R, LEN, NBITS, Discr>), dim3(T),
                               dim3(kNW * kWave), 0, s, in, out, n, dis, prefix, totals, ld, m,
                               nullptr);
            timing_end(ctx, s, tl, "rows_partition");
            return ctx_hip(ctx, hipGetLastError(), "array partition launch");
        }
    }
    hipLaunchKernelGGL((k_partition_rows<kNW, R, LEN, NBITS, Discr>), dim3(T), dim3(kNW * kWave),
                       0, s, in, out, n, dis, prefix, totals, ld, m, nullptr);
    timing_end(ctx, s, tl, "rows_partition");
    return ctx_hip(ctx, hipGetLastError(), "array partition launch");
}

template <int LEN, class Discr>
mdiscr_status rows_bits(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n, Discr dis,
                        uint32_t m, uint32_t* user_counts, hipStream_t s, uint32_t* ws) {
    switch (class_bits(m)) {
        case 0:
        case 1: return launch_rows<LEN, 1>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 2: return launch_rows<LEN, 2>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 3: return launch_rows<LEN, 3>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 4: return launch_rows<LEN, 4>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 5: return launch_rows<LEN, 5>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 6: return launch_rows<LEN, 6>(ctx, in, out, n, dis, m, user_counts, s, ws);
    }
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "m=%u", m);
}

template <class Discr>
mdiscr_status rows_len(mdiscr_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n,
                       uint32_t len, Discr dis, uint32_t m, uint32_t* user_counts, hipStream_t s,
                       uint32_t* ws) {
    switch (len) {
        case 2: return rows_bits<2>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 3: return rows_bits<3>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 4: return rows_bits<4>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 5: return rows_bits<5>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 6: return rows_bits<6>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 7: return rows_bits<7>(ctx, in, out, n, dis, m, user_counts, s, ws);
        case 8: return rows_bits<8>(ctx, in, out, n, dis, m, user_counts, s, ws);
    }
    return ctx_fail(ctx, MDISCR_EUNSUPPORTED, "array length %u", len);
}

}  // namespace

extern "C" mdiscr_status mdiscr_partition_arrays(mdiscr_ctx* ctx, const uint32_t* d_in,
                                                 uint32_t* d_out, uint64_t n, uint32_t len,
                                                 mdiscr_discr d, uint32_t* d_counts,
                                                 void* stream) {
    if (!ctx) return MDISCR_EINVAL;
    mdiscr_status st = check_discr(ctx, d);
    if (st) return st;
    if (len == 0 || len > 8)
        return ctx_fail(ctx, MDISCR_EINVAL, "array length %u outside [1,8]", len);
    if (len == 1)  // one-word elements: the plain partition
        return mdiscr_partition_u32(ctx, d_in, d_out, n, d, d_counts, stream);
    if (n > MDISCR_MAX_ELEMS) return ctx_fail(ctx, MDISCR_EINVAL, "n too large");
    if (n && (!d_in || !d_out)) return ctx_fail(ctx, MDISCR_EINVAL, "null data pointer");
    if (n && ranges_overlap(d_in, d_out, n * len * 4))
        return ctx_fail(ctx, MDISCR_EINVAL, "output overlaps input");
    hipStream_t s = (hipStream_t)stream;
    st = ctx_begin(ctx, s);
    if (st) return st;
    if (n == 0) {
        if (d_counts) {
            hipLaunchKernelGGL(k_fill_u32<>, dim3(1), dim3(64), 0, s, d_counts, 0u, d.m);
            return ctx_hip(ctx, hipGetLastError(), "fill launch");
        }
        return MDISCR_OK;
    }
    const uint64_t tile = (uint64_t)kNW * kWave * rounds_for((int)len);
    const uint64_t T = (n + tile - 1) / tile;
    const uint64_t ld = (T + 3) & ~3ull;
    void* ws = nullptr;
    st = ctx_workspace(ctx, (2 * ld + 1) * d.m * sizeof(uint32_t), &ws);  // counts, prefix, totals
    if (st) return st;
    uint32_t* w = static_cast<uint32_t*>(ws);
    if (d.kind == MDISCR_DISCR_MASK || (d.m & (d.m - 1)) == 0)
        return rows_len(ctx, d_in, d_out, (uint32_t)n, len, DiscrMask{d.m - 1}, d.m, d_counts, s,
                        w);
    return rows_len(ctx, d_in, d_out, (uint32_t)n, len, DiscrMod{mdiscr_divmagic_make(d.m)}, d.m,
                    d_counts, s, w);
}
