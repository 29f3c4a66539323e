// segmented_kernels.hpp -- batched segmented m-way stable partition on gfx950.
//
// Replaces the reference's 11-launch segmented pipeline (mdiscrSegm,
// MdiscrSegmHost.cu.h:36-230; kernels MdiscrSegmKernels.cu.h:9-114), whose
// semantics are "plain m-discrimination of every segment in place of its
// range" (validated per segment, MdiscrSegmTest.cu:21-38).
//
// Work is described by a device table of segment descriptors (input and
// output pointers, length, discriminator), so one launch can carry segments of
// one array (offsets) or of unrelated buffers (a serving batch).  A planning
// kernel builds the work queue on the device:
//   W item : the short segments (len <= 64*R) among NW consecutive segment
//            indices, one per wave; each wave partitions its segment alone
//            (no block barrier).
//   T item : tile j of a segment with len > 64*R.  A one-tile segment is
//            written straight from LDS; tiles of longer segments take their
//            offsets from a count pass + scan over "rows" (one row per tile
//            of a multi-tile segment, consecutive within a segment).
// count and partition are persistent: a fixed grid walks the queue, whose
// length is read from device memory, so the whole sequence can be captured
// in a graph once and replayed with new contents.
#pragma once

#include <type_traits>

#include "plain_kernels.hpp"

// Output stores of the segmented partition (1 = non-temporal hint).  Unlike
// the single-array kernels, write-back stores were slower here: 10,000
// lognormal segments, m=4, 0.315-0.321 -> 0.335-0.336 ms per call.
#ifndef MDISCR_SEG_NT_STORE
#define MDISCR_SEG_NT_STORE 1
#endif

namespace mdiscr {

struct SegDesc {
    const uint32_t* in;
    uint32_t* out;
    uint32_t* counts;   // m class counts of this segment, or null
    uint32_t len;
    uint32_t m;
    mdiscr_divmagic g;  // x % m
};

struct SegItem {
    uint32_t seg;     // (first) segment
    uint32_t j;       // T: tile index; W: bit w = segment seg+w is short
    uint32_t ntiles;  // T: tiles of the segment (>= 1); W: 0
    uint32_t row;     // T with ntiles > 1: count row of this tile
};

struct SegPlan {
    uint32_t n_items;
    uint32_t n_rows;
    uint32_t n_long;  // multi-tile segments, expanded into T items by k_seg_expand
};

// A multi-tile segment whose T items the planner reserved but did not write:
// items[item + j] = {seg, j, tiles, row + j} for j < tiles.
struct SegLong {
    uint32_t seg;
    uint32_t item;
    uint32_t tiles;
    uint32_t row;
};

// Up to this many segments the planner writes the T items of the multi-tile
// ones itself (one serving batch), and no expansion kernel runs.
constexpr uint32_t kPlanExpandMax = 256;

template <bool MASK>
struct SegDiscr;
template <>
struct SegDiscr<true> {
    static __device__ __forceinline__ DiscrMask make(const SegDesc& d) { return DiscrMask{d.m - 1u}; }
};
template <>
struct SegDiscr<false> {
    static __device__ __forceinline__ DiscrMod make(const SegDesc& d) { return DiscrMod{d.g}; }
};

// ---------------------------------------------------------------------------
// Segments of one array given by offsets: what every descriptor is built
// from.  With off null, the descriptors are already in the device table (a
// batch's, copied there by the planner).
struct OffSrc {
    const uint64_t* off;   // nseg + 1 offsets, or null
    const uint32_t* in;
    uint32_t* out;
    uint32_t* counts;      // m per segment, or null
    uint32_t m;
    mdiscr_divmagic g;
    uint64_t n;            // elements of in / out
};

// Segment s = [off[s], off[s+1]), clamped to a range inside [0, n): offsets
// that decrease or run past n (unchecked when the caller gave no host copy)
// give segments that are empty or cut at n, never an access outside the
// arrays.
__device__ __forceinline__ SegDesc desc_from_offsets(const OffSrc& os, uint32_t s) {
    const uint64_t a = min(os.off[s], os.n);
    const uint64_t b = min(max(os.off[s + 1], a), os.n);
    SegDesc d;
    d.in = os.in + a;
    d.out = os.out + a;
    d.counts = os.counts ? os.counts + (size_t)s * os.m : nullptr;
    d.len = (uint32_t)(b - a);
    d.m = os.m;
    d.g = os.g;
    return d;
}

// ---------------------------------------------------------------------------
// Planning: one workgroup of THREADS walks the segments in passes, with
// block scans for the item and row counts.
template <int THREADS>
__device__ __forceinline__ void block_excl_sum3(const uint32_t (&v)[3], uint32_t (*wsum)[3],
                                                uint32_t (&pre)[3], uint32_t (&total)[3]) {
    // Three exclusive block sums sharing one pair of barriers (small batch
    // calls 27.6 -> 27.1 us against one block sum, and its barriers, each).
    constexpr int NWAVE = THREADS / kWave;
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    uint32_t wex[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        wex[q] = wave_excl_scan(v[q], lane);
        if (lane == kWave - 1) wsum[wave][q] = wex[q] + v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        uint32_t p = 0, t = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) {
            p += (w < (int)wave) ? wsum[w][q] : 0u;
            t += wsum[w][q];
        }
        pre[q] = p + wex[q];
        total[q] = t;
    }
    __syncthreads();
}

template <int THREADS, int NW, int R>
__global__ __launch_bounds__(THREADS) void k_seg_plan(const SegDesc* __restrict__ staged,
                                                      OffSrc os,
                                                      SegDesc* __restrict__ desc, uint32_t nseg,
                                                      SegItem* __restrict__ items,
                                                      SegLong* __restrict__ longs,
                                                      SegPlan* __restrict__ plan,
                                                      uint32_t* __restrict__ counts, uint32_t ld,
                                                      uint32_t mrows, uint32_t expand_here,
                                                      uint32_t cap_items, uint32_t cap_rows) {
    // Segments are grouped by index, NW per group; a group's short segments
    // (len <= 64*R) form one W item whose bit w marks segment first+w.  Each
    // thread owns SPT consecutive segments (whole groups) per pass.  A
    // one-tile segment gets its T item here; a multi-tile segment gets its
    // items and count rows reserved and one SegLong record, and k_seg_expand
    // writes its T items with a whole workgroup (one thread writing hundreds
    // of items would hold up the block).
    constexpr uint32_t WT = kWave * R;          // one wave's elements
    constexpr uint32_t TILE = NW * kWave * R;   // one workgroup's elements
    constexpr uint32_t SPT = 16;
    constexpr uint32_t PASS = THREADS * SPT;    // segments per pass
    static_assert(SPT % NW == 0, "whole groups per thread");
    __shared__ uint32_t wsum[THREADS / kWave][3];
    // A pass's segment lengths.  Consecutive threads read consecutive
    // offsets or descriptors (every line fetched once, where one thread per
    // SPT segments touched a line per lane and load), and one word of padding
    // per SPT entries puts thread t's entries t*SPT+k in distinct banks.
    __shared__ uint32_t slen[PASS + PASS / SPT];
    // Few segments (expand_here, at most kPlanExpandMax): the multi-tile
    // records stay in LDS for the expansion at the end, which then reads them
    // without one dependent global load per segment (batch of 16 requests,
    // 23M elements: 0.0890 -> 0.0873 ms).
    __shared__ SegLong slong[kPlanExpandMax];
    // A batch's descriptors arrive in mapped pinned memory: copy them to the
    // device table the later kernels read.
    if (staged) {
        for (uint32_t i = threadIdx.x; i < nseg; i += THREADS) desc[i] = staged[i];
        __syncthreads();
    }
    uint32_t item_carry = 0, row_carry = 0, long_carry = 0;
    for (uint32_t base = 0; base < nseg; base += PASS) {
        const uint32_t cnt = min(nseg - base, PASS);
        if (os.off) {
            // Lengths from the offsets (consecutive 8-byte loads).  Few
            // segments: the planner also writes their descriptors, since no
            // expansion kernel runs.
            for (uint32_t i = threadIdx.x; i < cnt; i += THREADS) {
                if (expand_here) {
                    const SegDesc d = desc_from_offsets(os, base + i);
                    desc[base + i] = d;
                    slen[i + i / SPT] = d.len;
                } else {
                    const uint64_t a = min(os.off[base + i], os.n);
                    const uint64_t b = min(max(os.off[base + i + 1], a), os.n);
                    slen[i + i / SPT] = (uint32_t)(b - a);
                }
            }
        } else {
            for (uint32_t i = threadIdx.x; i < cnt; i += THREADS)
                slen[i + i / SPT] = desc[base + i].len;
        }
        __syncthreads();
        const uint32_t s0 = base + threadIdx.x * SPT;
        uint32_t len[SPT];
#pragma unroll
        for (uint32_t k = 0; k < SPT; ++k)
            len[k] = (s0 + k < nseg) ? slen[threadIdx.x * (SPT + 1) + k] : 0u;
        uint32_t nitems = 0, nrows = 0, nlong = 0;
#pragma unroll
        for (uint32_t g = 0; g < SPT; g += NW) {
            bool any_short = false;
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) {
                const uint32_t k = g + w;
                if (s0 + k >= nseg) continue;
                if (len[k] <= WT) {
                    any_short = true;
                } else {
                    const uint32_t tiles = (len[k] + TILE - 1) / TILE;
                    nitems += tiles;
                    nrows += tiles > 1 ? tiles : 0u;
                    nlong += tiles > 1 ? 1u : 0u;
                }
            }
            nitems += any_short ? 1u : 0u;
        }
        const uint32_t part[3] = {nitems, nrows, nlong};
        uint32_t pre[3], tot[3];
        block_excl_sum3<THREADS>(part, wsum, pre, tot);
        uint32_t ib = item_carry + pre[0];
        uint32_t rb = row_carry + pre[1];
        uint32_t lb = long_carry + pre[2];
        const uint32_t itot = tot[0], rtot = tot[1], ltot = tot[2];
#pragma unroll
        for (uint32_t g = 0; g < SPT; g += NW) {
            uint32_t mask = 0;
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w)
                if (s0 + g + w < nseg && len[g + w] <= WT) mask |= 1u << w;
            if (mask) {
                if (ib < cap_items) items[ib] = SegItem{s0 + g, mask, 0u, 0u};
                ++ib;
            }
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) {
                const uint32_t k = g + w;
                if (s0 + k >= nseg || len[k] <= WT) continue;
                const uint32_t tiles = (len[k] + TILE - 1) / TILE;
                if (tiles == 1) {
                    if (ib < cap_items) items[ib] = SegItem{s0 + k, 0u, 1u, 0u};
                } else {
                    const SegLong L{s0 + k, ib, tiles, rb};
                    if (expand_here) slong[lb++] = L;
                    else longs[lb++] = L;
                    rb += tiles;
                }
                ib += tiles;
            }
        }
        item_carry += itot;
        row_carry += rtot;
        long_carry += ltot;
    }
    // Segments whose lengths add up to more than the arrays hold (offsets that
    // overlap; unchecked on the host without a host copy) would need more
    // items or rows than the workspace has: such a plan is left empty, and no
    // later kernel of the call touches memory.
    const bool over = item_carry > cap_items || row_carry > cap_rows;
    if (threadIdx.x == 0)
        *plan = over ? SegPlan{0u, 0u, 0u} : SegPlan{item_carry, row_carry, long_carry};
    // A zero after the last row: the scan's exclusive prefix there is the total.
    if (!over)
        for (uint32_t c = threadIdx.x; c < mrows; c += THREADS) counts[(size_t)c * ld + row_carry] = 0u;
    // Few segments (a serving batch): write the multi-tile segments' T items
    // here, the whole block on one segment at a time, instead of launching
    // k_seg_expand.
    if (expand_here && !over) {
        __syncthreads();
        for (uint32_t k = 0; k < long_carry; ++k) {
            const SegLong L = slong[k];
            for (uint32_t j = threadIdx.x; j < L.tiles; j += THREADS)
                items[L.item + j] = SegItem{L.seg, j, L.tiles, L.row + j};
        }
    }
}

// The T items of the multi-tile segments (one workgroup per SegLong record)
// and, for segments given by offsets, every descriptor.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_seg_expand(const SegLong* __restrict__ longs,
                                                        const SegPlan* __restrict__ plan,
                                                        SegItem* __restrict__ items, OffSrc os,
                                                        SegDesc* __restrict__ desc,
                                                        uint32_t nseg) {
    // Segments given by offsets: every descriptor, grid-stride.
    if (os.off)
        for (uint32_t i = blockIdx.x * THREADS + threadIdx.x; i < nseg; i += gridDim.x * THREADS)
            desc[i] = desc_from_offsets(os, i);
    const uint32_t n_long = plan->n_long;
    for (uint32_t k = blockIdx.x; k < n_long; k += gridDim.x) {
        const SegLong L = longs[k];
        for (uint32_t j = threadIdx.x; j < L.tiles; j += THREADS)
            items[L.item + j] = SegItem{L.seg, j, L.tiles, L.row + j};
    }
}

// ---------------------------------------------------------------------------
// count: class counts of every tile of a multi-tile segment, counts[c*ld+row].
// With tiles of up to 8 rounds a workgroup takes the queue's items in pairs
// (2p, 2p+1).  The T items of one multi-tile segment are consecutive, so a
// pair is mostly two neighbouring tiles of one segment; the loads of both
// tiles are in flight together, which halves the item -> descriptor -> data
// latency chains per element.  (10k lognormal segments, m=4: the 4 x 8 tile's
// whole call 0.3153 -> 0.3076 ms; pairs of 4 x 16 tiles were slower, 0.3107 ->
// 0.3177 ms, so those take one item at a time.  Three or four items per step:
// 0.3101-0.3116 ms against 0.3080-0.3083 ms for pairs.)
#ifndef MDISCR_SEG_COUNT_P
#define MDISCR_SEG_COUNT_P 2u
#endif
template <int NW, int R, int NBITS, bool MASK>
__global__ __launch_bounds__(NW* kWave) void k_seg_count(const SegDesc* __restrict__ desc,
                                                         const SegItem* __restrict__ items,
                                                         const SegPlan* __restrict__ plan,
                                                         uint32_t* __restrict__ counts,
                                                         uint32_t ld) {
    constexpr int MC = 1 << NBITS;
    constexpr int THREADS = NW * kWave;
    constexpr uint32_t TILE = NW * kWave * R;
    constexpr bool PACKED = count_packed<NBITS, std::conditional_t<MASK, DiscrMask, DiscrMod>>();
    using HR = HistRep<MC>;
    constexpr int HW = PACKED ? MC : HR::kWords;  // histogram words per wave
    constexpr uint32_t P = R <= 8 ? MDISCR_SEG_COUNT_P : 1u;  // items per iteration
    __shared__ uint32_t hist[P][NW][HW];
    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t n_items = plan->n_items;
    const uint32_t n_groups = (n_items + P - 1u) / P;
    for (uint32_t p = blockIdx.x; p < n_groups; p += gridDim.x) {
        // An item that is not a tile of a multi-tile segment keeps a zero
        // descriptor (len 0): none of its loads, counts or writes happen.
        SegItem it[P];
        SegDesc d[P] = {};
        bool any = false;
#pragma unroll
        for (uint32_t t = 0; t < P; ++t) {
            it[t] = P * p + t < n_items ? items[P * p + t] : SegItem{0u, 0u, 0u, 0u};
            any |= it[t].ntiles > 1;  // block-uniform
        }
        if (!any) continue;
#pragma unroll
        for (uint32_t t = 0; t < P; ++t)
            if (it[t].ntiles > 1) d[t] = desc[it[t].seg];
        __syncthreads();  // hist reuse
        if (!PACKED) {
            for (int c = lane; c < HW; c += kWave)
#pragma unroll
                for (uint32_t t = 0; t < P; ++t) hist[t][wave][c] = 0;
            __builtin_amdgcn_wave_barrier();
        }
        uint32_t v[P][R];
#pragma unroll
        for (uint32_t t = 0; t < P; ++t) {
            const uint32_t tstart = d[t].len ? it[t].j * TILE : 0u;
            const bool full = tstart + TILE <= d[t].len;
            const uint32_t* __restrict__ lp = d[t].in + tstart + threadIdx.x;  // constant offsets
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t idx = tstart + k * THREADS + threadIdx.x;
                v[t][k] = (full || idx < d[t].len) ? __builtin_nontemporal_load(lp + k * THREADS) : 0u;
            }
        }
#pragma unroll
        for (uint32_t t = 0; t < P; ++t) {
            const uint32_t tstart = d[t].len ? it[t].j * TILE : 0u;
            const bool full = tstart + TILE <= d[t].len;
            const auto dis = SegDiscr<MASK>::make(d[t]);
            LaneCounts<NBITS> lc;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t idx = tstart + k * THREADS + threadIdx.x;
                if (full || idx < d[t].len) {
                    const uint32_t c = dis(v[t][k]);
                    if (PACKED) lc.add(c);
                    else atomicAdd(&hist[t][wave][HR::at(c, lane)], 1u);
                }
            }
            if (PACKED) {
                const uint32_t cnt = lc.wave_class_count(lane);
                if (lane < (uint32_t)MC) hist[t][wave][lane] = cnt;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t t = 0; t < P; ++t) {
            for (uint32_t c = threadIdx.x; c < d[t].m; c += THREADS) {
                uint32_t s = 0;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    if (PACKED) {
                        s += hist[t][w][c];
                    } else {
#pragma unroll
                        for (int k = 0; k < HR::kRep; ++k) s += hist[t][w][c * HR::kRep + k];
                    }
                }
                counts[(size_t)c * ld + it[t].row] = s;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// partition: walks the queue.  STAGED: place each tile in LDS grouped by class
// and write contiguous runs; !STAGED: store every element straight from
// registers to its destination (no staging pass; for few classes the
// destinations of one wave store form a handful of contiguous runs).
template <int NW, int R, int NBITS, bool MASK, bool STAGED>
__global__ __launch_bounds__(NW* kWave) void k_seg_partition(
    const SegDesc* __restrict__ desc, const SegItem* __restrict__ items,
    const SegPlan* __restrict__ plan, const uint32_t* __restrict__ prefix, uint32_t ld) {
    constexpr int MC = 1 << NBITS;
    constexpr int THREADS = NW * kWave;
    constexpr uint32_t WT = kWave * R;
    constexpr uint32_t TILE = NW * WT;
    __shared__ uint32_t stage[TILE];
    __shared__ uint32_t cursor[NW][MC];
    __shared__ uint32_t dpos[MC];

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t n_items = plan->n_items;
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const SegItem it = items[i];
        __syncthreads();  // LDS reuse across items
        if (it.ntiles == 0) {
            // ---- W item: wave w partitions segment it.seg + w on its own
            //      when bit w of it.j is set.
            if ((it.j >> wave) & 1u) {
                const SegDesc d = desc[it.seg + wave];
                const auto dis = SegDiscr<MASK>::make(d);
                uint32_t v[R], packed[R];
                load_rounds<R, false>(d.in, v, 0u, d.len, lane);
                for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
                __builtin_amdgcn_wave_barrier();
                rank_rounds<R, NBITS, false>(v, packed, cursor[wave], 0u, d.len, d.m, lane, dis);
                // class bases inside the segment (lane = class)
                const uint32_t cnt = lane < (uint32_t)MC ? cursor[wave][lane] : 0u;
                const uint32_t toff = wave_excl_scan(cnt, lane);
                if (lane < (uint32_t)MC) cursor[wave][lane] = toff;
                if (d.counts && lane < d.m)
                    d.counts[lane] = cnt - (lane == d.m - 1u ? WT - d.len : 0u);
                __builtin_amdgcn_wave_barrier();
                uint32_t* st = stage + wave * WT;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint32_t c = packed[r] >> 16;
                    const uint32_t slot = cursor[wave][c] + (packed[r] & 0xFFFFu);
                    if (STAGED) st[slot] = v[r];
                    else if (slot < d.len) st_out<MDISCR_SEG_NT_STORE>(v[r], d.out + slot);
                }
                if (STAGED) {
                    __builtin_amdgcn_wave_barrier();
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        const uint32_t s = k * kWave + lane;
                        if (s < d.len) st_out<MDISCR_SEG_NT_STORE>(st[s], d.out + s);
                    }
                }
            }
            continue;
        }
        // ---- T item: tile it.j of segment it.seg.
        const SegDesc d = desc[it.seg];
        const auto dis = SegDiscr<MASK>::make(d);
        const uint32_t tstart = it.j * TILE;
        const uint32_t wbase = tstart + wave * WT;
        const bool full = tstart + TILE <= d.len;
        const uint32_t valid = full ? TILE : d.len - tstart;
        uint32_t pj = 0, p0 = 0, pe = 0;
        if (wave == 0 && it.ntiles > 1 && lane < d.m) {
            const uint32_t row0 = it.row - it.j;
            const uint32_t* pr = prefix + (size_t)lane * ld;
            pj = pr[it.row];
            p0 = pr[row0];
            pe = pr[row0 + it.ntiles];
        }
        uint32_t v[R], packed[R];
        if (full) load_rounds<R, true>(d.in, v, wbase, d.len, lane);
        else load_rounds<R, false>(d.in, v, wbase, d.len, lane);
        for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
        __builtin_amdgcn_wave_barrier();
        if (full) rank_rounds<R, NBITS, true>(v, packed, cursor[wave], wbase, d.len, d.m, lane, dis);
        else rank_rounds<R, NBITS, false>(v, packed, cursor[wave], wbase, d.len, d.m, lane, dis);
        __syncthreads();
        if (wave == 0) {
            uint32_t running;
            if (it.ntiles > 1) {
                // segment-relative: class base from the segment totals, plus
                // this tile's prefix within the segment
                running = wave_excl_scan(pe - p0, lane) + (pj - p0);
                if (it.j == 0 && d.counts && lane < d.m) d.counts[lane] = pe - p0;
            } else {
                running = 0;  // one tile: LDS order is the output order
            }
            const uint32_t before = running;
            tile_offsets<NW, MC>(cursor, dpos, running, lane);
            if (it.ntiles == 1) {
                if (lane < (uint32_t)MC) dpos[lane] = 0u;
                const uint32_t tot = running - before;
                if (d.counts && lane < d.m)
                    d.counts[lane] = tot - (lane == d.m - 1u ? TILE - valid : 0u);
            }
        }
        __syncthreads();
        if (STAGED) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t c = packed[r] >> 16;
                stage[cursor[wave][c] + (packed[r] & 0xFFFFu)] = v[r];
            }
            __syncthreads();
            if (full) write_runs<R, THREADS, (R < 8 ? R : 8), true, MDISCR_SEG_NT_STORE>(stage, dpos, d.out, valid, dis);
            else write_runs<R, THREADS, (R < 8 ? R : 8), false, MDISCR_SEG_NT_STORE>(stage, dpos, d.out, valid, dis);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t c = packed[r] >> 16;
                const uint32_t slot = cursor[wave][c] + (packed[r] & 0xFFFFu);
                if (full || slot < valid)
                    st_out<MDISCR_SEG_NT_STORE>(v[r], d.out + (dpos[c] + slot));
            }
        }
    }
}

}  // namespace mdiscr
