// plain_kernels.hpp -- the plain m-way stable partition on gfx950.
//
// Replaces the reference's transpose + mapVctKernel + scanInc + writeMultiKernel
// pipeline (MdiscrHost_optimized_clean.cu.h:47-118,
// MdiscrKernels_optimized_clean.cu.h:10-183) with three launches:
//
//   chunk_count reads CT tiles per workgroup once and writes every tile's m
//               class counts (class-major, counts[c*ld + t]).        4 B/elem
//   tile_scan   exclusive scan of each class's counts over the T tiles into
//               a prefix array, a few workgroups per class; writes the
//               class totals.
//   partition   CT tiles per workgroup in order (one, as launched), carrying
//               each class's running global offset; reads each tile once,
//               ranks every element inside its wave with
//               NBITS ballots (stable: lower lanes and earlier rounds first),
//               scans the per-wave class counts into per-class LDS cursors,
//               stages the tile in LDS grouped by class, then writes each
//               class run contiguously to
//                   running[c] + (slot - tile_start[c]).               8 B/elem
//
// Where each class starts in the output is fixed by the global totals, which
// are known only after every element has been read once; chunk_count is that
// read.
//
// Tile geometry: NW waves x 64 lanes x R rounds, round r of wave w holding
// elements tile_base + w*64*R + r*64 + lane, so a wave's element order is
// (round, lane) = memory order and each load instruction reads 256 contiguous
// bytes.
#pragma once

#include <type_traits>

#include "device_common.hpp"

// Cache-policy switches for the plain kernels (1 = non-temporal hint).
#ifndef MDISCR_NT_LOAD
#define MDISCR_NT_LOAD 1
#endif
// Output stores of the quad-ranked partition (m <= 16) and of the
// ballot-ranked one (1 = non-temporal hint).  A tile's class runs start and
// end mid-line; with write-back stores the neighbouring tile, which runs on
// the same XCD at about the same time, completes those lines in L2 before
// they go to HBM, where a non-temporal store sends each part on its own.  So
// both write back: 2^28 keys, m=32 0.482 -> 0.444 ms, m=64 0.526 -> 0.504,
// m=2..16 1-2% faster.
#ifndef MDISCR_QUAD_NT_STORE
#define MDISCR_QUAD_NT_STORE 0
#endif
// The one-pass quad kernel under a costly discriminator (x % m) keeps each LDS
// slot's class as a byte for the write-out instead of evaluating x % m again
// there (2^28 keys: m=7 0.416 -> 0.397 ms, m=13 0.472 -> 0.458).  The
// two-digit kernel does not: its 8 KB of class bytes leave room for three
// workgroups per CU instead of four, and it got slower (m=33 0.504 -> 0.521,
// m=48 0.506 -> 0.531).
#ifndef MDISCR_QUAD_CLS
#define MDISCR_QUAD_CLS 1
#endif
#ifndef MDISCR_BALLOT_NT_STORE
#define MDISCR_BALLOT_NT_STORE 0
#endif
#ifndef MDISCR_COUNT_PREFETCH
#define MDISCR_COUNT_PREFETCH 1
#endif
#ifndef MDISCR_COUNT_NT_LOAD
#define MDISCR_COUNT_NT_LOAD 1
#endif

// Copies of each wave's class histogram in the atomic count: lane l adds into
// copy l % REP, so the lanes of one class spread over REP words instead of
// serialising on one (2^28 keys, x % 9: count 0.194 -> 0.150 ms with 4
// copies; x % 13: 0.165 -> 0.151).
#ifndef MDISCR_COUNT_REP
#define MDISCR_COUNT_REP 4
#endif
// Under x % m the count takes the atomic histogram at every class count (1),
// or packed byte counters up to 8 classes as under the mask (0).  The packed
// add (a 64-bit shift and add per key) on top of the modulus's two
// quarter-rate multiplies costs more VALU time than one LDS atomic: 2^28
// keys, count 0.173 -> 0.163 ms at x % 3 / 5 / 6 / 7 (0.160 under the mask);
// 2^26 SoA keys, x % 7: 0.049 -> 0.044 ms.
#ifndef MDISCR_MOD_ATOMIC_COUNT
#define MDISCR_MOD_ATOMIC_COUNT 1
#endif
// Partition tiles in reverse order within each XCD's range: the count pass
// reads every range front to back, so its last tiles may still sit in the
// memory-side cache when the partition starts on them.
#ifndef MDISCR_PLAIN_REVERSE
#define MDISCR_PLAIN_REVERSE 0
#endif

namespace mdiscr {

template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#if MDISCR_NT_LOAD
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
// NT: a non-temporal (streaming) store; else an ordinary write-back store.
template <bool NT, typename T>
__device__ __forceinline__ void st_out(T v, T* p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int NW, int R, int CT>
struct Geom {
    static constexpr int kThreads = NW * kWave;
    static constexpr int kTile = NW * kWave * R;   // elements per tile
    static constexpr int kChunkTiles = CT;           // tiles per workgroup
    static constexpr int kChunk = kTile * CT;        // elements per workgroup
};

// ---------------------------------------------------------------------------
// Per-lane class counters for the count pass.  NBITS <= 4: byte counters
// packed in one or two 64-bit words (one shift and a 64-bit add per element);
// wider class sets count with LDS atomics.
template <int NBITS>
// Packed byte counters pay a per-tile wave reduction (4 wave sums up to 8
// classes, 8 for 16); at 16 classes LDS atomics count faster (measured: SoA
// count over 2^26 keys 0.046 ms atomic vs 0.08 ms packed), at 8 or fewer the
// packed form does (plain 2^28, m=8: 0.167 ms packed vs 0.198 ms atomic).
#ifndef MDISCR_PACKED_MAX_BITS
#define MDISCR_PACKED_MAX_BITS 3
#endif
struct LaneCounts {
    static constexpr bool kPacked = NBITS <= MDISCR_PACKED_MAX_BITS;
    uint64_t lo = 0, hi = 0;
    __device__ __forceinline__ void add(uint32_t c) {
        const uint64_t inc = 1ull << ((c & 7u) * 8u);
        if (NBITS <= 3) {
            lo += inc;
        } else {
            const bool up = (c & 8u) != 0;
            lo += up ? 0ull : inc;
            hi += up ? inc : 0ull;
        }
    }
    // Sum over the wave; lane c (< 2^NBITS) receives the count of class c.
    __device__ __forceinline__ uint32_t wave_class_count(uint32_t lane) const {
        constexpr uint64_t kEven = 0x00FF00FF00FF00FFull;
        uint32_t e0 = (uint32_t)(lo & kEven), e1 = (uint32_t)((lo & kEven) >> 32);
        uint32_t o0 = (uint32_t)((lo >> 8) & kEven), o1 = (uint32_t)(((lo >> 8) & kEven) >> 32);
        e0 = wave_sum(e0); e1 = wave_sum(e1); o0 = wave_sum(o0); o1 = wave_sum(o1);
        uint32_t h0 = 0, h1 = 0, g0 = 0, g1 = 0;
        if (NBITS == 4) {
            h0 = wave_sum((uint32_t)(hi & kEven));
            h1 = wave_sum((uint32_t)((hi & kEven) >> 32));
            g0 = wave_sum((uint32_t)((hi >> 8) & kEven));
            g1 = wave_sum((uint32_t)(((hi >> 8) & kEven) >> 32));
        }
        // class c = byte c: even bytes in e/h, odd bytes in o/g; 16-bit fields.
        const uint32_t c = lane & 15u;
        const bool up = c >= 8u;
        const uint32_t b = c & 7u;
        const uint32_t w0 = (b & 1u) ? (up ? g0 : o0) : (up ? h0 : e0);
        const uint32_t w1 = (b & 1u) ? (up ? g1 : o1) : (up ? h1 : e1);
        const uint32_t f = b >> 1;  // field 0..3 across (w0, w1)
        const uint32_t w = (f < 2u) ? w0 : w1;
        return (w >> ((f & 1u) * 16u)) & 0xFFFFu;
    }
};

template <int NBITS, class Discr>
constexpr bool count_packed() {
    return LaneCounts<NBITS>::kPacked &&
           !(MDISCR_MOD_ATOMIC_COUNT && std::is_same_v<Discr, DiscrMod>);
}

// The atomic count's per-wave histogram: kRep copies of each class's word,
// lane l adding into copy l % kRep, at most 64 words.
template <int MC>
struct HistRep {
    static constexpr int kRep = (MC * MDISCR_COUNT_REP <= 64) ? MDISCR_COUNT_REP
                                : (64 / MC > 1 ? 64 / MC : 1);
    static constexpr int kWords = MC * kRep;
    __device__ __forceinline__ static uint32_t at(uint32_t c, uint32_t lane) {
        return c * kRep + (lane & (kRep - 1));
    }
};

// ---------------------------------------------------------------------------
// count pass: each workgroup reads CT consecutive tiles and writes every
// tile's class counts, counts[c*ld + t] (class-major rows of ld >= T entries).
// VEC: the input is 16-byte aligned, so full tiles use 16-byte loads (a count
// does not depend on element order).
template <int NW, int R, int CT, int NBITS, bool VEC, class Discr>
__global__ __launch_bounds__(NW* kWave) void k_chunk_count(
    const uint32_t* __restrict__ in, uint32_t n, Discr dis,
    uint32_t* __restrict__ counts, uint32_t ld, uint32_t m) {
    using G = Geom<NW, R, CT>;
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = G::kTile;
    constexpr int THREADS = G::kThreads;
    constexpr bool PACKED = count_packed<NBITS, Discr>();
    static_assert(R <= 255, "byte counters per lane");
    static_assert(R % 4 == 0, "16-byte loads");
    using HR = HistRep<MC>;
    __shared__ uint32_t hist[2][NW][PACKED ? MC : HR::kWords];  // double-buffered across tiles

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t chunk = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t T = (n + TILE - 1) / TILE;
    const uint32_t t0 = chunk * CT;
    const uint32_t t1 = min(t0 + (uint32_t)CT, T);

    // 16-byte loads of a full tile (one per-lane pointer, constant offsets).
    auto vec_load = [&](uint32_t tt, u32x4 (&dst)[R / 4]) {
        const u32x4* lp = reinterpret_cast<const u32x4*>(in + (size_t)tt * TILE) + threadIdx.x;
#pragma unroll
        for (int k = 0; k < R / 4; ++k) {
#if MDISCR_COUNT_NT_LOAD
            dst[k] = __builtin_nontemporal_load(lp + k * THREADS);
#else
            dst[k] = lp[k * THREADS];
#endif
        }
    };
    auto tile_full = [&](uint32_t tt) { return (uint64_t)(tt + 1) * TILE <= n; };
    // With MDISCR_COUNT_PREFETCH the next full tile's loads are issued before
    // the current tile is counted.
    u32x4 q[R / 4];
    if (VEC && MDISCR_COUNT_PREFETCH && t0 < t1 && tile_full(t0)) vec_load(t0, q);
    for (uint32_t t = t0; t < t1; ++t) {
        uint32_t (*h)[PACKED ? MC : HR::kWords] = hist[(t - t0) & 1u];
        if (!PACKED) {
            for (int c = lane; c < HR::kWords; c += kWave) h[wave][c] = 0;
            __builtin_amdgcn_wave_barrier();
        }
        LaneCounts<NBITS> lc;
        const uint32_t tbase = t * (uint32_t)TILE;
        const bool full = (uint64_t)tbase + TILE <= n;
        if (VEC && full) {
            u32x4 qn[R / 4];
            const bool pre = MDISCR_COUNT_PREFETCH && t + 1 < t1 && tile_full(t + 1);
            if (!MDISCR_COUNT_PREFETCH) vec_load(t, q);
            if (pre) vec_load(t + 1, qn);
#pragma unroll
            for (int k = 0; k < R / 4; ++k) {
                const uint32_t e[4] = {q[k].x, q[k].y, q[k].z, q[k].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t c = dis(e[j]);
                    if (PACKED) lc.add(c);
                    else atomicAdd(&h[wave][HR::at(c, lane)], 1u);
                }
            }
            if (pre) {
#pragma unroll
                for (int k = 0; k < R / 4; ++k) q[k] = qn[k];
            }
        } else {
            uint32_t v[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t idx = tbase + k * THREADS + threadIdx.x;
                v[k] = (full || idx < n) ? in[idx] : 0u;
            }
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const uint32_t idx = tbase + k * THREADS + threadIdx.x;
                if (full || idx < n) {
                    const uint32_t c = dis(v[k]);
                    if (PACKED) lc.add(c);
                    else atomicAdd(&h[wave][HR::at(c, lane)], 1u);
                }
            }
        }
        if (PACKED) {
            const uint32_t cnt = lc.wave_class_count(lane);
            if (lane < (uint32_t)MC) h[wave][lane] = cnt;
        }
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < m; c += THREADS) {
            uint32_t s = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                if (PACKED) {
                    s += h[w][c];
                } else {
#pragma unroll
                    for (int k = 0; k < HR::kRep; ++k) s += h[w][c * HR::kRep + k];
                }
            }
            counts[(size_t)c * ld + t] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// tile scan: prefix[c*ld + t] <- sum_{t' < t} counts[c*ld + t'] for t < T, and
// the class total to totals[c] (and user_counts[c] when given).  Workgroup
// (c, k) = (blockIdx.x, blockIdx.y) covers chunk k of class c's row, entries
// [k*chunk, (k+1)*chunk).  It first sums the counts before its chunk itself
// (read only, so no workgroup waits on another; a row has few chunks), then
// scans the chunk in passes of THREADS*PER entries: wave w owns the
// contiguous segment of 64*PER of them, and lane l reads and writes entries
// 256j + 4l .. 256j + 4l + 3 of that segment (j < PER/4), so each load and
// store instruction of a wave touches 1 KiB of consecutive entries.  Rows
// start 16-byte aligned (ld a multiple of 4) and chunk is a multiple of
// THREADS*PER.  prefix may be counts itself only with one chunk per row.
template <int THREADS, int PER>
__global__ __launch_bounds__(THREADS) void k_tile_scan(
    const uint32_t* counts, uint32_t* prefix, uint32_t T, uint32_t ld, uint32_t chunk,
    uint32_t* __restrict__ totals, uint32_t* __restrict__ user_counts,
    const uint32_t* __restrict__ rows_dev) {
    static_assert(PER % 4 == 0, "vector loads");
    constexpr int Q = PER / 4;                   // 16-byte accesses per lane per pass
    constexpr uint32_t WSEG = kWave * PER;       // entries per wave per pass
    constexpr int NWAVE = THREADS / kWave;
    // Row count from device memory (a planned queue): rows plus the zero after them.
    if (rows_dev) T = *rows_dev + 1u;
    const uint64_t lo64 = (uint64_t)blockIdx.y * chunk;
    if (lo64 >= T) return;
    const uint32_t lo = (uint32_t)lo64;
    const uint32_t hi = (uint32_t)min((uint64_t)T, lo64 + chunk);
    __shared__ uint32_t wsum[NWAVE];
    const uint32_t* row = counts + (size_t)blockIdx.x * ld;
    uint32_t* orow = prefix + (size_t)blockIdx.x * ld;
    const uint32_t lane = lane_id(), wave = wave_id();

    // The sum of the row before the chunk: 16-byte loads, four per lane in flight.
    uint32_t carry = 0;
    if (lo) {
        const u32x4* r4 = reinterpret_cast<const u32x4*>(row);
        const uint32_t nq = lo / 4;
        uint32_t part = 0, i = threadIdx.x;
        for (; i + 3 * THREADS < nq; i += 4 * THREADS) {
            const u32x4 a = r4[i], b = r4[i + THREADS], c = r4[i + 2 * THREADS],
                        d = r4[i + 3 * THREADS];
            part += ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w)) +
                    ((c.x + c.y) + (c.z + c.w)) + ((d.x + d.y) + (d.z + d.w));
        }
        for (; i < nq; i += THREADS) {
            const u32x4 a = r4[i];
            part += (a.x + a.y) + (a.z + a.w);
        }
        part = wave_sum(part);
        if (lane == 0) wsum[wave] = part;
        __syncthreads();
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) carry += wsum[w];
        __syncthreads();
    }
    for (uint32_t pbase = lo; pbase < hi; pbase += THREADS * PER) {
        const uint32_t seg = pbase + wave * WSEG;
        u32x4 q[Q];
        uint32_t pre[Q];  // entries of the segment before this lane's four, per access
        uint32_t run = 0;
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const uint32_t e0 = seg + j * (kWave * 4) + lane * 4;
            if (e0 + 4 <= hi) {
                q[j] = *reinterpret_cast<const u32x4*>(row + e0);
            } else {
                q[j].x = e0 < hi ? row[e0] : 0u;
                q[j].y = e0 + 1 < hi ? row[e0 + 1] : 0u;
                q[j].z = e0 + 2 < hi ? row[e0 + 2] : 0u;
                q[j].w = 0u;  // e0 + 3 >= hi here
            }
        }
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const uint32_t sum = q[j].x + q[j].y + q[j].z + q[j].w;
            const uint32_t ex = wave_excl_scan(sum, lane);
            pre[j] = run + ex;
            run += (uint32_t)__builtin_amdgcn_readlane((int)(ex + sum), kWave - 1);
        }
        if (lane == 0) wsum[wave] = run;
        __syncthreads();
        uint32_t wpre = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) {
            wpre += (w < (int)wave) ? wsum[w] : 0u;
            tot += wsum[w];
        }
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const uint32_t e0 = seg + j * (kWave * 4) + lane * 4;
            u32x4 o;
            o.x = carry + wpre + pre[j];
            o.y = o.x + q[j].x;
            o.z = o.y + q[j].y;
            o.w = o.z + q[j].z;
            if (e0 + 4 <= hi) {
                *reinterpret_cast<u32x4*>(orow + e0) = o;
            } else {
                if (e0 < hi) orow[e0] = o.x;
                if (e0 + 1 < hi) orow[e0 + 1] = o.y;
                if (e0 + 2 < hi) orow[e0 + 2] = o.z;
            }
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && hi == T) {
        if (totals) totals[blockIdx.x] = carry;
        if (user_counts) user_counts[blockIdx.x] = carry;
    }
}

constexpr int kScanThreads = 1024;
constexpr int kScanPer = 16;
constexpr uint32_t kScanMaxChunks = 16;  // chunks per row: bounds the re-summed prefix

// Launches the tile scan of `rows` class rows of at most t_max entries (the
// entry count comes from rows_dev when given, t_max bounding it).  An
// in-place scan (prefix == counts) runs one chunk per row.
inline void launch_tile_scan(const uint32_t* counts, uint32_t* prefix, uint32_t t_max,
                             uint32_t ld, uint32_t rows, uint32_t* totals, uint32_t* user_counts,
                             const uint32_t* rows_dev, hipStream_t s) {
    constexpr uint32_t kPass = (uint32_t)kScanThreads * kScanPer;
    const uint32_t passes = (uint32_t)(((uint64_t)t_max + kPass - 1) / kPass);  // per row
    uint32_t per_chunk = (passes + kScanMaxChunks - 1) / kScanMaxChunks;
    if (prefix == counts || per_chunk == 0) per_chunk = passes ? passes : 1u;
    const uint32_t chunk = per_chunk * kPass;
    const uint32_t nchunks = passes ? (passes + per_chunk - 1) / per_chunk : 1u;
    hipLaunchKernelGGL((k_tile_scan<kScanThreads, kScanPer>), dim3(rows, nchunks),
                       dim3(kScanThreads), 0, s, counts, prefix, t_max, ld, chunk, totals,
                       user_counts, rows_dev);
}

// ---------------------------------------------------------------------------
// Tile phases shared by the plain and struct-of-arrays partitions.

// Phase 1: in-wave stable rank of R rounds of 64 keys.  Round r of the wave
// holds elements wbase + r*64 + lane, so (round, lane) is memory order.  When
// !FULL, elements past n take class m-1: they sit at the end of the last tile,
// so they rank after every real class m-1 element and land in LDS slots
// [valid, TILE), which are never written out.  packed[r] = class<<16 | rank
// within (wave, class); cursor[c] ends as this wave's class-c count.
template <int R, int NBITS, bool FULL, class Discr>
__device__ __forceinline__ void rank_rounds(const uint32_t (&v)[R], uint32_t (&packed)[R],
                                            uint32_t* cursor, uint32_t wbase, uint32_t n,
                                            uint32_t m, uint32_t lane, Discr dis) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint32_t c = dis(v[r]);
        if (!FULL) c = (wbase + r * kWave + lane < n) ? c : m - 1u;
        uint32_t rk, cnt;
        wave_match_rank<NBITS>(c, lane, rk, cnt);
        const uint32_t run = cursor[c];
        if (rk == cnt - 1u) cursor[c] = run + cnt;
        packed[r] = (c << 16) | (run + rk);
        __builtin_amdgcn_wave_barrier();
    }
}

// Phase 2 (wave 0, lane = class): turn the per-wave class counts in
// cursor[w][c] into LDS slot bases (class-major, wave-minor), set
// dpos[c] = running[c] - tile_start[c], and advance running[c].
template <int NW, int MC>
__device__ __forceinline__ void tile_offsets(uint32_t (*cursor)[MC], uint32_t* dpos,
                                             uint32_t& running, uint32_t lane) {
    uint32_t tot = 0;
    if (lane < (uint32_t)MC) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t x = cursor[w][lane];
            cursor[w][lane] = tot;
            tot += x;
        }
    }
    const uint32_t toff = wave_excl_scan(tot, lane);
    if (lane < (uint32_t)MC) {
#pragma unroll
        for (int w = 0; w < NW; ++w) cursor[w][lane] += toff;
        dpos[lane] = running - toff;
    }
    running += tot;
}

template <int R, bool FULL>
__device__ __forceinline__ void load_rounds(const uint32_t* __restrict__ in, uint32_t (&v)[R],
                                            uint32_t wbase, uint32_t n, uint32_t lane) {
    // One per-lane pointer; round r is the constant offset r * 256 bytes, so
    // the loads share one address register pair.
    const uint32_t* __restrict__ lp = in + wbase + lane;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (FULL) v[r] = ld_stream(lp + r * kWave);
        else v[r] = wbase + r * kWave + lane < n ? ld_stream(lp + r * kWave) : 0u;
    }
}

// Phase 4: slot s = k*THREADS + tid goes to out[dpos[class] + s], in batches
// of WB slots per thread (the LDS reads of a batch are issued together).
template <int R, int THREADS, int WB, bool FULL, bool NT, class Discr>
__device__ __forceinline__ void write_runs(const uint32_t* stage, const uint32_t* dpos,
                                           uint32_t* __restrict__ out, uint32_t valid,
                                           Discr dis) {
    static_assert(R % WB == 0, "write batch");
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += WB) {
        uint32_t x[WB], d[WB];
#pragma unroll
        for (int k = 0; k < WB; ++k) {
            const uint32_t s = (k0 + k) * THREADS + threadIdx.x;
            x[k] = (FULL || s < valid) ? stage[s] : 0u;
        }
#pragma unroll
        for (int k = 0; k < WB; ++k) d[k] = dpos[dis(x[k])];
#pragma unroll
        for (int k = 0; k < WB; ++k) {
            const uint32_t s = (k0 + k) * THREADS + threadIdx.x;
            if (FULL || s < valid) st_out<NT>(x[k], out + (d[k] + s));
        }
    }
}

// write_runs with each slot's class read from cls (bytes in LDS) instead of
// evaluated from the key.
template <int R, int THREADS, int WB, bool FULL, bool NT>
__device__ __forceinline__ void write_runs_cls(const uint32_t* stage, const uint8_t* cls,
                                               const uint32_t* dpos, uint32_t* __restrict__ out,
                                               uint32_t valid) {
    static_assert(R % WB == 0, "write batch");
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += WB) {
        uint32_t x[WB], d[WB];
#pragma unroll
        for (int k = 0; k < WB; ++k) {
            const uint32_t s = (k0 + k) * THREADS + threadIdx.x;
            x[k] = (FULL || s < valid) ? stage[s] : 0u;
            d[k] = (FULL || s < valid) ? (uint32_t)cls[s] : 0u;
        }
#pragma unroll
        for (int k = 0; k < WB; ++k) d[k] = dpos[d[k]];
#pragma unroll
        for (int k = 0; k < WB; ++k) {
            const uint32_t s = (k0 + k) * THREADS + threadIdx.x;
            if (FULL || s < valid) st_out<NT>(x[k], out + (d[k] + s));
        }
    }
}

// ---------------------------------------------------------------------------
// partition: each workgroup runs its chunk's tiles in order, carrying the
// running global offset of every class across tiles (wave 0, lane c).
//
// SINGLE: the input is one tile (n <= TILE), launched as one workgroup with no
// count or scan pass before it: the tile's own class counts are the totals,
// and the kernel writes them to user_counts (when given).
template <int NW, int R, int CT, int NBITS, bool PREFETCH, int WB, int MINW, class Discr,
          bool SINGLE = false>
__global__ __launch_bounds__(NW* kWave, MINW) void k_partition(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
    Discr dis, const uint32_t* __restrict__ prefix,
    const uint32_t* __restrict__ totals, uint32_t ld, uint32_t m,
    uint32_t* __restrict__ user_counts, uint64_t* __restrict__ offsets) {
    using G = Geom<NW, R, CT>;
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = G::kTile;
    constexpr int THREADS = G::kThreads;
    __shared__ uint32_t stage[TILE];
    __shared__ uint32_t cursor[NW][MC];  // per-wave running count, then LDS base
    __shared__ uint32_t dpos[MC];        // global destination - LDS slot, per class

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
#if MDISCR_PLAIN_REVERSE
    const uint32_t chunk = xcd_tile_reversed(blockIdx.x, gridDim.x);
#else
    const uint32_t chunk = xcd_tile(blockIdx.x, gridDim.x);
#endif
    const uint32_t T = (n + TILE - 1) / TILE;
    const uint32_t t0 = chunk * CT;
    const uint32_t t1 = min(t0 + (uint32_t)CT, T);

    // Wave 0, lane c: the first tile's class-c prefix and the class-c total;
    // the running position of the next class-c element is formed at the first
    // tile's phase 2, after these loads have landed.
    uint32_t pre = 0, tot = 0;
    if (!SINGLE && wave == 0 && lane < m) {
        pre = prefix[(size_t)lane * ld + t0];
        tot = totals[lane];
    }

    uint32_t v[R];
    {
        const uint32_t wbase = t0 * (uint32_t)TILE + wave * (uint32_t)(kWave * R);
        if ((uint64_t)(t0 + 1) * TILE <= n) load_rounds<R, true>(in, v, wbase, n, lane);
        else load_rounds<R, false>(in, v, wbase, n, lane);
    }

    uint32_t running = 0;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t tbase = t * (uint32_t)TILE;
        const uint32_t wbase = tbase + wave * (uint32_t)(kWave * R);
        const bool full = (uint64_t)tbase + TILE <= n;
        const uint32_t valid = full ? (uint32_t)TILE : n - tbase;

        for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
        __builtin_amdgcn_wave_barrier();

        // 1. In-wave stable rank.
        uint32_t packed[R];
        if (full) rank_rounds<R, NBITS, true>(v, packed, cursor[wave], wbase, n, m, lane, dis);
        else rank_rounds<R, NBITS, false>(v, packed, cursor[wave], wbase, n, m, lane, dis);
        __syncthreads();

        // Prefetch the next tile while this one is placed and written.
        uint32_t vn[R];
        if (PREFETCH && t + 1 < t1) {
            const uint32_t nb = wbase + (uint32_t)TILE;
            if ((uint64_t)(tbase + TILE) + TILE <= n) load_rounds<R, true>(in, vn, nb, n, lane);
            else load_rounds<R, false>(in, vn, nb, n, lane);
        }

        // 2. Per-class LDS cursors and global offsets.
        if (wave == 0) {
            if constexpr (SINGLE) {
                // Elements past n were ranked as class m-1: not counted.
                uint32_t ct = 0;
                if (lane < (uint32_t)MC)
#pragma unroll
                    for (int w = 0; w < NW; ++w) ct += cursor[w][lane];
                running = wave_excl_scan(ct, lane);
                if (user_counts && lane < m)
                    user_counts[lane] = ct - (lane == m - 1u ? (uint32_t)TILE - n : 0u);
            } else if (t == t0) {
                running = wave_excl_scan(tot, lane) + pre;
            }
            if (offsets && t == 0) {  // the class starts (tile 0's running offsets)
                if (lane < m) offsets[lane] = running;
                if (lane == 0) offsets[m] = n;
            }
            tile_offsets<NW, MC>(cursor, dpos, running, lane);
        }
        __syncthreads();

        // 3. Stage the tile in LDS grouped by class.
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t c = packed[r] >> 16;
            stage[cursor[wave][c] + (packed[r] & 0xFFFFu)] = v[r];
        }
        __syncthreads();

        // 4. Contiguous class runs to global memory.
        if (full) write_runs<R, THREADS, WB, true, MDISCR_BALLOT_NT_STORE>(stage, dpos, out, valid, dis);
        else write_runs<R, THREADS, WB, false, MDISCR_BALLOT_NT_STORE>(stage, dpos, out, valid, dis);

        if (t + 1 < t1) {
            if (PREFETCH) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = vn[r];
            } else {
                const uint32_t nb = wbase + (uint32_t)TILE;
                if ((uint64_t)(tbase + TILE) + TILE <= n) load_rounds<R, true>(in, v, nb, n, lane);
                else load_rounds<R, false>(in, v, nb, n, lane);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Quad partition: the plain partition with 16-byte loads, for up to 16
// classes and a 16-byte aligned input.  Lane l of wave w holds Q = R/4
// quads; quad q holds elements wbase + 256q + 4l + j (j = 0..3), so each load
// instruction reads 1 KiB of consecutive input and the wave's element order
// (q, lane, j) is memory order.
//
// Rank inside a wave, one quad at a time and without ballots: each lane adds
// its four classes into packed byte counters (class c in byte c & 7 of 64-bit
// word c >> 3).  A DPP scan of those words over the wave leaves in byte c the
// number of class-c elements in lower lanes, at most 63 * 4 = 252, so no byte
// carries into the next (only lane 63's inclusive sum can reach 256, and it is
// never used as such).  An element's rank in its (wave, quad, class) group is
// that byte plus the class-c elements before it in its own quad.  The group
// sizes go to LDS, and wave 0 turns them into LDS slot bases ordered by class,
// then wave, then quad: within a class, stage order is memory order.
//
// OUT16: each class run is staged at base_c + 3c + pad_c with pad_c < 4 chosen
// so that LDS slot and output index agree mod 4; the write-out reads four
// slots per lane and stores them with one 16-byte store when all four belong
// to one run, element by element otherwise.  Runs stay disjoint: run c-1 ends
// at or before base_c + 3c.
//
// offsets (optional, device, m+1 entries): the class starts, written by the
// workgroup of tile 0, and offsets[m] = n.
template <int NBITS>
struct QuadRank {
    static_assert(NBITS <= 4, "quad ranking packs at most 16 classes");
    static constexpr int kWords = NBITS <= 3 ? 1 : 2;       // 64-bit counter words
    static constexpr int kScan = NBITS <= 2 ? 1 : 2 * kWords;  // 32-bit halves that can be non-zero

    // byte c of a packed counter set
    static __device__ __forceinline__ uint32_t byte_of(const uint64_t (&w)[2], uint32_t c) {
        // both words are read first: a select between values, not addresses
        const uint64_t w0 = w[0], w1 = w[1];
        const uint64_t x = (NBITS > 3 && (c & 8u)) ? w1 : w0;
        return (uint32_t)(x >> ((c & 7u) * 8u)) & 0xFFu;
    }
    static __device__ __forceinline__ void add(uint64_t (&w)[2], uint32_t c) {
        const uint64_t inc = 1ull << ((c & 7u) * 8u);
        if (NBITS <= 3) {
            w[0] += inc;
        } else {
            const bool up = (c & 8u) != 0;
            w[0] += up ? 0ull : inc;
            w[1] += up ? inc : 0ull;
        }
    }
    static __device__ __forceinline__ uint32_t half(const uint64_t (&w)[2], int i) {
        return (uint32_t)(w[i >> 1] >> ((i & 1) * 32));
    }
    static __device__ __forceinline__ void set_half(uint64_t (&w)[2], int i, uint32_t x) {
        const uint64_t keep = (i & 1) ? 0xFFFFFFFFull : 0xFFFFFFFF00000000ull;
        w[i >> 1] = (w[i >> 1] & keep) | ((uint64_t)x << ((i & 1) * 32));
    }
};

// Inclusive scans of K independent words across the wave, step by step in
// lockstep so that each DPP move's source was written K instructions earlier
// (the wait states a DPP read of a just-written register needs are filled
// with the other words' work).
template <int K>
__device__ __forceinline__ void wave_incl_scan_dpp_n(uint32_t (&x)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0x111, 0xf, 0xf, false);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0x112, 0xf, 0xf, false);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0x114, 0xf, 0xf, false);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0x118, 0xf, 0xf, false);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0x142, 0xa, 0xf, false);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x[k], 0x143, 0xc, 0xf, false);
}

// Phase 1 of the quad partition for one wave: the classes (ck) and the ranks
// within their (wave, quad, class) group (rk) of Q quads, one byte per
// element; group sizes to qb[q][c].  !FULL: elements at or past n (index
// e_base + 256q + 4 lane + j) rank last as class m-1.
template <int Q, int NBITS, bool FULL, class Discr>
__device__ __forceinline__ void quad_rank(const u32x4 (&v)[Q], uint32_t (&ck)[Q], uint32_t (&rk)[Q],
                                          uint32_t (*qb)[1 << NBITS], uint32_t e_base, uint32_t n,
                                          uint32_t m, uint32_t lane, Discr dis) {
    using QR = QuadRank<NBITS>;
    constexpr int MC = 1 << NBITS;
    constexpr int K = QR::kScan;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint32_t x[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
        uint64_t h[2] = {0ull, 0ull};
        uint32_t c[4], lp[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] = dis(x[j]);
            if (!FULL) c[j] = (e_base + q * (kWave * 4) + lane * 4 + j < n) ? c[j] : m - 1u;
            lp[j] = j ? QR::byte_of(h, c[j]) : 0u;
            QR::add(h, c[j]);
        }
        // Exclusive counts of lower lanes: scan, then subtract the own word.
        uint32_t sc[K], own[K];
#pragma unroll
        for (int k = 0; k < K; ++k) own[k] = sc[k] = QR::half(h, k);
        wave_incl_scan_dpp_n<K>(sc);
        uint64_t ex[2] = {0ull, 0ull}, e63[2] = {0ull, 0ull}, h63[2] = {0ull, 0ull};
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t e = sc[k] - own[k];
            QR::set_half(ex, k, e);
            QR::set_half(e63, k, (uint32_t)__builtin_amdgcn_readlane((int)e, kWave - 1));
            QR::set_half(h63, k, (uint32_t)__builtin_amdgcn_readlane((int)own[k], kWave - 1));
        }
        uint32_t cw = 0, rw = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cw |= c[j] << (8 * j);
            rw |= (QR::byte_of(ex, c[j]) + lp[j]) << (8 * j);
        }
        // Materialised here, so the scan words die before the next quad.
        asm volatile("" : "+v"(cw), "+v"(rw));
        ck[q] = cw;
        rk[q] = rw;
        // Group size of class `lane`: lane 63's exclusive count plus its own.
        if (lane < (uint32_t)MC) qb[q][lane] = QR::byte_of(e63, lane) + QR::byte_of(h63, lane);
    }
}

// quad_rank for digits given as bytes j of dq[q] (digit < 2^NBITS of element
// j of quad q): the ranks within their (wave, quad, digit) group to rk, the
// group sizes to qb[q][d].
template <int Q, int NBITS>
__device__ __forceinline__ void quad_rank_digits(const uint32_t (&dq)[Q], uint32_t (&rk)[Q],
                                                 uint32_t (*qb)[1 << NBITS], uint32_t lane) {
    using QR = QuadRank<NBITS>;
    constexpr int MC = 1 << NBITS;
    constexpr int K = QR::kScan;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint64_t h[2] = {0ull, 0ull};
        uint32_t c[4], lp[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] = (dq[q] >> (8 * j)) & 0xFFu;
            lp[j] = j ? QR::byte_of(h, c[j]) : 0u;
            QR::add(h, c[j]);
        }
        uint32_t sc[K], own[K];
#pragma unroll
        for (int k = 0; k < K; ++k) own[k] = sc[k] = QR::half(h, k);
        wave_incl_scan_dpp_n<K>(sc);
        uint64_t ex[2] = {0ull, 0ull}, e63[2] = {0ull, 0ull}, h63[2] = {0ull, 0ull};
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t e = sc[k] - own[k];
            QR::set_half(ex, k, e);
            QR::set_half(e63, k, (uint32_t)__builtin_amdgcn_readlane((int)e, kWave - 1));
            QR::set_half(h63, k, (uint32_t)__builtin_amdgcn_readlane((int)own[k], kWave - 1));
        }
        uint32_t rw = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) rw |= (QR::byte_of(ex, c[j]) + lp[j]) << (8 * j);
        asm volatile("" : "+v"(rw));
        rk[q] = rw;
        if (lane < (uint32_t)MC) qb[q][lane] = QR::byte_of(e63, lane) + QR::byte_of(h63, lane);
    }
}

template <int NW, int R, int NBITS, int WB, bool OUT16, int MINW, class Discr, bool SINGLE = false>
__global__ __launch_bounds__(NW* kWave, MINW) void k_partition_quad(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n, Discr dis,
    const uint32_t* __restrict__ prefix, const uint32_t* __restrict__ totals, uint32_t ld,
    uint32_t m, uint32_t* __restrict__ user_counts, uint64_t* __restrict__ offsets) {
    constexpr int MC = 1 << NBITS;
    constexpr int Q = R / 4;
    constexpr int THREADS = NW * kWave;
    constexpr int TILE = THREADS * R;
    constexpr int SLOTS = OUT16 ? ((TILE + 3 * MC + 3) & ~3) : TILE;
    static_assert(R % 4 == 0, "whole quads per lane");
    __shared__ __attribute__((aligned(16))) uint32_t stage[SLOTS];
    __shared__ uint32_t qbase[NW][Q][MC];  // group sizes, then LDS slot bases
    __shared__ uint32_t dpos[MC];          // output index - LDS slot, per class
    __shared__ uint32_t meta[MC];          // OUT16: run start | valid run length << 16
    constexpr bool CLS = MDISCR_QUAD_CLS && CostlyDiscr<Discr>::value && !OUT16;
    __shared__ uint8_t cls[CLS ? TILE : 1];  // class of each LDS slot

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

    // Loads: one per-lane pointer, quad q at the constant offset q KiB.
    u32x4 v[Q];
    if (full) {
        const u32x4* lp = reinterpret_cast<const u32x4*>(in + wbase) + lane;
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = ld_stream(lp + q * kWave);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t e = wbase + q * (kWave * 4) + lane * 4;
            v[q].x = e < n ? in[e] : 0u;
            v[q].y = e + 1 < n ? in[e + 1] : 0u;
            v[q].z = e + 2 < n ? in[e + 2] : 0u;
            v[q].w = e + 3 < n ? in[e + 3] : 0u;
        }
    }

    // 1. Rank every element in its (wave, quad, class) group.  ck/rk hold the
    // quad's four classes / ranks in bytes j.
    uint32_t ck[Q], rk[Q];
    if (full) quad_rank<Q, NBITS, true>(v, ck, rk, qbase[wave], wbase, n, m, lane, dis);
    else quad_rank<Q, NBITS, false>(v, ck, rk, qbase[wave], wbase, n, m, lane, dis);
    __syncthreads();

    // 2. LDS slot bases per (wave, quad, class) and output offsets per class.
    if (wave == 0) {
        uint32_t g[NW * Q];
        uint32_t ct = 0;
        if (lane < (uint32_t)MC) {
#pragma unroll
            for (int i = 0; i < NW * Q; ++i) {
                g[i] = qbase[i / Q][i % Q][lane];
                const uint32_t x = g[i];
                g[i] = ct;
                ct += x;
            }
        }
        uint32_t running;
        if constexpr (SINGLE) {
            running = wave_excl_scan(ct, lane);
            if (user_counts && lane < m)
                user_counts[lane] = ct - (lane == m - 1u ? (uint32_t)TILE - n : 0u);
        } else {
            running = wave_excl_scan(tot, lane) + pre;
        }
        if (offsets && t == 0) {
            if (lane < m) offsets[lane] = running;
            if (lane == 0) offsets[m] = n;
        }
        const uint32_t toff = wave_excl_scan(ct, lane);
        uint32_t base = toff;
        if (OUT16) {
            const uint32_t a = (uint32_t)(reinterpret_cast<uintptr_t>(out) >> 2) & 3u;
            base = toff + 3u * lane + ((running + a - toff - 3u * lane) & 3u);
            const uint32_t rl = ct - ((lane == m - 1u) ? (uint32_t)TILE - valid : 0u);
            if (lane < (uint32_t)MC) meta[lane] = base | (rl << 16);
        }
        if (lane < (uint32_t)MC) {
#pragma unroll
            for (int i = 0; i < NW * Q; ++i) qbase[i / Q][i % Q][lane] = g[i] + base;
            dpos[lane] = running - base;
        }
    }
    __syncthreads();

    // 3. Stage the tile in LDS grouped by class.
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint32_t x[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = (ck[q] >> (8 * j)) & 0xFFu;
            const uint32_t r = (rk[q] >> (8 * j)) & 0xFFu;
            const uint32_t sl = qbase[wave][q][c] + r;
            stage[sl] = x[j];
            if constexpr (CLS) cls[sl] = (uint8_t)c;
        }
    }
    __syncthreads();

    // 4. Class runs to global memory.
    if constexpr (CLS) {
        if (full) write_runs_cls<R, THREADS, WB, true, MDISCR_QUAD_NT_STORE>(stage, cls, dpos, out, valid);
        else write_runs_cls<R, THREADS, WB, false, MDISCR_QUAD_NT_STORE>(stage, cls, dpos, out, valid);
    } else if constexpr (!OUT16) {
        if (full) write_runs<R, THREADS, WB, true, MDISCR_QUAD_NT_STORE>(stage, dpos, out, valid, dis);
        else write_runs<R, THREADS, WB, false, MDISCR_QUAD_NT_STORE>(stage, dpos, out, valid, dis);
    } else {
        constexpr int GROUPS = SLOTS / 4;
        const u32x4* st4 = reinterpret_cast<const u32x4*>(stage);
#pragma unroll
        for (int k = 0; k < (GROUPS + THREADS - 1) / THREADS; ++k) {
            const uint32_t gi = k * THREADS + threadIdx.x;
            if ((k + 1) * THREADS > GROUPS && gi >= (uint32_t)GROUPS) break;
            const u32x4 x = st4[gi];
            const uint32_t s0 = 4u * gi;
            const uint32_t d0 = dis(x.x), d3 = dis(x.w);
            const uint32_t m0 = meta[d0], m3 = meta[d3];
            const bool in0 = s0 - (m0 & 0xFFFFu) < (m0 >> 16);
            const bool in3 = s0 + 3u - (m3 & 0xFFFFu) < (m3 >> 16);
            if (in0 && in3 && d0 == d3) {
                st_out<MDISCR_QUAD_NT_STORE>(x, reinterpret_cast<u32x4*>(out + (dpos[d0] + s0)));
            } else {
                const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t d = dis(xs[j]);
                    const uint32_t mt = meta[d];
                    if (s0 + j - (mt & 0xFFFFu) < (mt >> 16))
                        st_out<MDISCR_QUAD_NT_STORE>(xs[j], out + (dpos[d] + s0 + j));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The quad partition for 17..64 classes, as two stable digit passes inside
// the tile (least significant first).  Packed byte counters rank at most 16
// classes cheaply; 64 classes by ballots cost about 6 VALU per class bit per
// element.  So the tile is ranked and staged in LDS by the class's low 3 bits
// (8 digits), read back in that order, ranked and staged again by its high
// bits (4 or 8 digits).  Both passes are stable, so LDS then holds the tile
// grouped by class, each class in input order, and the class runs are
// written as in the one-pass kernel.  A class's run starts inside the tile at
// the sum of the tile's counts of the lower classes, which the count pass's
// scan already holds: prefix[c][t + 1] - prefix[c][t] (totals[c] for t + 1).

// Digit of the class: (class >> SHIFT) & MASKD.
template <class Discr, int SHIFT, uint32_t MASKD>
struct ClassDigit {
    Discr dis;
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        return (dis(x) >> SHIFT) & MASKD;
    }
};

// Wave 0, lane = digit d < DC: group sizes qb[w][q][d] (groups in memory
// order) become LDS slot bases ordered by digit, then group.
template <int NW, int Q, int DC>
__device__ __forceinline__ void digit_bases(uint32_t (*qb)[Q][DC], uint32_t lane) {
    uint32_t g[NW * Q];
    uint32_t ct = 0;
    if (lane < (uint32_t)DC) {
#pragma unroll
        for (int i = 0; i < NW * Q; ++i) {
            const uint32_t x = qb[i / Q][i % Q][lane];
            g[i] = ct;
            ct += x;
        }
    }
    const uint32_t start = wave_excl_scan(ct, lane);
    if (lane < (uint32_t)DC) {
#pragma unroll
        for (int i = 0; i < NW * Q; ++i) qb[i / Q][i % Q][lane] = g[i] + start;
    }
}

// Stage quad q's four elements at qb[q][digit] + rank (digits and ranks in
// bytes j of ck / rk).
template <int Q, int DC>
__device__ __forceinline__ void stage_quads(uint32_t* stage, const u32x4 (&v)[Q],
                                            const uint32_t (&ck)[Q], const uint32_t (&rk)[Q],
                                            const uint32_t (*qb)[DC]) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint32_t x[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = (ck[q] >> (8 * j)) & 0xFFu;
            const uint32_t r = (rk[q] >> (8 * j)) & 0xFFu;
            stage[qb[q][c] + r] = x[j];
        }
    }
}

// Tiles of NW waves x R keys per lane (quads as in k_partition_quad), on a
// 16-byte aligned input of at least two tiles (the count pass ran).  Keys past
// n are given the value m - 1, whose class is m - 1 for the mask and the
// modulus alike: both passes then rank them after every real key of class
// m - 1, in the LDS slots at or past `valid`, which are never written out.
template <int NW, int R, int NBITS, int WB, int MINW, class Discr, bool CLS = false>
__global__ __launch_bounds__(NW* kWave, MINW) void k_partition_lsd(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n, Discr dis,
    const uint32_t* __restrict__ prefix, const uint32_t* __restrict__ totals, uint32_t ld,
    uint32_t m, uint64_t* __restrict__ offsets) {
    static_assert(NBITS > 3 && NBITS <= 6, "two digits: 3 low bits, 1..3 high bits");
    constexpr int LB = 3, HB = NBITS - 3;
    constexpr int LC = 1 << LB, HC = 1 << HB;
    constexpr int Q = R / 4;
    constexpr int THREADS = NW * kWave;
    constexpr int TILE = THREADS * R;
    constexpr uint32_t WT = kWave * R;
    static_assert(R % 4 == 0, "whole quads per lane");
    __shared__ __attribute__((aligned(16))) uint32_t stage[TILE];
    __shared__ uint32_t qlo[NW][Q][LC];  // group sizes, then slot bases: low digit
    __shared__ uint32_t qhi[NW][Q][HC];  // the same for the high digit
    __shared__ uint32_t dpos[1 << NBITS];  // output index - LDS slot, per class
    // CLS: each slot's class as a byte, carried from pass 1 to pass 2 and to
    // the write-out, so that the class is evaluated once per key.
    __shared__ __attribute__((aligned(16))) uint8_t cls[CLS ? TILE : 4];

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tbase = t * (uint32_t)TILE;
    const uint32_t wbase = tbase + wave * WT;
    const bool full = (uint64_t)tbase + TILE <= n;
    const uint32_t valid = full ? (uint32_t)TILE : n - tbase;

    // Class offsets of this tile (wave 0, lane = class), issued first.
    uint32_t pre = 0, nxt = 0, tot = 0;
    if (wave == 0 && lane < m) {
        const uint32_t* pr = prefix + (size_t)lane * ld;
        pre = pr[t];
        tot = totals[lane];
        nxt = t + 1u < gridDim.x ? pr[t + 1u] : tot;
    }

    u32x4 v[Q];
    if (full) {
        const u32x4* lp = reinterpret_cast<const u32x4*>(in + wbase) + lane;
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = ld_stream(lp + q * kWave);
    } else {
        const uint32_t fill = m - 1u;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t e = wbase + q * (kWave * 4) + lane * 4;
            v[q].x = e < n ? in[e] : fill;
            v[q].y = e + 1 < n ? in[e + 1] : fill;
            v[q].z = e + 2 < n ? in[e + 2] : fill;
            v[q].w = e + 3 < n ? in[e + 3] : fill;
        }
    }

    // Pass 1: rank by the low digit, stage in (digit, memory) order.
    uint32_t ck[Q], rk[Q];
    [[maybe_unused]] uint32_t cp[CLS ? Q : 1];  // CLS: the classes, bytes j of cp[q]
    if constexpr (CLS) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            cp[q] = dis(v[q].x) | dis(v[q].y) << 8 | dis(v[q].z) << 16 | dis(v[q].w) << 24;
            ck[q] = cp[q] & 0x07070707u;
        }
        quad_rank_digits<Q, LB>(ck, rk, qlo[wave], lane);
    } else {
        quad_rank<Q, LB, true>(v, ck, rk, qlo[wave], 0u, 0u, m, lane,
                               ClassDigit<Discr, 0, LC - 1>{dis});
    }
    __syncthreads();
    if (wave == 0) {
        digit_bases<NW, Q, LC>(qlo, lane);
        const uint32_t running = wave_excl_scan(tot, lane) + pre;
        const uint32_t cstart = wave_excl_scan(nxt - pre, lane);  // class start in the tile
        if (lane < (1u << NBITS)) dpos[lane] = running - cstart;
        if (offsets && t == 0) {
            if (lane < m) offsets[lane] = running;
            if (lane == 0) offsets[m] = n;
        }
    }
    __syncthreads();
    if constexpr (CLS) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t x[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t sl = qlo[wave][q][(ck[q] >> (8 * j)) & 0xFFu] + ((rk[q] >> (8 * j)) & 0xFFu);
                stage[sl] = x[j];
                cls[sl] = (uint8_t)(cp[q] >> (8 * j));
            }
        }
    } else {
        stage_quads<Q, LC>(stage, v, ck, rk, qlo[wave]);
    }
    __syncthreads();

    // Pass 2: the wave's share of that order, ranked by the high digit.
    const u32x4* sp = reinterpret_cast<const u32x4*>(stage + wave * WT) + lane;
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = sp[q * kWave];
    if constexpr (CLS) {
        // The classes of the same four keys: one 32-bit LDS read per quad.
        const uint32_t* cw = reinterpret_cast<const uint32_t*>(cls + wave * WT) + lane;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            cp[q] = cw[q * kWave];
            ck[q] = (cp[q] >> LB) & (0x01010101u * (HC - 1u));
        }
        quad_rank_digits<Q, HB>(ck, rk, qhi[wave], lane);
    } else {
        quad_rank<Q, HB, true>(v, ck, rk, qhi[wave], 0u, 0u, m, lane,
                               ClassDigit<Discr, LB, HC - 1>{dis});
    }
    __syncthreads();  // every wave has read its share of stage (and cls)
    if (wave == 0) digit_bases<NW, Q, HC>(qhi, lane);
    __syncthreads();
    if constexpr (CLS) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t x[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t sl = qhi[wave][q][(ck[q] >> (8 * j)) & 0xFFu] + ((rk[q] >> (8 * j)) & 0xFFu);
                stage[sl] = x[j];
                cls[sl] = (uint8_t)(cp[q] >> (8 * j));
            }
        }
        __syncthreads();
        if (full) write_runs_cls<R, THREADS, WB, true, MDISCR_QUAD_NT_STORE>(stage, cls, dpos, out, valid);
        else write_runs_cls<R, THREADS, WB, false, MDISCR_QUAD_NT_STORE>(stage, cls, dpos, out, valid);
    } else {
        stage_quads<Q, HC>(stage, v, ck, rk, qhi[wave]);
        __syncthreads();
        if (full) write_runs<R, THREADS, WB, true, MDISCR_QUAD_NT_STORE>(stage, dpos, out, valid, dis);
        else write_runs<R, THREADS, WB, false, MDISCR_QUAD_NT_STORE>(stage, dpos, out, valid, dis);
    }
}

// d_counts[c] = 0 for n == 0, and other small fills (a kernel rather than a
// memset: the graphs this library captures hold kernel nodes only).
template <int = 0>
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// sizes[0,n) = 0 except sizes[off[k]] = counts[k] for non-empty classes k
// (tupleToSizeArray, HelpersCommon.h:295-311).
template <int = 0>
__global__ void k_sizes_from_counts(const uint32_t* __restrict__ counts,
                                    uint32_t m, uint32_t* __restrict__ sizes,
                                    uint32_t n) {
    __shared__ uint32_t off[kMaxClasses];
    __shared__ uint32_t cnt[kMaxClasses];
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t k = 0; k < m; ++k) {
            off[k] = acc;
            cnt[k] = counts[k];
            acc += cnt[k];
        }
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        uint32_t s = 0;
        for (uint32_t k = 0; k < m; ++k)
            if (off[k] == i && cnt[k] != 0) s = cnt[k];
        sizes[i] = s;
    }
}

}  // namespace mdiscr
