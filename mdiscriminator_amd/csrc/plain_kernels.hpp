// plain_kernels.hpp -- the plain m-way stable partition on gfx950.
//
// Replaces the reference's transpose + mapVctKernel + scanInc + writeMultiKernel
// pipeline (MdiscrHost_optimized_clean.cu.h:47-118,
// MdiscrKernels_optimized_clean.cu.h:10-183) with three launches:
//
//   tile_hist   reads each tile once and writes its m class counts
//               (class-major, counts[c*T + t]).                      4 B/elem
//   tile_scan   one workgroup per class: exclusive scan of that class's
//               counts over the T tiles, in place; writes the class total.
//   partition   reads each tile once, ranks every element inside its wave with
//               NBITS ballots (stable: lower lanes and earlier rounds first),
//               scans the per-wave class counts into per-class LDS cursors,
//               stages the tile in LDS grouped by class, then writes each
//               class run contiguously to
//                   base[c] + prefix[c][t] + (slot - tile_start[c]).  8 B/elem
//
// The class of the output rows is fixed by the global totals, which are known
// only after every element has been read once; tile_hist is that read.
//
// Tile geometry: NW waves x 64 lanes x R rounds, round r of wave w holding
// elements tile_base + w*64*R + r*64 + lane, so a wave's element order is
// (round, lane) = memory order and each load instruction reads 256 contiguous
// bytes.
#pragma once

#include "device_common.hpp"

namespace mdiscr {

template <int NW, int R>
struct TileGeom {
    static constexpr int kThreads = NW * kWave;
    static constexpr int kTile = NW * kWave * R;
};

// ---------------------------------------------------------------------------
// tile_hist: counts[c*T + t] = number of class-c elements of tile t.
template <int NW, int R, int NBITS, class Discr>
__global__ __launch_bounds__(NW* kWave) void k_tile_hist(
    const uint32_t* __restrict__ in, uint32_t n, Discr dis,
    uint32_t* __restrict__ counts, uint32_t T, uint32_t m) {
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = TileGeom<NW, R>::kTile;
    __shared__ uint32_t hist[NW][MC];

    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tbase = tile * (uint32_t)TILE;
    const uint32_t wbase = tbase + wave * (uint32_t)(kWave * R);

    for (int c = lane; c < MC; c += kWave) hist[wave][c] = 0;
    __builtin_amdgcn_wave_barrier();

    uint32_t v[R];
    const bool full = (uint64_t)tbase + TILE <= n;
    if (full) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = __builtin_nontemporal_load(in + wbase + r * kWave + lane);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t idx = wbase + r * kWave + lane;
            v[r] = idx < n ? in[idx] : 0u;
        }
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t mine = 0;  // NBITS <= 3: lane k accumulates class k's count
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t idx = wbase + r * kWave + lane;
        const bool ok = full || idx < n;
        const uint32_t c = dis(v[r]);
        if (NBITS <= 3) {
            // Few classes: one ballot per class, counted on the scalar unit.
#pragma unroll
            for (int k = 0; k < MC; ++k) {
                const uint64_t B = ballot64(ok && c == (uint32_t)k);
                if ((uint32_t)k == lane) mine += (uint32_t)__popcll(B);
            }
        } else {
            if (ok) atomicAdd(&hist[wave][c], 1u);
        }
    }
    if (NBITS <= 3 && lane < (uint32_t)MC) hist[wave][lane] = mine;
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < m; c += NW * kWave) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += hist[w][c];
        counts[(size_t)c * T + tile] = s;
    }
}

// ---------------------------------------------------------------------------
// tile_scan: for class c = blockIdx.x, counts[c*T + t] <- sum_{t'<t} counts,
// totals[c] <- sum over all tiles; user_counts[c] too when given.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_tile_scan(
    uint32_t* __restrict__ counts, uint32_t T, uint32_t* __restrict__ totals,
    uint32_t* __restrict__ user_counts) {
    constexpr int NWAVE = THREADS / kWave;
    __shared__ uint32_t wsum[NWAVE];
    uint32_t* row = counts + (size_t)blockIdx.x * T;
    const uint32_t chunk = (T + THREADS - 1) / THREADS;
    const uint32_t beg = threadIdx.x * chunk;
    const uint32_t end = min(beg + chunk, T);
    uint32_t s = 0;
    for (uint32_t t = beg; t < end; ++t) s += row[t];
    const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
    const uint32_t wex = wave_excl_scan(s, lane);
    if (lane == kWave - 1) wsum[wave] = wex + s;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) {
        wpre += (w < (int)wave) ? wsum[w] : 0u;
        tot += wsum[w];
    }
    uint32_t acc = wpre + wex;
    for (uint32_t t = beg; t < end; ++t) {
        const uint32_t x = row[t];
        row[t] = acc;
        acc += x;
    }
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = tot;
        if (user_counts) user_counts[blockIdx.x] = tot;
    }
}

// ---------------------------------------------------------------------------
// partition: stable scatter of one tile per workgroup.
template <int NW, int R, int NBITS, class Discr>
__global__ __launch_bounds__(NW* kWave) void k_partition(
    const uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n,
    Discr dis, const uint32_t* __restrict__ prefix,
    const uint32_t* __restrict__ totals, uint32_t T, uint32_t m) {
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = TileGeom<NW, R>::kTile;
    constexpr int THREADS = NW * kWave;
    __shared__ uint32_t stage[TILE];
    __shared__ uint32_t cursor[NW][MC];  // per-wave running count, then LDS base
    __shared__ uint32_t dpos[MC];        // global destination - LDS slot, per class

    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tbase = tile * (uint32_t)TILE;
    const uint32_t wbase = tbase + wave * (uint32_t)(kWave * R);
    const bool full = (uint64_t)tbase + TILE <= n;
    const uint32_t valid = full ? (uint32_t)TILE : n - tbase;

    for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
    __builtin_amdgcn_wave_barrier();

    // Wave 0 fetches this tile's class prefixes and the class totals early.
    uint32_t my_prefix = 0, my_total = 0;
    if (wave == 0 && lane < m) {
        my_prefix = prefix[(size_t)lane * T + tile];
        my_total = totals[lane];
    }

    uint32_t v[R];
    if (full) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = __builtin_nontemporal_load(in + wbase + r * kWave + lane);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t idx = wbase + r * kWave + lane;
            v[r] = idx < n ? in[idx] : 0u;
        }
    }

    // 1. In-wave stable rank.  Elements past n take class m-1: they sit at the
    //    end of the last tile, so they rank after every real class m-1 element
    //    and land in LDS slots [valid, TILE), which are never written out.
    uint32_t packed[R];  // class << 16 | rank within (wave, class)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t idx = wbase + r * kWave + lane;
        const uint32_t c = (full || idx < n) ? dis(v[r]) : m - 1u;
        uint32_t rk, cnt;
        wave_match_rank<NBITS>(c, lane, rk, cnt);
        const uint32_t run = cursor[wave][c];
        if (rk == cnt - 1u) cursor[wave][c] = run + cnt;
        packed[r] = (c << 16) | (run + rk);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // 2. Per-class LDS cursors (class-major, wave-minor) and global offsets.
    if (wave == 0) {
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
        const uint32_t cbase = wave_excl_scan(my_total, lane);
        if (lane < (uint32_t)MC) {
#pragma unroll
            for (int w = 0; w < NW; ++w) cursor[w][lane] += toff;
            dpos[lane] = cbase + my_prefix - toff;
        }
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
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint32_t s = k * THREADS + threadIdx.x;
        if (s < valid) {
            const uint32_t x = stage[s];
            const uint32_t c = dis(x);
            __builtin_nontemporal_store(x, out + (dpos[c] + s));
        }
    }
}

// d_counts[c] = 0 for n == 0 (a fill kernel rather than a memset, so callers
// may capture the call in a graph).
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = v;
}

// sizes[0,n) = 0 except sizes[off[k]] = counts[k] for non-empty classes k
// (tupleToSizeArray, HelpersCommon.h:295-311).
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
