// soa_kernels.hpp -- struct-of-arrays m-way stable partition on gfx950.
//
// Replaces the reference's three-column tuple path (MdiscrHost_tuples.cu.h:
// 262-429, MdiscrKernels_tuples.cu.h:4-236): the class comes from one key
// column (ModTuple<N>::apply, MdiscrHost_tuples.cu.h:44-45) and every column
// moves to the same destination, as raw bits.  A key source carries an
// element stride (the count pass over array elements reads word 0 at stride
// LEN); the partition's columns are contiguous.  A
// 1-byte key column lets a caller-computed class array be the discriminator;
// a select source reads the class from one of two columns, chosen per element
// by a third (the tuple discriminator's z ? x % N : y % N form).
//
// Same three launches as the plain path; the count pass reads only the key
// column.  The partition ranks the key exactly as the plain kernel does,
// records each element's class in LDS beside its slot, then moves the columns
// one at a time through one LDS staging buffer of element type ST (4 bytes
// when no column is wider, so larger tiles fit): load (same indices as the
// keys) -> stage at the element's slot -> write contiguous class runs.
// Bytes per record: key bytes (count) + 2 * sum(column bytes).
#pragma once

#include <type_traits>

#include "plain_kernels.hpp"

// Cache policy of the SoA column stores (1 = non-temporal hint).  Write-back
// stores let neighbouring tiles' partial lines at class-run ends merge in L2:
// 2^26 (key, f32, index) records, m=4 / 16 / 32 went 0.316 -> 0.288,
// 0.325 -> 0.298, 0.373 -> 0.307 ms.
#ifndef MDISCR_SOA_NT_STORE
#define MDISCR_SOA_NT_STORE 0
#endif

namespace mdiscr {

constexpr int kMaxCols = 8;

struct SoaCols {
    const void* in[kMaxCols];
    void* out[kMaxCols];
    uint32_t bytes[kMaxCols];
    uint32_t ncols;
    uint32_t key;
};

template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
    return __builtin_nontemporal_load(p);
}

// ---------------------------------------------------------------------------
// Class sources: the 32-bit value the discriminator sees for element idx.

// Every class source loads R elements at once, first + r*step (r < R), with
// elements at or past n reading 0 unless `full`: gather().  Loads of one
// column are issued back to back.

// Loads R elements of a contiguous column of T through one per-lane pointer.
template <typename T, int R>
__device__ __forceinline__ void gather_col(const T* __restrict__ p, uint32_t first, uint32_t step,
                                           uint32_t n, bool full, uint32_t (&v)[R]) {
    const T* __restrict__ lp = p + first;
#pragma unroll
    for (int r = 0; r < R; ++r)
        v[r] = (full || first + r * step < n) ? (uint32_t)ld_nt(lp + (size_t)r * step) : 0u;
}

// One key column (4-byte keys or 1-byte classes) with an element stride.  The
// values in registers are that column's elements, so the partition writes the
// key column from registers.
template <typename KT>
struct KeySrc {
    using KeyT = KT;
    static constexpr bool kKeyInRegs = true;
    const KT* key;
    uint32_t stride;
    template <int R>
    __device__ __forceinline__ void gather(uint32_t first, uint32_t step, uint32_t n, bool full,
                                           uint32_t (&v)[R]) const {
        if (stride == 1) {
            gather_col<KT, R>(key, first, step, n, full, v);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t idx = first + r * step;
                v[r] = (full || idx < n) ? (uint32_t)ld_nt(key + (size_t)idx * stride) : 0u;
            }
        }
    }
};

// z ? x : y over three contiguous columns of 1- or 4-byte values.  With one
// discriminator D, z ? D(x) : D(y) = D(z ? x : y), so the select form of the
// tuple discriminator costs one D per element.  Each column's width is a
// wave-uniform branch taken once per gather, not per element.
struct SelectSrc {
    using KeyT = uint32_t;
    static constexpr bool kKeyInRegs = false;
    const void* sel;
    const void* a;
    const void* b;
    uint32_t sel4, a4, b4;  // nonzero when that column holds 4-byte values
    uint32_t sel_col, a_col, b_col;  // their indices among the moved columns
    template <int R>
    __device__ __forceinline__ static void col(const void* p, uint32_t w4, uint32_t first,
                                               uint32_t step, uint32_t n, bool full,
                                               uint32_t (&v)[R]) {
        if (w4) gather_col<uint32_t, R>(static_cast<const uint32_t*>(p), first, step, n, full, v);
        else gather_col<uint8_t, R>(static_cast<const uint8_t*>(p), first, step, n, full, v);
    }
    template <int R>
    __device__ __forceinline__ void gather(uint32_t first, uint32_t step, uint32_t n, bool full,
                                           uint32_t (&v)[R]) const {
        uint32_t z[R], y[R];
        col<R>(sel, sel4, first, step, n, full, z);
        col<R>(a, a4, first, step, n, full, v);
        col<R>(b, b4, first, step, n, full, y);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = z[r] ? v[r] : y[r];
    }
    // gather, keeping the three columns' values (zero-extended) as well.
    template <int R>
    __device__ __forceinline__ void gather_keep(uint32_t first, uint32_t step, uint32_t n,
                                                bool full, uint32_t (&v)[R], uint32_t (&z)[R],
                                                uint32_t (&x)[R], uint32_t (&y)[R]) const {
        col<R>(sel, sel4, first, step, n, full, z);
        col<R>(a, a4, first, step, n, full, x);
        col<R>(b, b4, first, step, n, full, y);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = z[r] ? x[r] : y[r];
    }
};

template <class Src>
struct IsSelect : std::false_type {};
template <>
struct IsSelect<SelectSrc> : std::true_type {};

// The select source's partition writes its three columns from the values it
// loaded for the classes instead of loading them again.
#ifndef MDISCR_SEL_KEEP
#define MDISCR_SEL_KEEP 1
#endif

// ---------------------------------------------------------------------------
// count pass over a class source (a strided or 1-byte key column, or a select).
template <int NW, int R, int CT, int NBITS, class Src, class Discr>
__global__ __launch_bounds__(NW* kWave) void k_src_count(
    Src src, uint32_t n, Discr dis, uint32_t* __restrict__ counts, uint32_t ld, uint32_t m) {
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = NW * kWave * R;
    constexpr int THREADS = NW * kWave;
    constexpr bool PACKED = count_packed<NBITS, Discr>();
    static_assert(R <= 255, "byte counters per lane");
    using HR = HistRep<MC>;
    __shared__ uint32_t hist[2][NW][PACKED ? MC : HR::kWords];
    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t chunk = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t T = (n + TILE - 1) / TILE;
    const uint32_t t0 = chunk * CT;
    const uint32_t t1 = min(t0 + (uint32_t)CT, T);
    for (uint32_t t = t0; t < t1; ++t) {
        uint32_t (*h)[PACKED ? MC : HR::kWords] = hist[(t - t0) & 1u];
        if (!PACKED) {
            for (int c = lane; c < HR::kWords; c += kWave) h[wave][c] = 0;
            __builtin_amdgcn_wave_barrier();
        }
        LaneCounts<NBITS> lc;
        const uint32_t tbase = t * (uint32_t)TILE;
        const bool full = (uint64_t)tbase + TILE <= n;
        uint32_t v[R];
        src.template gather<R>(tbase + threadIdx.x, THREADS, n, full, v);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t idx = tbase + k * THREADS + threadIdx.x;
            if (full || idx < n) {
                const uint32_t c = dis(v[k]);
                if (PACKED) lc.add(c);
                else atomicAdd(&h[wave][HR::at(c, lane)], 1u);
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
// Where the tile's slot s = k*THREADS + tid goes: dpos[class of s] + s.  Kept
// in registers (one per slot the thread writes), or, with LDS, in an LDS word
// per slot that only the thread owning the slot writes and reads.  LDS frees
// R registers (107 -> 85 at 8 waves x 12 rounds) for one LDS read per slot and
// column: it pays only for the select source (2^26 (uint, uchar, bool)
// records, m=8: partition 0.267 -> 0.254 ms; 4-byte key columns 2-5% slower).
// Against looking the class and dpos up again for every column (two LDS
// reads), the LDS word took the select's partition 0.234 -> 0.222 ms at
// m=32 and 0.245 -> 0.233 at m=64.
template <int R, bool LDS>
struct SlotDest {
    uint32_t reg[LDS ? 1 : R];
    const uint32_t* dslot;
    __device__ __forceinline__ uint32_t at(int k, uint32_t s) const {
        if constexpr (LDS) return dslot[s];
        else return reg[k];
    }
};

// Moves one column of the current tile: load at the tile's element indices,
// stage at each element's LDS slot, write slot s to out[dest(s)].
// FULL: a whole tile, loaded and written without per-element bounds tests.
template <typename T, int NW, int R, bool FULL, bool DL>
__device__ __forceinline__ void soa_move_column(const void* __restrict__ src_v,
                                                void* __restrict__ dst_v, uint32_t wbase,
                                                uint32_t n, const uint32_t (&slot)[R],
                                                const SlotDest<R, DL>& dest, uint32_t valid,
                                                void* stage_v, uint32_t lane) {
    constexpr int THREADS = NW * kWave;
    // one per-lane pointer; round r is a constant offset
    const T* __restrict__ lp = static_cast<const T*>(src_v) + wbase + lane;
    T* __restrict__ dst = static_cast<T*>(dst_v);
    T* stage = static_cast<T*>(stage_v);
    T x[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        x[r] = (FULL || wbase + r * kWave + lane < n) ? ld_nt(lp + r * kWave) : T(0);
#pragma unroll
    for (int r = 0; r < R; ++r) stage[slot[r]] = x[r];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint32_t s = k * THREADS + threadIdx.x;
        if (FULL || s < valid) st_out<MDISCR_SOA_NT_STORE>(stage[s], dst + dest.at(k, s));
    }
    __syncthreads();
}

// The key column is already in registers: stage and write it without a reload.
template <typename KT, int NW, int R, bool FULL, bool DL>
__device__ __forceinline__ void soa_write_keys(const uint32_t (&v)[R], void* __restrict__ dst_v,
                                               const uint32_t (&slot)[R],
                                               const SlotDest<R, DL>& dest, uint32_t valid,
                                               void* stage_v) {
    constexpr int THREADS = NW * kWave;
    KT* __restrict__ dst = static_cast<KT*>(dst_v);
    KT* stage = static_cast<KT*>(stage_v);
#pragma unroll
    for (int r = 0; r < R; ++r) stage[slot[r]] = (KT)v[r];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const uint32_t s = k * THREADS + threadIdx.x;
        if (FULL || s < valid) st_out<MDISCR_SOA_NT_STORE>(stage[s], dst + dest.at(k, s));
    }
    __syncthreads();
}

// SINGLE: the records are one tile, launched as one workgroup with no count or
// scan pass: the tile's own class counts are the totals, written to
// user_counts (when given).
template <int NW, int R, int CT, int NBITS, class Src, typename ST, class Discr,
          bool SINGLE = false>
__global__ __launch_bounds__(NW* kWave) void k_partition_soa(
    SoaCols cols, Src src, uint32_t n, Discr dis, const uint32_t* __restrict__ prefix,
    const uint32_t* __restrict__ totals, uint32_t ld, uint32_t m,
    uint32_t* __restrict__ user_counts) {
    constexpr int MC = 1 << NBITS;
    constexpr int TILE = NW * kWave * R;
    constexpr int THREADS = NW * kWave;
    __shared__ ST stage[TILE];           // widest column element
    __shared__ uint8_t cls[TILE];        // class of each LDS slot
    constexpr bool kDestLds = std::is_same_v<Src, SelectSrc>;
    __shared__ uint32_t dslot[kDestLds ? TILE : 1];  // slot destinations (own thread's)
    __shared__ uint32_t cursor[NW][MC];
    __shared__ uint32_t dpos[MC];

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t chunk = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t T = (n + TILE - 1) / TILE;
    const uint32_t t0 = chunk * CT;
    const uint32_t t1 = min(t0 + (uint32_t)CT, T);

    uint32_t pre = 0, tot = 0;
    if (!SINGLE && wave == 0 && lane < m) {
        pre = prefix[(size_t)lane * ld + t0];
        tot = totals[lane];
    }

    uint32_t running = 0;
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t tbase = t * (uint32_t)TILE;
        const uint32_t wbase = tbase + wave * (uint32_t)(kWave * R);
        const bool full = (uint64_t)tbase + TILE <= n;
        const uint32_t valid = full ? (uint32_t)TILE : n - tbase;

        uint32_t v[R];
        constexpr bool kSelKeep = MDISCR_SEL_KEEP && std::is_same_v<Src, SelectSrc>;
        [[maybe_unused]] uint32_t sz[kSelKeep ? R : 1], sa[kSelKeep ? R : 1], sb[kSelKeep ? R : 1];
        if constexpr (kSelKeep)
            src.template gather_keep<R>(wbase + lane, kWave, n, full, v, sz, sa, sb);
        else if constexpr (Src::kKeyInRegs)  // the partition's key column is contiguous
            gather_col<typename Src::KeyT, R>(src.key, wbase + lane, kWave, n, full, v);
        else
            src.template gather<R>(wbase + lane, kWave, n, full, v);
        for (int c = lane; c < MC; c += kWave) cursor[wave][c] = 0;
        __builtin_amdgcn_wave_barrier();

        // 1. In-wave stable rank of the keys (as the plain kernel).
        uint32_t slot[R];  // class << 16 | rank, then the LDS slot
        if (full) rank_rounds<R, NBITS, true>(v, slot, cursor[wave], wbase, n, m, lane, dis);
        else rank_rounds<R, NBITS, false>(v, slot, cursor[wave], wbase, n, m, lane, dis);
        __syncthreads();

        // 2. LDS cursors and global offsets (wave 0, lane = class).
        if (wave == 0) {
            if constexpr (SINGLE) {
                // Records past n were ranked as class m-1: not counted.
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
            tile_offsets<NW, MC>(cursor, dpos, running, lane);
        }
        __syncthreads();

        // 3. Slots, the class of every slot, and each output slot's destination.
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t c = slot[r] >> 16;
            slot[r] = cursor[wave][c] + (slot[r] & 0xFFFFu);
            cls[slot[r]] = (uint8_t)c;
        }
        __syncthreads();
        SlotDest<R, kDestLds> dest;
        dest.dslot = dslot;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t s = k * THREADS + threadIdx.x;
            const uint32_t d = (full || s < valid) ? dpos[cls[s]] + s : 0u;
            if constexpr (kDestLds) dslot[s] = d;
            else dest.reg[k] = d;
        }

        // 4. Every column through the staging buffer, the key column first:
        // its values are already in registers, which are then free for the
        // other columns' loads.  FULL: a whole tile, no bounds tests.
        auto columns = [&](auto full_tag) {
            constexpr bool FULL = decltype(full_tag)::value;
            if constexpr (Src::kKeyInRegs)
                soa_write_keys<typename Src::KeyT, NW, R, FULL>(v, cols.out[cols.key], slot,
                                                                dest, valid, stage);
            if constexpr (kSelKeep) {
                // The select's columns from registers (a column named twice is
                // written once).
                auto put = [&](uint32_t j, uint32_t w4, const uint32_t (&x)[R]) {
                    if (w4)
                        soa_write_keys<uint32_t, NW, R, FULL>(x, cols.out[j], slot, dest, valid,
                                                              stage);
                    else
                        soa_write_keys<uint8_t, NW, R, FULL>(x, cols.out[j], slot, dest, valid,
                                                             stage);
                };
                put(src.a_col, src.a4, sa);
                if (src.b_col != src.a_col) put(src.b_col, src.b4, sb);
                if (src.sel_col != src.a_col && src.sel_col != src.b_col)
                    put(src.sel_col, src.sel4, sz);
            }
            for (uint32_t j = 0; j < cols.ncols; ++j) {
                if constexpr (Src::kKeyInRegs) {
                    if (j == cols.key) continue;
                }
                if constexpr (kSelKeep) {
                    if (j == src.a_col || j == src.b_col || j == src.sel_col) continue;
                }
                switch (cols.bytes[j]) {
                    case 1:
                        soa_move_column<uint8_t, NW, R, FULL>(cols.in[j], cols.out[j], wbase, n,
                                                              slot, dest, valid, stage, lane);
                        break;
                    case 2:
                        soa_move_column<uint16_t, NW, R, FULL>(cols.in[j], cols.out[j], wbase,
                                                               n, slot, dest, valid, stage,
                                                               lane);
                        break;
                    case 4:
                        soa_move_column<uint32_t, NW, R, FULL>(cols.in[j], cols.out[j], wbase,
                                                               n, slot, dest, valid, stage,
                                                               lane);
                        break;
                    default:
                        if constexpr (sizeof(ST) >= 8)
                            soa_move_column<uint64_t, NW, R, FULL>(cols.in[j], cols.out[j],
                                                                   wbase, n, slot, dest, valid,
                                                                   stage, lane);
                        break;
                }
            }
        };
        // The select source's whole tiles go without bounds tests (2^26
        // (uint, uchar, bool) records, m=32: partition 0.223 -> 0.217 ms, m=64
        // unchanged); for key columns the tested form is faster (m=16: 0.291
        // against 0.308 ms).
        if constexpr (kDestLds) {
            if (full) columns(std::true_type{});
            else columns(std::false_type{});
        } else {
            columns(std::false_type{});
        }
    }
}

}  // namespace mdiscr
