// soa_quad_kernels.hpp -- the struct-of-arrays partition up to 16 classes in
// quad layout.
//
// Same contract as k_partition_soa (MdiscrHost_tuples.cu.h:262-429, every
// column moved to its record's destination as raw bits), but every column is
// read and written four records at a time:
//   - lane l of a wave holds quad q = records wbase + 256q + 4l + j (j < 4), so
//     a 4-byte column loads as one 16-byte word per quad, a 2-byte column as
//     8 bytes and a 1-byte column as one 4-byte word (byte-wide global
//     instructions cost as much as word-wide ones);
//   - the class source is ranked by quads with packed byte counters (the plain
//     quad partition's quad_rank), so each record's LDS slot is known in the
//     same layout and no slot table is needed;
//   - each class run is staged at an LDS offset that agrees with its output
//     index mod 4 (up to 3 padding slots between runs), so a group of four
//     slots inside one run goes to HBM as one 4-element store; groups at run
//     ends store their valid elements one by one.
// Every column pointer (in and out) must be aligned to 4 elements.
#pragma once

#include "soa_kernels.hpp"

namespace mdiscr {

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Quad q of this lane: elements wbase + 256q + 4 lane + j, zero-extended to
// x[q][j]; elements at or past n read 0 unless `full`.
template <typename T, int Q>
__device__ __forceinline__ void quad_load(const T* __restrict__ col, uint32_t wbase,
                                          uint32_t lane, uint32_t n, bool full,
                                          uint32_t (&x)[Q][4]) {
    if (full) {
        const T* __restrict__ base = col + wbase;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if constexpr (sizeof(T) == 4) {
                const u32x4 w = ld_nt(reinterpret_cast<const u32x4*>(base) + lane + q * kWave);
                x[q][0] = w.x;
                x[q][1] = w.y;
                x[q][2] = w.z;
                x[q][3] = w.w;
            } else if constexpr (sizeof(T) == 2) {
                const u32x2 w = ld_nt(reinterpret_cast<const u32x2*>(base) + lane + q * kWave);
                x[q][0] = w.x & 0xFFFFu;
                x[q][1] = w.x >> 16;
                x[q][2] = w.y & 0xFFFFu;
                x[q][3] = w.y >> 16;
            } else {
                const uint32_t w = ld_nt(reinterpret_cast<const uint32_t*>(base) + lane + q * kWave);
                x[q][0] = w & 0xFFu;
                x[q][1] = (w >> 8) & 0xFFu;
                x[q][2] = (w >> 16) & 0xFFu;
                x[q][3] = w >> 24;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint32_t e = wbase + q * (kWave * 4) + lane * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) x[q][j] = e + j < n ? (uint32_t)col[e + j] : 0u;
        }
    }
}

// A column of 1-, 2- or 4-byte elements given by its width (wave-uniform).
template <int Q>
__device__ __forceinline__ void quad_load_w(const void* col, uint32_t bytes, uint32_t wbase,
                                            uint32_t lane, uint32_t n, bool full,
                                            uint32_t (&x)[Q][4]) {
    if (bytes == 4) quad_load<uint32_t, Q>(static_cast<const uint32_t*>(col), wbase, lane, n, full, x);
    else if (bytes == 2) quad_load<uint16_t, Q>(static_cast<const uint16_t*>(col), wbase, lane, n, full, x);
    else quad_load<uint8_t, Q>(static_cast<const uint8_t*>(col), wbase, lane, n, full, x);
}

// The class source's values of this lane's quads (what the discriminator sees).
template <int Q, class Src>
struct QuadSource;

template <int Q, typename KT>
struct QuadSource<Q, KeySrc<KT>> {
    uint32_t key[Q][4];
    __device__ __forceinline__ void load(const KeySrc<KT>& s, uint32_t wbase, uint32_t lane,
                                         uint32_t n, bool full) {
        quad_load<KT, Q>(s.key, wbase, lane, n, full, key);
    }
    __device__ __forceinline__ uint32_t value(int q, int j) const { return key[q][j]; }
};

template <int Q>
struct QuadSource<Q, SelectSrc> {
    uint32_t z[Q][4], x[Q][4], y[Q][4];
    __device__ __forceinline__ void load(const SelectSrc& s, uint32_t wbase, uint32_t lane,
                                         uint32_t n, bool full) {
        quad_load_w<Q>(s.sel, s.sel4 ? 4u : 1u, wbase, lane, n, full, z);
        quad_load_w<Q>(s.a, s.a4 ? 4u : 1u, wbase, lane, n, full, x);
        quad_load_w<Q>(s.b, s.b4 ? 4u : 1u, wbase, lane, n, full, y);
    }
    __device__ __forceinline__ uint32_t value(int q, int j) const {
        return z[q][j] ? x[q][j] : y[q][j];
    }
};

// Per-tile state the column moves share.
template <int Q, int MC>
struct QuadTile {
    static constexpr int KG = Q + 1;  // 4-slot groups per thread (the last one partial)
    uint32_t slot[Q][4];              // LDS slot of each of this lane's records
    uint32_t gdest[KG];               // output index of group g's first slot, ~0u if the
                                      // group is not four valid slots of one run
};

// Stages this lane's values of one column at their slots.
template <typename T, int Q, int MC>
__device__ __forceinline__ void quad_stage(const QuadTile<Q, MC>& qt, const uint32_t (&x)[Q][4],
                                           uint32_t* stage) {
    T* st = reinterpret_cast<T*>(stage);
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) st[qt.slot[q][j]] = (T)x[q][j];
}

// Writes the staged column: one 4-element store per group of four valid slots
// of one run, then the valid slots of the other groups from the tile's
// boundary list (slot, output index).
template <typename T, int THREADS, int Q, int MC, int GROUPS>
__device__ __forceinline__ void quad_write(const QuadTile<Q, MC>& qt, const uint32_t* stage,
                                           const uint2* blist, uint32_t nb,
                                           void* __restrict__ out_v) {
    T* __restrict__ out = static_cast<T*>(out_v);
    const T* st = reinterpret_cast<const T*>(stage);
#pragma unroll
    for (int k = 0; k < QuadTile<Q, MC>::KG; ++k) {
        const uint32_t g = k * THREADS + threadIdx.x;
        const uint32_t gd = qt.gdest[k];
        if (gd == ~0u) continue;
        if constexpr (sizeof(T) == 4) {
            const u32x4 w = reinterpret_cast<const u32x4*>(st)[g];
            st_out<MDISCR_SOA_NT_STORE>(w, reinterpret_cast<u32x4*>(out + gd));
        } else if constexpr (sizeof(T) == 2) {
            const u32x2 w = reinterpret_cast<const u32x2*>(st)[g];
            st_out<MDISCR_SOA_NT_STORE>(w, reinterpret_cast<u32x2*>(out + gd));
        } else {
            const uint32_t w = reinterpret_cast<const uint32_t*>(st)[g];
            st_out<MDISCR_SOA_NT_STORE>(w, reinterpret_cast<uint32_t*>(out + gd));
        }
    }
    for (uint32_t i = threadIdx.x; i < nb; i += THREADS) {
        const uint2 e = blist[i];
        st_out<MDISCR_SOA_NT_STORE>(st[e.x], out + e.y);
    }
}

template <int THREADS, int Q, int MC, int GROUPS>
__device__ __forceinline__ void quad_write_w(uint32_t bytes, const QuadTile<Q, MC>& qt,
                                             const uint32_t* stage, const uint2* blist,
                                             uint32_t nb, void* out) {
    if (bytes == 4) quad_write<uint32_t, THREADS, Q, MC, GROUPS>(qt, stage, blist, nb, out);
    else if (bytes == 2) quad_write<uint16_t, THREADS, Q, MC, GROUPS>(qt, stage, blist, nb, out);
    else quad_write<uint8_t, THREADS, Q, MC, GROUPS>(qt, stage, blist, nb, out);
}

template <int Q, int MC>
__device__ __forceinline__ void quad_stage_w(uint32_t bytes, const QuadTile<Q, MC>& qt,
                                             const uint32_t (&x)[Q][4], uint32_t* stage) {
    if (bytes == 4) quad_stage<uint32_t>(qt, x, stage);
    else if (bytes == 2) quad_stage<uint16_t>(qt, x, stage);
    else quad_stage<uint8_t>(qt, x, stage);
}

// NW waves x 64 lanes x Q quads = one tile per workgroup.  SINGLE: the records
// are one tile, launched as one workgroup with no count or scan pass; its own
// class counts are the totals, written to user_counts (when given).
template <int NW, int Q, int NBITS, class Src, class Discr, bool SINGLE = false>
__global__ __launch_bounds__(NW* kWave) void k_partition_soa_quad(
    SoaCols cols, Src src, uint32_t n, Discr dis, const uint32_t* __restrict__ prefix,
    const uint32_t* __restrict__ totals, uint32_t ld, uint32_t m,
    uint32_t* __restrict__ user_counts) {
    static_assert(NBITS <= 4, "quad ranking packs at most 16 classes");
    constexpr int MC = 1 << NBITS;
    constexpr int THREADS = NW * kWave;
    constexpr int TILE = THREADS * Q * 4;
    constexpr int GROUPS = TILE / 4 + MC;  // runs padded by at most 3 slots each
    constexpr int SLOTS = 4 * GROUPS;
    static_assert(GROUPS - THREADS * Q <= THREADS, "one partial group round");
    constexpr bool kSel = IsSelect<Src>::value;
    __shared__ __attribute__((aligned(16))) uint32_t stage[SLOTS];
    __shared__ __attribute__((aligned(16))) uint8_t cls[SLOTS];
    __shared__ uint32_t qbase[NW][Q][MC];  // group sizes, then LDS slot bases
    __shared__ uint32_t dpos[MC];          // output index - LDS slot, per class
    __shared__ uint32_t meta[MC];          // run start | valid run length << 16
    // Valid slots of groups that are not four valid slots of one run: such a
    // group holds the first or the last valid slot of a run, so there are at
    // most 2 MC of them, with at most 4 valid slots each.
    constexpr int kBoundMax = 8 * MC;
    __shared__ uint2 blist[kBoundMax];     // (slot, output index)
    __shared__ uint32_t nbound;

    const uint32_t lane = lane_id();
    const uint32_t wave = wave_id();
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tbase = t * (uint32_t)TILE;
    const uint32_t wbase = tbase + wave * (uint32_t)(kWave * 4 * Q);
    const bool full = (uint64_t)tbase + TILE <= n;
    const uint32_t valid = full ? (uint32_t)TILE : n - tbase;

    uint32_t pre = 0, tot = 0;
    if (!SINGLE && wave == 0 && lane < m) {
        pre = prefix[(size_t)lane * ld + t];
        tot = totals[lane];
    }

    // 1. Class source by quads, ranked in its (wave, quad, class) groups.
    QuadSource<Q, Src> qs;
    qs.load(src, wbase, lane, n, full);
    uint32_t ck[Q], rk[Q];
    {
        u32x4 v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] = u32x4{qs.value(q, 0), qs.value(q, 1), qs.value(q, 2),
                                                 qs.value(q, 3)};
        if (full) quad_rank<Q, NBITS, true>(v, ck, rk, qbase[wave], wbase, n, m, lane, dis);
        else quad_rank<Q, NBITS, false>(v, ck, rk, qbase[wave], wbase, n, m, lane, dis);
    }
    __syncthreads();

    // 2. LDS slot bases per (wave, quad, class), padded so that LDS slot and
    // output index agree mod 4; output offsets and valid run lengths per class.
    if (wave == 0) {
        if (lane == 0) nbound = 0;
        uint32_t g[NW * Q];
        uint32_t ct = 0;
        if (lane < (uint32_t)MC) {
#pragma unroll
            for (int i = 0; i < NW * Q; ++i) {
                const uint32_t x = qbase[i / Q][i % Q][lane];
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
        const uint32_t toff = wave_excl_scan(ct, lane);
        const uint32_t base = toff + 3u * lane + ((running - toff - 3u * lane) & 3u);
        const uint32_t rl = ct - ((lane == m - 1u) ? (uint32_t)TILE - valid : 0u);
        if (lane < (uint32_t)MC) {
            meta[lane] = base | (rl << 16);
#pragma unroll
            for (int i = 0; i < NW * Q; ++i) qbase[i / Q][i % Q][lane] = g[i] + base;
            dpos[lane] = running - base;
        }
    }
    __syncthreads();

    // 3. Slots and the class of every slot.
    QuadTile<Q, MC> qt;
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t c = (ck[q] >> (8 * j)) & 0xFFu;
            qt.slot[q][j] = qbase[wave][q][c] + ((rk[q] >> (8 * j)) & 0xFFu);
            cls[qt.slot[q][j]] = (uint8_t)c;
        }

    // 4. Columns, the class source's first (from registers).  Once the class
    // bytes are in LDS (the first staging barrier), each thread forms its
    // groups' destinations, and the valid slots of the other groups (those
    // holding a run's first or last valid slot: at most 2 per class) go to
    // the boundary list.
    if constexpr (kSel) quad_stage_w<Q, MC>(src.a4 ? 4u : 1u, qt, qs.x, stage);
    else quad_stage<typename Src::KeyT>(qt, qs.key, stage);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < QuadTile<Q, MC>::KG; ++k) {
        const uint32_t g = k * THREADS + threadIdx.x;
        uint32_t gd = ~0u;
        if (k < QuadTile<Q, MC>::KG - 1 || g < (uint32_t)GROUPS) {
            const uint32_t cw = reinterpret_cast<const uint32_t*>(cls)[g];
            const uint32_t c = cw & 0xFFu;
            const uint32_t mt = meta[c & (uint32_t)(MC - 1)];
            const uint32_t rs = 4u * g - (mt & 0xFFFFu);
            if (c < (uint32_t)MC && cw == c * 0x01010101u && rs < (mt >> 16) &&
                rs + 3u < (mt >> 16)) {
                gd = dpos[c] + 4u * g;
            } else {
                // Slots outside every run hold stale class bytes: a slot is
                // valid only inside the valid range of the class it names.
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t p = 4u * g + j;
                    const uint32_t cj = (cw >> (8 * j)) & (uint32_t)(MC - 1);
                    const uint32_t mj = meta[cj];
                    if (p - (mj & 0xFFFFu) < (mj >> 16)) {
                        const uint32_t i = atomicAdd(&nbound, 1u);
                        blist[i] = make_uint2(p, dpos[cj] + p);
                    }
                }
            }
        }
        qt.gdest[k] = gd;
    }
    __syncthreads();
    const uint32_t nb = nbound;
    if constexpr (kSel) {
        quad_write_w<THREADS, Q, MC, GROUPS>(src.a4 ? 4u : 1u, qt, stage, blist, nb,
                                             cols.out[src.a_col]);
    } else {
        quad_write<typename Src::KeyT, THREADS, Q, MC, GROUPS>(qt, stage, blist, nb,
                                                               cols.out[cols.key]);
    }
    __syncthreads();
    // One more column: stage, barrier, write, barrier.
    auto put = [&](uint32_t j, uint32_t bytes, const uint32_t (&x)[Q][4]) {
        quad_stage_w<Q, MC>(bytes, qt, x, stage);
        __syncthreads();
        quad_write_w<THREADS, Q, MC, GROUPS>(bytes, qt, stage, blist, nb, cols.out[j]);
        __syncthreads();
    };
    if constexpr (kSel) {
        // a column named twice is written once
        if (src.b_col != src.a_col) put(src.b_col, src.b4 ? 4u : 1u, qs.y);
        if (src.sel_col != src.a_col && src.sel_col != src.b_col)
            put(src.sel_col, src.sel4 ? 4u : 1u, qs.z);
    }
    for (uint32_t j = 0; j < cols.ncols; ++j) {
        if constexpr (kSel) {
            if (j == src.a_col || j == src.b_col || j == src.sel_col) continue;
        } else {
            if (j == cols.key) continue;
        }
        uint32_t x[Q][4];
        quad_load_w<Q>(cols.in[j], cols.bytes[j], wbase, lane, n, full, x);
        put(j, cols.bytes[j], x);
    }
}

}  // namespace mdiscr
