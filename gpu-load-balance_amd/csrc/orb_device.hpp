// orb_device.hpp -- wave64 and block primitives shared by the kernels.
#pragma once

#include <type_traits>

#include "orb_internal.hpp"

namespace orb {

// ---------------------------------------------------------------------------
// wave64 / block primitives
// ---------------------------------------------------------------------------
// A value the same in every lane, kept in a scalar register (the compiler
// cannot tell that a value loaded for a block-uniform index is uniform).
__device__ inline uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline float uniform(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// A compiler-only barrier (no instruction) on four values: the loads that
// produce them complete above it, so a later branch cannot pull them below
// itself (the compiler would otherwise split a record load around a branch
// on one of its fields, or issue it only after an unrelated check: one more
// memory round trip per chunk).
__device__ inline void keep_values(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    asm volatile("" ::"s"(a), "s"(b), "s"(c), "s"(d));
}
__device__ inline void keep_values(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e,
                                   uint32_t f, uint32_t g, uint32_t h)
{
    asm volatile("" ::"s"(a), "s"(b), "s"(c), "s"(d), "s"(e), "s"(f), "s"(g), "s"(h));
}

// The record of chunk q of the level.  A loop over the level's chunks loads
// it before it compares q with the chunk count (so both loads are in flight
// at once) and uses it only when q is below that count; an index past the
// allocation reads its last record instead (no branch around the load).
__device__ inline uint4 chunk_record(const uint4 *chunk_desc, uint32_t q, uint32_t max_chunks)
{
    return chunk_desc[q < max_chunks ? q : max_chunks - 1u];
}

template <class T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <class T>
__device__ inline T wave_inclusive_scan(T v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// f(std::integral_constant<int, I>) for I in [I0, N), unrolled at compile time.
template <int I0, int N, class F>
__device__ inline void static_for(F &&f)
{
    if constexpr (I0 < N) {
        f(std::integral_constant<int, I0>{});
        static_for<I0 + 1, N>(f);
    }
}

// Sum over a 256-thread block; every thread gets the result.
template <class T>
__device__ inline T block_sum(T v, T *smem4)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) smem4[w] = v;
    __syncthreads();
    T r = smem4[0] + smem4[1] + smem4[2] + smem4[3];
    __syncthreads();
    return r;
}

// Exclusive scan over a 256-thread block; *total gets the block sum.
template <class T>
__device__ inline T block_exclusive_scan(T v, T *smem4, T *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T inc = wave_inclusive_scan(v);
    if (lane == 63) smem4[w] = inc;
    __syncthreads();
    T pre = 0;
    for (int i = 0; i < w; ++i) pre += smem4[i];
    *total = smem4[0] + smem4[1] + smem4[2] + smem4[3];
    __syncthreads();
    return pre + inc - v;
}

// Stores the root cell at k == 0 and a zero cell elsewhere, word by word
// (a select between two whole structs would go through scratch memory).
__device__ inline void store_root_or_zero(orb_cell *dst, const orb_cell &root, bool is_root)
{
    uint32_t *w = reinterpret_cast<uint32_t *>(dst);
    const uint32_t *rw = reinterpret_cast<const uint32_t *>(&root);
#pragma unroll
    for (int j = 0; j < (int)(sizeof(orb_cell) / 4); ++j) w[j] = is_root ? rw[j] : 0u;
}

// The share of wave w in a chunk [b, e) whose window keys the second
// histogram round packs: the quarter [1024w, 1024w + 1024) of the chunk
// (offset from b, and length), which holds that wave's keys and its slots.
__device__ inline void wave_share(uint32_t b, uint32_t e, int w, uint32_t &off, uint32_t &cnt)
{
    const uint32_t len = e - b, q = (uint32_t)(kChunk / (kBlock / 64));
    off = (uint32_t)w * q < len ? (uint32_t)w * q : len;
    cnt = len - off < q ? len - off : q;
}

// Records an undefined cell in a build's status pair (the earliest in the
// reference's order wins, whatever the order the checks land in).
__device__ inline void report_error(int32_t *status_pair, int32_t level, int32_t kind, int32_t id)
{
    atomicMax(reinterpret_cast<unsigned long long *>(status_pair),
              (unsigned long long)error_word(level, kind, id));
}

// Per-cell kernels give a wave `group` consecutive cells (a power of two
// <= 64): the wave reads their records in one load, a ballot keeps the
// ones that need work, and the wave works through those one by one.  A level
// of many finished cells then costs few waves.
// (block: this block's index among nblocks blocks that share the work)
template <class Active, class Body>
__device__ inline void for_cell_groups(int32_t nCells, int32_t group, uint32_t block,
                                       uint32_t nblocks, Active active, Body body)
{
    const int lane = threadIdx.x & 63;
    const int32_t wave = (int32_t)((block * blockDim.x + threadIdx.x) >> 6);
    const int32_t nwaves = (int32_t)((nblocks * blockDim.x) >> 6);
    for (int32_t g0 = wave * group; g0 < nCells; g0 += nwaves * group) {
        const int32_t c = g0 + lane;
        uint64_t m = __ballot(lane < group && c < nCells && active(c));
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1ull;
            body(g0 + j);
        }
    }
}
template <class Active, class Body>
__device__ inline void for_cell_groups(int32_t nCells, int32_t group, Active active, Body body)
{
    for_cell_groups(nCells, group, blockIdx.x, gridDim.x, active, body);
}

// chunk q -> its cell (binary search in the exclusive chunk offsets)
__device__ inline uint32_t chunk_to_cell(const uint32_t *first, int32_t nCells, uint32_t q)
{
    int32_t lo = 0, hi = nCells - 1;
    while (lo < hi) {  // last cell with first[cell] <= q
        int32_t mid = (lo + hi + 1) >> 1;
        if (first[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return (uint32_t)lo;
}

// Chunk q's cell and record: first key, one past the last, cell, axis
// (kChunkIdle when the cell has its cut already, belongs to a subtree or has
// no axis).
__device__ inline void fill_chunk(uint32_t q, const uint32_t *first, int32_t nCells,
                                  const CellRec *recs, uint32_t *chunk_cell, uint4 *chunk_desc)
{
    const uint32_t cell = chunk_to_cell(first, nCells, q);
    chunk_cell[q] = cell;
    const CellRec &r = recs[cell];
    const uint32_t b = r.begin + (q - first[cell]) * (uint32_t)kChunk;
    const uint32_t e = r.end - b < (uint32_t)kChunk ? r.end : b + (uint32_t)kChunk;
    const bool live = !r.found && !r.flags && r.axis >= 0 && r.axis <= 2;
    chunk_desc[q] = make_uint4(b, e, cell, live ? (uint32_t)r.axis : kChunkIdle);
}

// Hoare's last step for one cell (partition.cpp:61-62): after the pair
// swaps, rows i and n-1 of the cell [begin, end) trade places (x, y, z and
// the payload arrays that are on).  Nothing happens for an empty cell or one
// the split reported (TL = 0).
__device__ inline void hoare_final_rows(uint32_t begin, uint32_t end, const SplitRec &s, float *x,
                                        float *y, float *z, uint32_t *pid, uint32_t *w)
{
    const uint32_t n = end > begin ? end - begin : 0u;
    if (n == 0 || s.TL == 0) return;  // reported by k_hoare_split
    const uint32_t pa = begin + s.i, pb = begin + n - 1;
    if (pa == pb) return;
    const float ax = x[pa], ay = y[pa], az = z[pa];
    x[pa] = x[pb];
    y[pa] = y[pb];
    z[pa] = z[pb];
    x[pb] = ax;
    y[pb] = ay;
    z[pb] = az;
    if (pid) {
        const uint32_t ia = pid[pa];
        pid[pa] = pid[pb];
        pid[pb] = ia;
    }
    if (w) {
        const uint32_t wa = w[pa];
        w[pa] = w[pb];
        w[pb] = wa;
    }
}

__device__ inline const float *axis_ptr(const float *x, const float *y, const float *z, int a)
{
    return a == 0 ? x : (a == 1 ? y : z);
}

}  // namespace orb
