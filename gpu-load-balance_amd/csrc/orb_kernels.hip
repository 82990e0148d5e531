// orb_kernels.hip -- HIP kernels for gfx950 behind the plain (service) form:
// level tables, the segmented count-left pass, the exact Hoare partition as
// scans, the stable <= partition, the axis gather and a device scan.
//
// Layout in HBM: particles are float32 SoA, x | y | z, each array padded to a
// multiple of 64 floats so every array base is 256-byte aligned.  A level is a
// list of cells with disjoint [begin, end) ranges; each cell is cut into
// chunks of kChunk = 4096 keys (16 KiB), a chunk never spans two cells, and
// one 256-thread block (four wave64s) handles one chunk.  The chunk table is
// built on the device from the range map (no host block table).
#include "orb_device.hpp"

namespace orb {

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

// ---------------------------------------------------------------------------
// utilities
// ---------------------------------------------------------------------------
__global__ void k_zero_u32(uint32_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_zero_u64(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

static inline unsigned grid_for(size_t n, unsigned cap = 65536)
{
    size_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

int launch_zero_u32(hipStream_t s, uint32_t *p, size_t n)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

int launch_zero_u64(hipStream_t s, uint64_t *p, size_t n)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

// ---------------------------------------------------------------------------
// exclusive scan (three kernels; m <= kScanTile^2)
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048

// Entries at or past *lim (when given) count as zero and are not read; the
// output is written up to and including index *lim (the grand total).
__device__ inline size_t scan_limit(size_t m, const uint32_t *lim)
{
    if (!lim) return m;
    size_t l = (size_t)(*lim);
    return l < m ? l : m;
}

template <class T>
__global__ void k_scan_reduce(const T *in, size_t m, T *partials, const uint32_t *lim)
{
    __shared__ T sm[4];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    const size_t mr = scan_limit(m, lim);
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < mr) s += in[base + k];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <class T>
__global__ void k_scan_partials(T *partials, int np)
{
    __shared__ T sm[4];
    T v[kScanItems];
    int base = threadIdx.x * kScanItems;
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < np) ? partials[base + k] : T(0);
        s += v[k];
    }
    T total;
    T pre = block_exclusive_scan(s, sm, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < np) partials[base + k] = pre;
        pre += v[k];
    }
}

template <class T>
__global__ void k_scan_apply(const T *in, T *out, size_t m, const T *partials, const uint32_t *lim)
{
    __shared__ T sm[4];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    const size_t mr = scan_limit(m, lim);
    const size_t mw = lim ? (mr + 1 < m ? mr + 1 : m) : m;
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < mr) ? in[base + k] : T(0);
        s += v[k];
    }
    T total;
    T pre = block_exclusive_scan(s, sm, &total) + (partials ? partials[blockIdx.x] : T(0));
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < mw) out[base + k] = pre;
        pre += v[k];
    }
}

template <class T>
static int scan_exclusive(Ctx &c, const T *in, T *out, size_t m, DevBuf<uint64_t> &scratch,
                          const uint32_t *lim = nullptr)
{
    if (m == 0) return ORB_OK;
    size_t np = (m + kScanTile - 1) / kScanTile;
    if (np > (size_t)kScanTile) return ORB_ERR_ARG;
    size_t words = (np * sizeof(T) + sizeof(uint64_t) - 1) / sizeof(uint64_t);
    ORB_TRY(scratch.alloc(words));
    T *partials = reinterpret_cast<T *>(scratch.p);
    if (np == 1) {
        // one tile: the block scan alone (offset 0)
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(kBlock), 0, c.stream, in, out, m,
                           (const T *)nullptr, lim);
        ORB_HIP(hipGetLastError());
        return ORB_OK;
    }
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)np), dim3(kBlock), 0, c.stream, in, m, partials,
                       lim);
    hipLaunchKernelGGL(k_scan_partials<T>, dim3(1), dim3(kBlock), 0, c.stream, partials, (int)np);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)np), dim3(kBlock), 0, c.stream, in, out, m,
                       (const T *)partials, lim);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

unsigned chunk_grid(const Ctx &c)
{
    // chunk kernels loop over the level's chunks: a bounded grid keeps a
    // level with few chunks from paying for a million empty blocks
    int64_t cap = 65536;
    if (c.cur_level >= 0 && c.cur_level < (int32_t)c.tune_chunk_blocks.size() &&
        c.tune_chunk_blocks[(size_t)c.cur_level] > 0)
        cap = c.tune_chunk_blocks[(size_t)c.cur_level];
    return (unsigned)(c.max_chunks < cap ? (c.max_chunks > 0 ? c.max_chunks : 1) : cap);
}

int preallocate_scratch(Ctx &c)
{
    // the largest scan: the chunk counts of a level (max_chunks + 1 entries)
    size_t m = (size_t)c.max_chunks + 1;
    size_t np = (m + kScanTile - 1) / kScanTile;
    ORB_TRY(c.scan_partials.alloc(np));
    if (c.part == ORB_PART_STABLE_LE) ORB_TRY(ensure_stable_buffers(c));
    return ORB_OK;
}

int ensure_stable_buffers(Ctx &c)
{
    // the other buffer has the particle buffer's shape, slack included
    ORB_TRY(c.tmp.alloc(3 * (size_t)c.stride + 64));
    if (c.track_ids) ORB_TRY(c.pid_tmp.alloc((size_t)c.stride));
    return ORB_OK;
}

__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t n4 = n / 4;
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    for (size_t k = i; k < n4; k += stride) d4[k] = s4[k];
    for (size_t k = 4 * n4 + i; k < n; k += stride) dst[k] = src[k];
}

// src and dst 16-byte aligned (allocation bases)
int launch_copy_u32(hipStream_t s, const uint32_t *src, uint32_t *dst, size_t n)
{
    if (n == 0) return ORB_OK;
    size_t g = (n / 4 + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)g), dim3(kBlock), 0, s, src, dst, n);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

__global__ void k_iota(uint32_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = (uint32_t)k;
}

int launch_iota(hipStream_t s, uint32_t *p, size_t n)
{
    if (n == 0) return ORB_OK;
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)g), dim3(kBlock), 0, s, p, n);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

// Brings the particles (and ids) back to the primary buffer after a stable
// build that ended in the other one: one copy per build instead of one per
// level.
int particles_home(Ctx &c)
{
    if (!c.in_tmp) return ORB_OK;
    ORB_TRY(launch_copy_u32(c.stream, reinterpret_cast<const uint32_t *>(c.tmp.p),
                            reinterpret_cast<uint32_t *>(c.pos.p), 3 * (size_t)c.stride));
    if (c.track_ids) ORB_TRY(launch_copy_u32(c.stream, c.pid_tmp.p, c.pid.p, (size_t)c.stride));
    c.in_tmp = false;
    return ORB_OK;
}

int scan_exclusive_u32(Ctx &c, const uint32_t *in, uint32_t *out, size_t m)
{
    return scan_exclusive<uint32_t>(c, in, out, m, c.scan_partials);
}

int scan_exclusive_u64(Ctx &c, const uint64_t *in, uint64_t *out, size_t m, const uint32_t *lim)
{
    return scan_exclusive<uint64_t>(c, in, out, m, c.scan_partials, lim);
}

// ---------------------------------------------------------------------------
// level records and the chunk table (ServiceCopyCells, copyCells.cu:10-64)
// ---------------------------------------------------------------------------
__global__ void k_level_recs(const orb_cell *cells, int32_t nCells, const uint32_t *ranges,
                             CellRec *recs, uint32_t *nchunks, int32_t cap)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nCells) return;
    if (i == nCells) {
        nchunks[nCells] = 0u;
        return;
    }
    const orb_cell &cl = cells[i];
    CellRec r;
    r.begin = ranges[2 * (size_t)cl.id];
    r.end = ranges[2 * (size_t)cl.id + 1];
    if (r.end < r.begin) r.end = r.begin;
    r.cut = cell_get_cut(cl);
    r.axis = cl.cutAxis;
    r.found = cl.foundCut ? 1 : 0;
    r.nLeaf = cl.nLeafCells;
    r.margin_l = cl.cutMarginLeft;
    r.margin_r = cl.cutMarginRight;
    r.id = cl.id;
    r.flags = 0;
    r.passes = 0;
    r.hb = 0;
    r.capr = cap;
    r.lp = 0;
    r.st = 0;
    r.lmax = 0;
    recs[i] = r;
    nchunks[i] = (r.end - r.begin + (uint32_t)kChunk - 1u) / (uint32_t)kChunk;
}

__global__ void k_fill_chunks(const uint32_t *first, int32_t nCells, uint32_t *chunk_cell,
                              uint32_t max_chunks)
{
    uint32_t total = first[nCells];
    if (total > max_chunks) total = max_chunks;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += gridDim.x * blockDim.x)
        chunk_cell[q] = chunk_to_cell(first, nCells, q);
}

int ensure_level_capacity(Ctx &c, int64_t nCells)
{
    if (nCells <= c.max_level_cells && c.recs.p) return ORB_OK;
    int64_t m = nCells > c.max_level_cells ? nCells : c.max_level_cells;
    c.max_level_cells = m;
    c.max_chunks = c.n / kChunk + m + 1;
    ORB_TRY(c.dcells.alloc(m));
    ORB_TRY(c.recs.alloc(m));
    ORB_TRY(c.cell_nchunks.alloc(m + 1));
    ORB_TRY(c.cell_first.alloc(m + 1));
    ORB_TRY(c.chunk_cell.alloc(c.max_chunks));
    ORB_TRY(c.cnt.alloc(2 * m));
    ORB_TRY(c.stamps.alloc(m));
    ORB_TRY(c.chunk_ltle.alloc(c.max_chunks + 1));
    ORB_TRY(c.chunk_cnt.alloc(c.max_chunks + 1));
    ORB_TRY(c.chunk_pre.alloc(c.max_chunks + 1));
    ORB_TRY(c.splits.alloc(m));
    ORB_TRY(c.chunk_comp.alloc(8 * ((size_t)c.max_chunks + 1)));
    ORB_TRY(c.ids_dev.alloc(m));
    ORB_TRY(c.hcells.alloc(m));
    ORB_TRY(c.hcnt.alloc(2 * m));
    return ORB_OK;
}

int upload_level(Ctx &c, const orb_cell *cells, int32_t nCells)
{
    ORB_TRY(ensure_level_capacity(c, nCells));
    // the previous upload may still be reading the pinned staging buffer
    if (c.ev_upload) ORB_HIP(hipEventSynchronize(c.ev_upload));
    std::memcpy(c.hcells.p, cells, sizeof(orb_cell) * (size_t)nCells);
    ORB_HIP(hipMemcpyAsync(c.dcells.p, c.hcells.p, sizeof(orb_cell) * (size_t)nCells,
                           hipMemcpyHostToDevice, c.stream));
    if (c.ev_upload) ORB_HIP(hipEventRecord(c.ev_upload, c.stream));
    unsigned g = (unsigned)((nCells + 1 + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_level_recs, dim3(g), dim3(kBlock), 0, c.stream, c.dcells.p, nCells,
                       c.ranges.p, c.recs.p, c.cell_nchunks.p, (int32_t)c.cap);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

int invalidate_stamps(Ctx &c, int32_t nCells);

// with_stamps: the plain form reuses a count pass's chunk counts for the
// partition when the stamp matches; the batched form takes them from its
// histograms and needs no stamps.
int build_chunk_table(Ctx &c, int32_t nCells, bool with_stamps)
{
    ORB_TRY(scan_exclusive_u32(c, c.cell_nchunks.p, c.cell_first.p, (size_t)nCells + 1));
    hipLaunchKernelGGL(k_fill_chunks, dim3(grid_for((size_t)c.max_chunks, 16384)), dim3(kBlock), 0,
                       c.stream, c.cell_first.p, nCells, c.chunk_cell.p, (uint32_t)c.max_chunks);
    ORB_HIP(hipGetLastError());
    return with_stamps ? invalidate_stamps(c, nCells) : ORB_OK;
}

// chunk q of the level: its cell and [b, e).  Returns false past the table.
__device__ inline bool chunk_range(uint32_t q, const uint32_t *first, int32_t nCells,
                                   const uint32_t *chunk_cell, const CellRec *recs,
                                   uint32_t &cell, uint32_t &b, uint32_t &e)
{
    uint32_t total = first[nCells];
    if (q >= total) return false;
    cell = chunk_cell[q];
    const CellRec &r = recs[cell];
    b = r.begin + (q - first[cell]) * (uint32_t)kChunk;
    e = b + (uint32_t)kChunk;
    if (e > r.end) e = r.end;
    return true;
}

// ---------------------------------------------------------------------------
// count-left pass (ServiceCountLeft / ServiceCountLeftGPU / ...GPUAxis)
//
// One block per chunk streams the chunk's keys once with 16-byte loads (head
// and tail of a misaligned chunk scalar) and counts x < cut and x <= cut
// together.  A wave64 shuffle reduction and four LDS partials give the block
// sum, written to the chunk's slot: no atomics (a same-address atomic per
// block serialises at L2 and costs more than the stream itself when a level
// has few cells).  k_cell_totals then sums each cell's chunk slots.
// Roofline: HBM, 4 bytes per key counted.
// ---------------------------------------------------------------------------
__device__ inline void count_range(const float *key, uint32_t b, uint32_t e, float cut,
                                   uint32_t &lt, uint32_t &le)
{
    uint32_t ab = (b + 3u) & ~3u;
    if (ab > e) ab = e;
    uint32_t ae = e & ~3u;
    if (ae < ab) ae = ab;
    const uint32_t t = threadIdx.x;
    if (t < ab - b) {
        float v = key[b + t];
        lt += v < cut;
        le += v <= cut;
    }
    if (t < e - ae) {
        float v = key[ae + t];
        lt += v < cut;
        le += v <= cut;
    }
    const float4 *k4 = reinterpret_cast<const float4 *>(key + ab);
    const uint32_t nq = (ae - ab) >> 2;  // <= kChunk/4 = 1024
    float4 v[kPerThread / 4];
#pragma unroll
    for (int u = 0; u < kPerThread / 4; ++u) {
        uint32_t j = t + (uint32_t)u * kBlock;
        if (j < nq) v[u] = k4[j];
        else v[u] = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
    }
#pragma unroll
    for (int u = 0; u < kPerThread / 4; ++u) {
        lt += (v[u].x < cut) + (v[u].y < cut) + (v[u].z < cut) + (v[u].w < cut);
        le += (v[u].x <= cut) + (v[u].y <= cut) + (v[u].z <= cut) + (v[u].w <= cut);
    }
}

__device__ inline void k_count_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                                  const uint32_t *first, const uint32_t *chunk_cell,
                                                  const float *x, const float *y, const float *z,
                                                  const float *axisT, int skip_found,
                                                  uint32_t *chunk_ltle)
{
    __shared__ uint32_t sm[4];
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    const CellRec r = recs[cell];
    if (skip_found && r.found) return;
    if (r.axis < 0 || r.axis > 2) return;
    const float *key = axisT ? axisT : axis_ptr(x, y, z, r.axis);
    uint32_t lt = 0, le = 0;
    count_range(key, b, e, r.cut, lt, le);
    // both counts <= 4096: one packed word, one reduction
    uint32_t packed = block_sum(lt | (le << 16), sm);
    if (threadIdx.x == 0) chunk_ltle[q] = packed;
}

__global__ __launch_bounds__(kBlock) void k_count(const CellRec *recs, int32_t nCells,
                                                  const uint32_t *first, const uint32_t *chunk_cell,
                                                  const float *x, const float *y, const float *z,
                                                  const float *axisT, int skip_found,
                                                  uint32_t *chunk_ltle)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_count_chunk(q, recs, nCells, first, chunk_cell, x, y, z, axisT, skip_found, chunk_ltle);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}


// Per-cell sums of the chunk slots; stamps the cut and axis the counts were
// taken at (a later partition at the same cut reuses the chunk slots).  One
// block per cell; 16-byte loads with eight in flight per thread keep the
// root cell's 32K slots to a few load round trips.
__global__ __launch_bounds__(kBlock) void k_cell_totals(const CellRec *recs, int32_t nCells,
                                                        const uint32_t *first,
                                                        const uint32_t *chunk_ltle,
                                                        int skip_found, int stamp,
                                                        uint32_t *cnt, CountStamp *stamps)
{
    __shared__ uint32_t sm[4];
    for (int32_t c = blockIdx.x; c < nCells; c += gridDim.x) {
        const CellRec r = recs[c];
        if (skip_found && r.found) continue;  // uniform per block
        const uint32_t q0 = first[c], q1 = first[c + 1];
        uint32_t lt = 0, le = 0;
        uint32_t a0 = (q0 + 3u) & ~3u;
        if (a0 > q1) a0 = q1;
        uint32_t a1 = q1 & ~3u;
        if (a1 < a0) a1 = a0;
        const uint32_t t = threadIdx.x;
        if (t < a0 - q0) {
            uint32_t p = chunk_ltle[q0 + t];
            lt += p & 0xffffu;
            le += p >> 16;
        }
        if (t < q1 - a1) {
            uint32_t p = chunk_ltle[a1 + t];
            lt += p & 0xffffu;
            le += p >> 16;
        }
        const uint4 *v4 = reinterpret_cast<const uint4 *>(chunk_ltle + a0);
        const uint32_t nv = (a1 - a0) >> 2;
        for (uint32_t base = 0; base < nv; base += 8u * kBlock) {
            uint4 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                uint32_t j = base + t + (uint32_t)u * kBlock;
                v[u] = j < nv ? v4[j] : make_uint4(0u, 0u, 0u, 0u);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                lt += (v[u].x & 0xffffu) + (v[u].y & 0xffffu) + (v[u].z & 0xffffu) + (v[u].w & 0xffffu);
                le += (v[u].x >> 16) + (v[u].y >> 16) + (v[u].z >> 16) + (v[u].w >> 16);
            }
        }
        lt = block_sum(lt, sm);
        le = block_sum(le, sm);
        if (t == 0) {
            cnt[2 * c] = lt;
            cnt[2 * c + 1] = le;
            CountStamp st;
            st.cut = r.cut;
            st.axis = stamp ? r.axis : -1;
            stamps[c] = st;
        }
    }
}

__global__ void k_invalidate_stamps(CountStamp *stamps, int32_t n)
{
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) stamps[i] = CountStamp{0.0f, -1};
}

int invalidate_stamps(Ctx &c, int32_t nCells)
{
    if (nCells <= 0) return ORB_OK;
    hipLaunchKernelGGL(k_invalidate_stamps, dim3((nCells + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       c.stream, c.stamps.p, nCells);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

int launch_count(Ctx &c, int32_t nCells, bool skip_found, const float *axis_override,
                 hipEvent_t after_count)
{
    hipLaunchKernelGGL(k_count, dim3(chunk_grid(c)), dim3(kBlock), 0, c.stream,
                       c.recs.p, nCells, c.cell_first.p, c.chunk_cell.p, c.x(), c.y(), c.z(),
                       axis_override, skip_found ? 1 : 0, c.chunk_ltle.p);
    ORB_HIP(hipGetLastError());
    if (after_count) ORB_HIP(hipEventRecord(after_count, c.stream));
    unsigned g = (unsigned)(nCells < 65536 ? nCells : 65536);
    hipLaunchKernelGGL(k_cell_totals, dim3(g), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_ltle.p, skip_found ? 1 : 0,
                       axis_override ? 0 : 1, c.cnt.p, c.stamps.p);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

// ---------------------------------------------------------------------------
// Hoare partition as scans (partition.cpp:28-60).
//
// For a cell of n keys and cut c let the L-stoppers be the positions with
// key >= c, left to right L_0 < L_1 < ..., and the R-stoppers the positions
// with key <= c, right to left R_0 > R_1 > ....  The serial loop swaps
// (L_t, R_t) for every t < K = #{t : L_t < R_t} (these transpositions are
// disjoint), stops with i = min(L_K, R_{K-1}) (i = L_0 when K == 0), and then
// swaps rows i and n-1.  The children are [0, i) and [i, n).
//
//   P1a  per chunk: #L and #R; taken from the last count pass when the cell
//        was counted at this cut (stamp), else one key read
//   scan per-chunk counts -> each chunk's rank offsets inside its cell
//   P1b  per chunk: block scan of the flags, stopper positions staged in LDS
//        and written contiguously to lpos / rpos at the cell's base + rank
//   P2   per cell (one wave): K by a 64-way search on lpos[t] < rpos[TR-1-t],
//        i, the children's ranges, undefined-input check
//   P3   per pair t < K: swap x, y, z of lpos[t] and rpos[TR-1-t]
//   P4   per cell: swap rows i and n-1
// ---------------------------------------------------------------------------
__device__ inline void k_stopper_counts_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                                           const uint32_t *first,
                                                           const uint32_t *chunk_cell,
                                                           const float *x, const float *y,
                                                           const float *z, int le_only,
                                                           const CountStamp *stamps,
                                                           const uint32_t *chunk_ltle,
                                                           uint64_t *chunk_cnt, uint32_t max_chunks)
{
    __shared__ uint32_t sm[4];
    uint32_t cell, b, e;
    (void)max_chunks;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    const CellRec r = recs[cell];
    const int axis = r.axis < 0 ? 0 : (r.axis > 2 ? 2 : r.axis);
    const CountStamp st = stamps[cell];
    uint32_t lt, le;
    if (st.axis == r.axis && __float_as_uint(st.cut) == __float_as_uint(r.cut)) {
        uint32_t p = chunk_ltle[q];  // counted at this cut by the last pass
        lt = p & 0xffffu;
        le = p >> 16;
    } else {
        lt = 0;
        le = 0;
        count_range(axis_ptr(x, y, z, axis), b, e, r.cut, lt, le);
        uint32_t packed = block_sum(lt | (le << 16), sm);
        lt = packed & 0xffffu;
        le = packed >> 16;
    }
    if (threadIdx.x == 0) {
        uint32_t len = e - b;
        uint32_t nl = le_only ? le : len - lt;        // <=-partition: keys going left
        uint32_t nr = le_only ? len - le : le;        // ... keys going right
        chunk_cnt[q] = (uint64_t)nl | ((uint64_t)nr << 32);
    }
}

__global__ __launch_bounds__(kBlock) void k_stopper_counts(const CellRec *recs, int32_t nCells,
                                                           const uint32_t *first,
                                                           const uint32_t *chunk_cell,
                                                           const float *x, const float *y,
                                                           const float *z, int le_only,
                                                           const CountStamp *stamps,
                                                           const uint32_t *chunk_ltle,
                                                           uint64_t *chunk_cnt, uint32_t max_chunks)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_stopper_counts_chunk(q, recs, nCells, first, chunk_cell, x, y, z, le_only, stamps, chunk_ltle, chunk_cnt, max_chunks);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}


// Each wave owns a contiguous quarter of the chunk (1024 keys, 16 per lane in
// registers, one coalesced dword load each).  Pass 1 counts its stoppers with
// ballots; the four wave counts give each wave its rank offset; pass 2 writes
// each stopper's position at its rank: lanes of one ballot write consecutive
// entries, so the stores stay contiguous without LDS staging.
__device__ inline void k_stopper_positions_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                                 const uint32_t *first,
                                                 const uint32_t *chunk_cell,
                                                 const uint64_t *chunk_pre,
                                                 const float *x, const float *y,
                                                 const float *z, uint32_t *lpos,
                                                 uint32_t *rpos)
{
    constexpr int kSeg = kChunk / (kBlock / 64);  // 1024 keys per wave
    constexpr int kIt = kSeg / 64;                 // 16 per lane
    __shared__ uint32_t s_cnt[kBlock / 64];
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    const CellRec r = recs[cell];
    const float cut = r.cut;
    const float *key = axis_ptr(x, y, z, r.axis);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t sb = b + (uint32_t)(w * kSeg);
    float v[kIt];
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
        const uint32_t p = sb + (uint32_t)(64 * j + lane);
        v[j] = p < e ? key[p] : 0.0f;
    }
    uint32_t nl = 0, nr = 0;
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
        const bool in = sb + (uint32_t)(64 * j + lane) < e;
        nl += (uint32_t)__popcll(__ballot(in && v[j] >= cut));
        nr += (uint32_t)__popcll(__ballot(in && v[j] <= cut));
    }
    if (lane == 0) s_cnt[w] = nl | (nr << 16);  // each <= 1024
    __syncthreads();
    uint32_t ol = 0, orr = 0;
    for (int u = 0; u < w; ++u) {
        ol += s_cnt[u] & 0xffffu;
        orr += s_cnt[u] >> 16;
    }
    const uint64_t base = chunk_pre[q] - chunk_pre[first[cell]];
    uint32_t *dl = lpos + r.begin + (uint32_t)(base & 0xffffffffull) + ol;
    uint32_t *dr = rpos + r.begin + (uint32_t)(base >> 32) + orr;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t off = sb - r.begin;  // cell-local position of this wave's first key
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
        const uint32_t k = (uint32_t)(64 * j + lane);
        const bool in = sb + k < e;
        const bool isL = in && v[j] >= cut, isR = in && v[j] <= cut;
        const uint64_t mL = __ballot(isL), mR = __ballot(isR);
        if (isL) dl[__popcll(mL & below)] = off + k;
        if (isR) dr[__popcll(mR & below)] = off + k;
        dl += __popcll(mL);
        dr += __popcll(mR);
    }
}

__global__ __launch_bounds__(kBlock) void k_stopper_positions(const CellRec *recs, int32_t nCells,
                                                              const uint32_t *first,
                                                              const uint32_t *chunk_cell,
                                                              const uint64_t *chunk_pre,
                                                              const float *x, const float *y,
                                                              const float *z, uint32_t *lpos,
                                                              uint32_t *rpos)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_stopper_positions_chunk(q, recs, nCells, first, chunk_cell, chunk_pre, x, y, z, lpos, rpos);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}


__device__ inline void report_error(int32_t *status, int32_t id)
{
    if (atomicCAS(&status[0], 0, 1) == 0) status[1] = id;
}

// one wave64 per cell
__global__ __launch_bounds__(kBlock) void k_hoare_split(const CellRec *recs, int32_t nCells,
                                                        const uint32_t *first,
                                                        const uint64_t *chunk_pre,
                                                        const uint32_t *lpos, const uint32_t *rpos,
                                                        SplitRec *splits, uint32_t *ranges,
                                                        int32_t *status, int64_t cap)
{
    const int lane = threadIdx.x & 63;
    const int32_t c = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (c >= nCells) return;  // uniform per wave
    const CellRec r = recs[c];
    if (r.flags) {  // handled elsewhere (batched form's subtree kernel)
        if (lane == 0) splits[c] = SplitRec{0u, 0u, 0u, 0u};
        return;
    }
    const uint64_t tot = chunk_pre[first[c + 1]] - chunk_pre[first[c]];
    const uint32_t TL = (uint32_t)(tot & 0xffffffffull), TR = (uint32_t)(tot >> 32);
    const uint32_t n = r.end - r.begin;
    SplitRec s{TL, TR, 0u, 0u};
    (void)cap;
    const int64_t lid = left_child_id(r.id), rid = right_child_id(r.id);
    if (n == 0 || TL == 0 || r.axis < 0 || r.axis > 2 || rid >= r.capr) {
        // empty cell or no key >= cut: the reference's loop leaves the cell
        if (lane == 0) {
            report_error(status + r.st, r.id);
            s.TL = 0u;  // k_hoare_final skips the cell
            splits[c] = s;
        }
        return;
    }
    const uint32_t *L = lpos + r.begin;
    const uint32_t *R = rpos + r.begin;  // left-to-right order; R_t = R[TR-1-t]
    // K = #{t : L[t] < R[TR-1-t]}; the predicate is true then false in t
    uint32_t lo = 0, hi = TL < TR ? TL : TR;
    while (hi - lo > 64u) {
        const uint32_t span = hi - lo;
        const uint32_t t = lo + (uint32_t)(((uint64_t)(lane + 1) * span) / 65u);
        const bool pr = L[t] < R[TR - 1 - t];
        const uint64_t m = __ballot(pr);
        if (m == 0ull) {
            hi = lo + (uint32_t)(((uint64_t)1 * span) / 65u);           // pred(t_0) false
        } else {
            const int j = 63 - __clzll((long long)m);                    // last true probe
            const uint32_t tj = lo + (uint32_t)(((uint64_t)(j + 1) * span) / 65u);
            const uint32_t tn = j < 63 ? lo + (uint32_t)(((uint64_t)(j + 2) * span) / 65u) : hi;
            lo = tj + 1;
            hi = tn;
        }
    }
    const uint32_t t = lo + (uint32_t)lane;
    const bool pr = t < hi && L[t] < R[TR - 1 - t];
    const uint32_t K = lo + (uint32_t)__popcll(__ballot(pr));
    if (lane != 0) return;
    uint32_t i;
    if (K == 0) {
        i = L[0];
    } else {
        i = R[TR - K];
        if (K < TL && L[K] < i) i = L[K];
    }
    s.K = K;
    s.i = i;
    splits[c] = s;
    const int64_t gl = r.hb + lid, gr = r.hb + rid;
    ranges[2 * gl] = r.begin;
    ranges[2 * gl + 1] = r.begin + i;
    ranges[2 * gr] = r.begin + i;
    ranges[2 * gr + 1] = r.end;
}

// Pairs in batches of kSwapBatch per thread: all index loads of a batch, then
// all six-array gathers, then the stores, so each thread keeps several HBM
// round trips in flight instead of one dependent chain per pair.
constexpr int kSwapBatch = 8;

__device__ inline void k_hoare_swap_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                          const uint32_t *first, const uint32_t *chunk_cell,
                                          const SplitRec *splits, const uint32_t *lpos,
                                          const uint32_t *rpos, float *x, float *y, float *z,
                                          uint32_t *pid)
{
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    const CellRec r = recs[cell];
    const SplitRec s = splits[cell];
    uint32_t t0 = b - r.begin;  // this chunk handles pair indices [t0, t0 + len)
    uint32_t t1 = e - r.begin;
    if (t1 > s.K) t1 = s.K;
    if (t0 >= t1) return;
    const uint32_t *L = lpos + r.begin;
    const uint32_t *R = rpos + r.begin;
    for (uint32_t t = t0 + threadIdx.x; t < t1; t += kBlock * kSwapBatch) {
        uint32_t pa[kSwapBatch], pb[kSwapBatch];
#pragma unroll
        for (int u = 0; u < kSwapBatch; ++u) {
            const uint32_t tt = t + (uint32_t)(u * kBlock);
            pa[u] = tt < t1 ? r.begin + L[tt] : 0u;
            pb[u] = tt < t1 ? r.begin + R[s.TR - 1 - tt] : 0u;
        }
        float ax[kSwapBatch], ay[kSwapBatch], az[kSwapBatch];
        float bx[kSwapBatch], by[kSwapBatch], bz[kSwapBatch];
#pragma unroll
        for (int u = 0; u < kSwapBatch; ++u) {
            if (t + (uint32_t)(u * kBlock) < t1) {
                ax[u] = x[pa[u]];
                ay[u] = y[pa[u]];
                az[u] = z[pa[u]];
                bx[u] = x[pb[u]];
                by[u] = y[pb[u]];
                bz[u] = z[pb[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kSwapBatch; ++u) {
            if (t + (uint32_t)(u * kBlock) < t1) {
                x[pa[u]] = bx[u];
                y[pa[u]] = by[u];
                z[pa[u]] = bz[u];
                x[pb[u]] = ax[u];
                y[pb[u]] = ay[u];
                z[pb[u]] = az[u];
            }
        }
        if (pid) {  // uniform
            uint32_t ia[kSwapBatch], ib[kSwapBatch];
#pragma unroll
            for (int u = 0; u < kSwapBatch; ++u) {
                if (t + (uint32_t)(u * kBlock) < t1) {
                    ia[u] = pid[pa[u]];
                    ib[u] = pid[pb[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < kSwapBatch; ++u) {
                if (t + (uint32_t)(u * kBlock) < t1) {
                    pid[pa[u]] = ib[u];
                    pid[pb[u]] = ia[u];
                }
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_hoare_swap(const CellRec *recs, int32_t nCells,
                                                       const uint32_t *first,
                                                       const uint32_t *chunk_cell,
                                                       const SplitRec *splits,
                                                       const uint32_t *lpos, const uint32_t *rpos,
                                                       float *x, float *y, float *z,
                                                       uint32_t *pid)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_hoare_swap_chunk(q, recs, nCells, first, chunk_cell, splits, lpos, rpos, x, y, z, pid);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}


__global__ void k_hoare_final(const CellRec *recs, int32_t nCells, const SplitRec *splits,
                              float *x, float *y, float *z, uint32_t *pid)
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nCells) return;
    const CellRec r = recs[c];
    const SplitRec s = splits[c];
    uint32_t n = r.end - r.begin;
    if (n == 0 || s.TL == 0) return;  // reported by k_hoare_split
    uint32_t pa = r.begin + s.i, pb = r.begin + n - 1;
    if (pa == pb) return;
    float ax = x[pa], ay = y[pa], az = z[pa];
    x[pa] = x[pb]; y[pa] = y[pb]; z[pa] = z[pb];
    x[pb] = ax; y[pb] = ay; z[pb] = az;
    if (pid) {
        const uint32_t ia = pid[pa];
        pid[pa] = pid[pb];
        pid[pb] = ia;
    }
}

// Stopper counts of a chunk from the histogram round that found its cell's
// cut at candidate j: keys below the cut = the bins below j (< or <=
// histogram), and the keys equal to candidate j turn one count into the
// other.  One wave per chunk: the row is 512 bytes, so the work is the chain
// of loads that find it, and four independent waves per block hide it.  A
// cell that hit the pass cap was cut at a midpoint no round counted: its
// chunks read their keys (the whole wave).
__global__ __launch_bounds__(kBlock) void k_b_stopper_counts(
    const CellRec *recs, int32_t nCells, const uint32_t *first, const uint32_t *chunk_cell,
    const float *x, const float *y, const float *z, int le_only, const uint16_t *hist,
    const uint16_t *tie, int hist_le, const float *pack, const uint32_t *comp,
    uint64_t *chunk_cnt)
{
    constexpr int kW = kBlock / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x * kW + (uint32_t)w; q < total; q += gridDim.x * kW) {
        const uint32_t cell = chunk_cell[q];
        const CellRec r = recs[cell];
        const uint32_t b = r.begin + (q - first[cell]) * (uint32_t)kChunk;
        uint32_t e = b + (uint32_t)kChunk;
        if (e > r.end) e = r.end;
        uint32_t lt = 0, le = 0;
        if (r.found >= 2) {
            const int j = r.found - 1;
            const uint16_t *row = hist + (size_t)q * kBins;
            uint32_t s = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int bin = 4 * lane + i;
                if (bin < j) s += row[bin];
            }
            const uint32_t below = wave_sum(s);
            const uint32_t eq = tie[(size_t)q * kBins + (uint32_t)j];
            if (hist_le) {
                le = below;
                lt = below - eq;
            } else {
                lt = below;
                le = below + eq;
            }
        } else if (r.passes > kHistK) {
            // the cell went through the packing round: its keys outside the
            // packing window lie below or above every later cut, so the
            // packed keys and the count below decide
            const float cut = r.cut;
            for (int sh = 0; sh < kW; ++sh) {
                uint32_t off, cnt;
                wave_share(b, e, sh, off, cnt);
                const uint32_t nc = comp[8 * (size_t)q + sh];
                for (uint32_t k = (uint32_t)lane; k < nc; k += 64) {
                    const float v = pack[b + off + k];
                    lt += v < cut ? 1u : 0u;
                    le += v <= cut ? 1u : 0u;
                }
            }
            const uint32_t bel = comp[8 * (size_t)q + 4];
            lt = wave_sum(lt) + bel;
            le = wave_sum(le) + bel;
        } else {
            const int axis = r.axis < 0 ? 0 : (r.axis > 2 ? 2 : r.axis);
            const float *key = axis_ptr(x, y, z, axis);
            const float cut = r.cut;
            for (uint32_t p = b + (uint32_t)lane; p < e; p += 64) {
                const float v = key[p];
                lt += v < cut ? 1u : 0u;
                le += v <= cut ? 1u : 0u;
            }
            lt = wave_sum(lt);
            le = wave_sum(le);
        }
        if (lane == 0) {
            const uint32_t len = e - b;
            const uint32_t nl = le_only ? le : len - lt;
            const uint32_t nr = le_only ? len - le : le;
            chunk_cnt[q] = (uint64_t)nl | ((uint64_t)nr << 32);
        }
    }
}

int launch_stopper_counts(Ctx &c, int32_t nCells, bool le_only, const HistCounts *hc)
{
    if (hc && hc->hist)
        hipLaunchKernelGGL(k_b_stopper_counts, dim3((chunk_grid(c) + 3) / 4), dim3(kBlock), 0, c.stream,
                           c.recs.p, nCells, c.cell_first.p, c.chunk_cell.p, c.x(), c.y(), c.z(),
                           le_only ? 1 : 0, hc->hist, hc->tie, hc->le, hc->pack, hc->comp,
                           c.chunk_cnt.p);
    else
        hipLaunchKernelGGL(k_stopper_counts, dim3(chunk_grid(c)), dim3(kBlock), 0,
                           c.stream, c.recs.p, nCells, c.cell_first.p, c.chunk_cell.p, c.x(), c.y(),
                           c.z(), le_only ? 1 : 0, c.stamps.p, c.chunk_ltle.p, c.chunk_cnt.p,
                           (uint32_t)c.max_chunks);
    ORB_HIP(hipGetLastError());
    return scan_exclusive_u64(c, c.chunk_cnt.p, c.chunk_pre.p, (size_t)c.max_chunks + 1,
                              c.cell_first.p + nCells);
}

int launch_hoare(Ctx &c, int32_t nCells, int32_t *status, const HistCounts *hc)
{
    if (!status) status = c.status.p;
    const unsigned gq = chunk_grid(c);
    ORB_TRY(launch_stopper_counts(c, nCells, false, hc));
    hipLaunchKernelGGL(k_stopper_positions, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_cell.p, c.chunk_pre.p, c.x(), c.y(), c.z(), c.lpos.p,
                       c.rpos.p);
    ORB_HIP(hipGetLastError());
    unsigned gw = (unsigned)((nCells + 3) / 4);
    hipLaunchKernelGGL(k_hoare_split, dim3(gw), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_pre.p, c.lpos.p, c.rpos.p, c.splits.p, c.ranges.p,
                       status, (int64_t)c.cap);
    ORB_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_hoare_swap, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_cell.p, c.splits.p, c.lpos.p, c.rpos.p, c.x(), c.y(),
                       c.z(), c.ids());
    ORB_HIP(hipGetLastError());
    unsigned gc = (unsigned)((nCells + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_hoare_final, dim3(gc), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.splits.p, c.x(), c.y(), c.z(), c.ids());
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

// ---------------------------------------------------------------------------
// stable <= partition (ServicePartitionGPU semantics, deterministic order)
// ---------------------------------------------------------------------------
// Keys <= cut first, then the rest, each in input order.  Each wave owns a
// contiguous quarter of the chunk with 16 keys per lane in registers; pass 1
// counts its keys <= cut with ballots, the four wave counts give the wave
// offsets, pass 2 writes each particle at its rank: the lanes of one ballot
// write consecutive slots (mbcnt), so the x, y, z (and id) stores coalesce.
__device__ inline void k_stable_scatter_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                              const uint32_t *first, const uint32_t *chunk_cell,
                                              const uint64_t *chunk_pre, const float *x,
                                              const float *y, const float *z, const uint32_t *pid,
                                              float *ox, float *oy, float *oz, uint32_t *opid)
{
    constexpr int kSeg = kChunk / (kBlock / 64);
    constexpr int kIt = kSeg / 64;
    __shared__ uint32_t s_cnt[kBlock / 64];
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    const CellRec r = recs[cell];
    const float cut = r.cut;
    const float *key = axis_ptr(x, y, z, r.axis);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t sb = b + (uint32_t)(w * kSeg);
    float v[kIt];
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
        const uint32_t p = sb + (uint32_t)(64 * j + lane);
        v[j] = p < e ? key[p] : 0.0f;
    }
    uint32_t nle = 0, nin = 0;
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
        const bool in = sb + (uint32_t)(64 * j + lane) < e;
        nle += (uint32_t)__popcll(__ballot(in && v[j] <= cut));
        nin += (uint32_t)__popcll(__ballot(in));
    }
    if (lane == 0) s_cnt[w] = nle | ((nin - nle) << 16);  // each <= 1024
    __syncthreads();
    uint32_t ol = 0, orr = 0;
    for (int u = 0; u < w; ++u) {
        ol += s_cnt[u] & 0xffffu;
        orr += s_cnt[u] >> 16;
    }
    const uint64_t base = chunk_pre[q] - chunk_pre[first[cell]];
    const uint64_t tot = chunk_pre[first[cell + 1]] - chunk_pre[first[cell]];
    const uint32_t cell_nle = (uint32_t)(tot & 0xffffffffull);
    uint32_t dl = r.begin + (uint32_t)(base & 0xffffffffull) + ol;
    uint32_t dr = r.begin + cell_nle + (uint32_t)(base >> 32) + orr;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
        const uint32_t p = sb + (uint32_t)(64 * j + lane);
        const bool in = p < e;
        const bool left = in && v[j] <= cut, right = in && !(v[j] <= cut);
        const uint64_t mL = __ballot(left), mR = __ballot(right);
        if (in) {
            const uint32_t d = left ? dl + (uint32_t)__popcll(mL & below)
                                    : dr + (uint32_t)__popcll(mR & below);
            ox[d] = x[p];
            oy[d] = y[p];
            oz[d] = z[p];
            if (opid) opid[d] = pid[p];
        }
        dl += (uint32_t)__popcll(mL);
        dr += (uint32_t)__popcll(mR);
    }
}

__global__ __launch_bounds__(kBlock) void k_stable_scatter(const CellRec *recs, int32_t nCells,
                                                           const uint32_t *first,
                                                           const uint32_t *chunk_cell,
                                                           const uint64_t *chunk_pre,
                                                           const float *x, const float *y,
                                                           const float *z, const uint32_t *pid,
                                                           float *ox, float *oy, float *oz,
                                                           uint32_t *opid)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_stable_scatter_chunk(q, recs, nCells, first, chunk_cell, chunk_pre, x, y, z, pid, ox,
                               oy, oz, opid);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}

// The level's ranges back from the other buffer (a level that does not hold
// every particle cannot flip buffers).
__device__ inline void k_copy_back_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                         const uint32_t *first, const uint32_t *chunk_cell,
                                         const float *ox, const float *oy, const float *oz,
                                         const uint32_t *opid, float *x, float *y, float *z,
                                         uint32_t *pid)
{
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    for (uint32_t p = b + threadIdx.x; p < e; p += kBlock) {
        x[p] = ox[p];
        y[p] = oy[p];
        z[p] = oz[p];
        if (pid) pid[p] = opid[p];
    }
}

__global__ __launch_bounds__(kBlock) void k_copy_back(const CellRec *recs, int32_t nCells,
                                                      const uint32_t *first,
                                                      const uint32_t *chunk_cell, const float *ox,
                                                      const float *oy, const float *oz,
                                                      const uint32_t *opid, float *x, float *y,
                                                      float *z, uint32_t *pid)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x)
        k_copy_back_chunk(q, recs, nCells, first, chunk_cell, ox, oy, oz, opid, x, y, z, pid);
}

__global__ void k_stable_ranges(const CellRec *recs, int32_t nCells, const uint32_t *first,
                                const uint64_t *chunk_pre, SplitRec *splits, uint32_t *ranges,
                                int32_t *status, int64_t cap)
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nCells) return;
    const CellRec r = recs[c];
    if (r.flags) return;
    uint64_t tot = chunk_pre[first[c + 1]] - chunk_pre[first[c]];
    uint32_t nle = (uint32_t)(tot & 0xffffffffull);
    (void)cap;
    int64_t lid = left_child_id(r.id), rid = right_child_id(r.id);
    if (r.axis < 0 || r.axis > 2 || rid >= r.capr) {
        report_error(status + r.st, r.id);
        return;
    }
    splits[c] = SplitRec{nle, (uint32_t)(tot >> 32), 0u, nle};
    const int64_t gl = r.hb + lid, gr = r.hb + rid;
    ranges[2 * gl] = r.begin;
    ranges[2 * gl + 1] = r.begin + nle;
    ranges[2 * gr] = r.begin + nle;
    ranges[2 * gr + 1] = r.end;
}

int launch_stable_le(Ctx &c, int32_t nCells, int32_t *status, bool full_cover,
                     const HistCounts *hc)
{
    if (!status) status = c.status.p;
    ORB_TRY(ensure_stable_buffers(c));
    const unsigned gq = chunk_grid(c);
    ORB_TRY(launch_stopper_counts(c, nCells, true, hc));
    float *ob = c.other();
    float *ox = ob, *oy = ob + c.stride, *oz = ob + 2 * c.stride;
    hipLaunchKernelGGL(k_stable_scatter, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_cell.p, c.chunk_pre.p, c.x(), c.y(), c.z(), c.ids(),
                       ox, oy, oz, c.other_ids());
    ORB_HIP(hipGetLastError());
    if (full_cover) {
        // every particle was written to the other buffer: it becomes current
        c.in_tmp = !c.in_tmp;
    } else {
        hipLaunchKernelGGL(k_copy_back, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                           c.cell_first.p, c.chunk_cell.p, ox, oy, oz, c.other_ids(), c.x(), c.y(),
                           c.z(), c.ids());
        ORB_HIP(hipGetLastError());
    }
    unsigned gc = (unsigned)((nCells + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_stable_ranges, dim3(gc), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_pre.p, c.splits.p, c.ranges.p, status,
                       (int64_t)c.cap);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

// ---------------------------------------------------------------------------
// ServiceMakeAxis gather and ServiceCount sizes
// ---------------------------------------------------------------------------
__device__ inline void k_make_axis_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                                      const uint32_t *first,
                                                      const uint32_t *chunk_cell, const float *x,
                                                      const float *y, const float *z, float *axisT)
{
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    const CellRec r = recs[cell];
    if (r.axis < 0 || r.axis > 2) return;
    const float *key = axis_ptr(x, y, z, r.axis);
    for (uint32_t p = b + threadIdx.x; p < e; p += kBlock) axisT[p] = key[p];
}

__global__ __launch_bounds__(kBlock) void k_make_axis(const CellRec *recs, int32_t nCells,
                                                      const uint32_t *first,
                                                      const uint32_t *chunk_cell, const float *x,
                                                      const float *y, const float *z, float *axisT)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_make_axis_chunk(q, recs, nCells, first, chunk_cell, x, y, z, axisT);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}


int launch_make_axis(Ctx &c, int32_t nCells)
{
    ORB_TRY(c.axisT.alloc((size_t)c.stride + 64));
    hipLaunchKernelGGL(k_make_axis, dim3(chunk_grid(c)), dim3(kBlock), 0, c.stream,
                       c.recs.p, nCells, c.cell_first.p, c.chunk_cell.p, c.x(), c.y(), c.z(),
                       c.axisT.p);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

__global__ void k_cell_sizes(const CellRec *recs, int32_t nCells, uint32_t *out)
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nCells) out[c] = recs[c].end - recs[c].begin;
}

int launch_cell_sizes(Ctx &c, int32_t nCells, uint32_t *dst)
{
    unsigned gc = (unsigned)((nCells + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_cell_sizes, dim3(gc), dim3(kBlock), 0, c.stream, c.recs.p, nCells, dst);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

}  // namespace orb
