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
#include <algorithm>

#include "orb_device.hpp"

namespace orb {


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
// exclusive scan (two kernels; m <= kScanTile^2)
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

// raw: partials holds each tile's sum (not yet scanned); the block adds up
// those of the tiles before it (at most kScanTile of them, kScanItems per
// thread), which saves a launch over scanning them first.
template <class T>
__global__ void k_scan_apply(const T *in, T *out, size_t m, const T *partials, const uint32_t *lim,
                             int raw)
{
    __shared__ T sm[4];
    T offset = 0;
    if (partials) {
        if (raw) {
            T p = 0;
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) {
                const uint32_t i = threadIdx.x * kScanItems + k;
                if (i < blockIdx.x) p += partials[i];
            }
            offset = block_sum(p, sm);
        } else {
            offset = partials[blockIdx.x];
        }
    }
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
    T pre = block_exclusive_scan(s, sm, &total) + offset;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < mw) out[base + k] = pre;
        pre += v[k];
    }
}

// partials: room for one T per tile of kScanTile entries
template <class T>
static int scan_exclusive_into(hipStream_t stream, const T *in, T *out, size_t m, T *partials,
                               const uint32_t *lim)
{
    if (m == 0) return ORB_OK;
    size_t np = (m + kScanTile - 1) / kScanTile;
    if (np > (size_t)kScanTile) return ORB_ERR_ARG;
    if (np == 1) {
        // one tile: the block scan alone (offset 0)
        hipLaunchKernelGGL(k_scan_apply<T>, dim3(1), dim3(kBlock), 0, stream, in, out, m,
                           (const T *)nullptr, lim, 0);
        ORB_HIP(hipGetLastError());
        return ORB_OK;
    }
    hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)np), dim3(kBlock), 0, stream, in, m, partials,
                       lim);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)np), dim3(kBlock), 0, stream, in, out, m,
                       (const T *)partials, lim, 1);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
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
    return scan_exclusive_into<T>(c.stream, in, out, m, reinterpret_cast<T *>(scratch.p), lim);
}

int scan_exclusive_u32_into(hipStream_t s, const uint32_t *in, uint32_t *out, size_t m,
                            uint32_t *partials)
{
    return scan_exclusive_into<uint32_t>(s, in, out, m, partials, nullptr);
}

unsigned chunk_grid(const Ctx &c)
{
    // chunk kernels loop over the level's chunks: a bounded grid keeps a
    // level with few chunks from paying for a million empty blocks
    int64_t cap = 65536;
    if (c.cur_level >= 0 && c.cur_level < (int32_t)c.tune_chunk_blocks.size() &&
        c.tune_chunk_blocks[(size_t)c.cur_level] > 0)
        cap = c.tune_chunk_blocks[(size_t)c.cur_level];
    else if (c.cur_level >= 0 && c.cur_level < (int32_t)c.level_chunk_hint.size() &&
             c.level_chunk_hint[(size_t)c.cur_level] >= 0)
        // a block per chunk the level is expected to hold, at least one per
        // CU (a level whose cells all sit in finished subtrees has none)
        cap = std::max<int64_t>(256, std::min<int64_t>(cap, c.level_chunk_hint[(size_t)c.cur_level]));
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
    if (c.weighted) ORB_TRY(c.wt_tmp.alloc((size_t)c.stride));
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

// Non-finite input check: NaN compares false both ways, so it is neither
// below nor above any cut, which the histogram rounds cannot count the way
// the reference's comparisons treat it.  Builds refuse such input: any NaN
// in x, y or z raises *word to kErrInputWord (one atomic per wave that saw
// one, none otherwise).
__global__ void k_find_nan(const float *x, const float *y, const float *z, size_t n,
                           unsigned long long *word)
{
    bool any = false;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
        any = any || isnan(x[k]) || isnan(y[k]) || isnan(z[k]);
    if (__ballot(any) != 0ull && (threadIdx.x & 63) == 0) atomicMax(word, (unsigned long long)kErrInputWord);
}

int launch_find_nan(hipStream_t s, const float *x, const float *y, const float *z, size_t n,
                    int32_t *status_pair)
{
    if (n == 0) return ORB_OK;
    hipLaunchKernelGGL(k_find_nan, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, x, y, z, n,
                       reinterpret_cast<unsigned long long *>(status_pair));
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
    if (c.weighted) ORB_TRY(launch_copy_u32(c.stream, c.wt_tmp.p, c.wt.p, (size_t)c.stride));
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

// Each chunk's cell, and its record: first key, one past the last, cell,
// axis (kChunkIdle when the cell has its cut already, belongs to a subtree
// or has no axis).  A kernel that takes a chunk per block or wave reaches
// its keys after one load of the record (issued together with the level's
// chunk count) instead of the chain chunk count -> chunk -> cell -> cell
// record.
__global__ void k_fill_chunks(const uint32_t *first, int32_t nCells, const CellRec *recs,
                              uint32_t *chunk_cell, uint4 *chunk_desc, uint32_t max_chunks)
{
    uint32_t total = first[nCells];
    if (total > max_chunks) total = max_chunks;
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += gridDim.x * blockDim.x)
        fill_chunk(q, first, nCells, recs, chunk_cell, chunk_desc);
}

int ensure_level_capacity(Ctx &c, int64_t nCells, int64_t chunks)
{
    if (nCells <= c.max_level_cells && chunks <= c.max_chunks && c.recs.p) return ORB_OK;
    int64_t m = nCells > c.max_level_cells ? nCells : c.max_level_cells;
    c.max_level_cells = m;
    // disjoint cells (a level's) hold at most n / kChunk + m chunks; a
    // service call's cells may overlap and ask for more
    const int64_t level_chunks = c.n / kChunk + m + 1;
    int64_t mc = level_chunks > c.max_chunks ? level_chunks : c.max_chunks;
    c.max_chunks = chunks > mc ? chunks : mc;
    ORB_TRY(c.dcells.alloc(m));
    ORB_TRY(c.recs.alloc(m));
    ORB_TRY(c.cell_nchunks.alloc(m + 1));
    ORB_TRY(c.cell_first.alloc(m + 1));
    ORB_TRY(c.chunk_cell.alloc(c.max_chunks));
    ORB_TRY(c.chunk_desc.alloc(c.max_chunks));
    ORB_TRY(c.cnt.alloc(2 * m));
    ORB_TRY(c.stamps.alloc(m));
    ORB_TRY(c.chunk_ltle.alloc(c.max_chunks + 1));
    ORB_TRY(c.chunk_cnt.alloc(c.max_chunks + 1));
    ORB_TRY(c.chunk_pre.alloc(c.max_chunks + 1));
    ORB_TRY(c.masks.alloc(2 * (size_t)kMaskWords * (size_t)c.max_chunks));
    ORB_TRY(c.cell_cand.alloc((size_t)m * kBins));
    ORB_TRY(c.splits.alloc(m));
    ORB_TRY(c.chunk_comp.alloc(8 * ((size_t)c.max_chunks + 1)));
    ORB_TRY(c.sel_lists.alloc((size_t)c.max_chunks + 1));
    ORB_TRY(c.swap_plan.alloc(2 * (size_t)c.max_chunks));
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
int build_chunk_table(Ctx &c, int32_t nCells, bool with_stamps, bool scanned)
{
    if (!scanned)
        ORB_TRY(scan_exclusive_u32(c, c.cell_nchunks.p, c.cell_first.p, (size_t)nCells + 1));
    hipLaunchKernelGGL(k_fill_chunks, dim3(grid_for((size_t)c.max_chunks, 16384)), dim3(kBlock), 0,
                       c.stream, c.cell_first.p, nCells, c.recs.p, c.chunk_cell.p, c.chunk_desc.p,
                       (uint32_t)c.max_chunks);
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
    if (r.axis < 0 || r.axis > 2) {  // no cut axis: nothing is left of the cut
        if (threadIdx.x == 0) chunk_ltle[q] = 0u;
        return;
    }
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
    const uint32_t total = first[nCells];  // the caller keeps it within the table
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

// Weighted count pass (workload weights w(p_i), documentation/notes.tex:
// 259-271): per chunk the weight of the keys < cut and <= cut (total_only:
// of every key, the cell's weight for ServiceCount), 64-bit sums.  One block
// per chunk, coalesced loads of keys and weights, wave shuffles and LDS.
// Roofline: HBM, 8 B per key (key + weight).
__global__ __launch_bounds__(kBlock) void k_count_w(const CellRec *recs, int32_t nCells,
                                                    const uint32_t *first,
                                                    const uint32_t *chunk_cell, const float *x,
                                                    const float *y, const float *z,
                                                    const uint32_t *wgt, int skip_found,
                                                    int total_only, uint64_t *chunk_w)
{
    __shared__ uint64_t sm[2][kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        uint32_t cell, b, e;
        chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e);
        const CellRec r = recs[cell];
        if (!total_only && skip_found && r.found) continue;  // uniform per block
        const bool axis_ok = r.axis >= 0 && r.axis <= 2;  // no axis: nothing left of the cut
        const float *key = axis_ptr(x, y, z, axis_ok ? r.axis : 0);
        uint64_t lt = 0, le = 0;
        for (uint32_t p = b + threadIdx.x; p < e; p += kBlock) {
            const uint64_t w = wgt[p];
            if (total_only) {
                lt += w;
            } else if (axis_ok) {
                const float v = key[p];
                lt += v < r.cut ? w : 0ull;
                le += v <= r.cut ? w : 0ull;
            }
        }
        if (total_only) le = lt;
        lt = wave_sum(lt);
        le = wave_sum(le);
        if (lane == 0) {
            sm[0][wv] = lt;
            sm[1][wv] = le;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            chunk_w[2 * (size_t)q] = sm[0][0] + sm[0][1] + sm[0][2] + sm[0][3];
            chunk_w[2 * (size_t)q + 1] = sm[1][0] + sm[1][1] + sm[1][2] + sm[1][3];
        }
        __syncthreads();  // the next chunk rewrites sm
    }
}

// Per-cell sums of the weighted chunk slots (one block per cell).
__global__ __launch_bounds__(kBlock) void k_cell_totals_w(const CellRec *recs, int32_t nCells,
                                                          const uint32_t *first,
                                                          const uint64_t *chunk_w, int skip_found,
                                                          uint64_t *cnt_w)
{
    __shared__ uint64_t sm[4];
    for (int32_t c = blockIdx.x; c < nCells; c += gridDim.x) {
        if (skip_found && recs[c].found) continue;  // uniform per block
        uint64_t lt = 0, le = 0;
        for (uint32_t q = first[c] + threadIdx.x; q < first[c + 1]; q += kBlock) {
            lt += chunk_w[2 * (size_t)q];
            le += chunk_w[2 * (size_t)q + 1];
        }
        lt = block_sum(lt, sm);
        le = block_sum(le, sm);
        if (threadIdx.x == 0) {
            cnt_w[2 * (size_t)c] = lt;
            cnt_w[2 * (size_t)c + 1] = le;
        }
    }
}

int launch_count_w(Ctx &c, int32_t nCells, bool skip_found, bool total_only)
{
    if (!c.weighted) return ORB_ERR_STATE;
    ORB_TRY(c.chunk_w.alloc(2 * ((size_t)c.max_chunks + 1)));
    ORB_TRY(c.cnt_w.alloc(2 * (size_t)c.max_level_cells));
    ORB_TRY(c.hcnt_w.alloc(2 * (size_t)c.max_level_cells));
    hipLaunchKernelGGL(k_count_w, dim3(chunk_grid(c)), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_cell.p, c.x(), c.y(), c.z(), c.weights(),
                       skip_found ? 1 : 0, total_only ? 1 : 0, c.chunk_w.p);
    const unsigned g = (unsigned)(nCells < 65536 ? (nCells > 0 ? nCells : 1) : 65536);
    hipLaunchKernelGGL(k_cell_totals_w, dim3(g), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                       c.cell_first.p, c.chunk_w.p, (skip_found && !total_only) ? 1 : 0, c.cnt_w.p);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

int launch_count_explicit(hipStream_t s, const CellRec *recs, int32_t nCells, const uint32_t *first,
                          uint32_t *chunk_cell, uint4 *chunk_desc, uint32_t max_chunks,
                          const float *x, const float *y, const float *z, bool skip_found,
                          uint32_t *chunk_ltle, uint32_t *cnt, CountStamp *stamps,
                          unsigned chunk_blocks)
{
    if (nCells <= 0) return ORB_OK;
    hipLaunchKernelGGL(k_fill_chunks, dim3(grid_for((size_t)max_chunks, 16384)), dim3(kBlock), 0, s,
                       first, nCells, recs, chunk_cell, chunk_desc, max_chunks);
    hipLaunchKernelGGL(k_count, dim3(chunk_blocks > 0 ? chunk_blocks : 1u), dim3(kBlock), 0, s, recs,
                       nCells, first, chunk_cell, x, y, z, (const float *)nullptr, skip_found ? 1 : 0,
                       chunk_ltle);
    const unsigned g = (unsigned)(nCells < 65536 ? nCells : 65536);
    hipLaunchKernelGGL(k_cell_totals, dim3(g), dim3(kBlock), 0, s, recs, nCells, first, chunk_ltle,
                       skip_found ? 1 : 0, 0, cnt, stamps);
    ORB_HIP(hipGetLastError());
    return ORB_OK;
}

// Per-chunk counts of the keys going left / right (stable partition):
// taken from the last count pass when the cell was counted at this cut
// (stamp), else one key read.
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


// ---------------------------------------------------------------------------
// Hoare partition over stopper masks (partition.cpp:28-60).
//
// For a cell of n keys and cut c the L-stoppers are the positions whose key
// is not < c (where `while (key[i] < cut) i++` stops), left to right
// L_0 < L_1 < ..., and the R-stoppers those whose key is not > c, right to
// left R_0 > R_1 > ....  The serial loop swaps (L_t, R_t) for every
// t < K = #{t : L_t < R_t} (disjoint transpositions), stops with
// i = min(L_K, R_{K-1}) (i = L_0 when K == 0) and then swaps rows i and
// n-1.  The children are [0, i) and [i, n).
//
//   masks  per chunk: the two stopper bitmasks (bit j of word k = chunk
//          position 64k + j, one ballot per word) and their popcounts
//   scan   per-chunk counts -> each chunk's index offsets inside its cell
//   split  per cell (one wave): K from the chunk offsets and one chunk's
//          masks, then i
//   swap   per chunk of the cell's sparser side: the positions of its pairs
//          selected from the masks into LDS, then x, y, z swapped
//   final  per cell: swap rows i and n-1
//
// The masks are 2 bits per key, so the positions never go through HBM.
// ---------------------------------------------------------------------------

// Stopper masks from the keys: one block per chunk, wave w builds words
// 16w .. 16w + 15 (one ballot per 64 keys), lane j keeps word 16w + j.
__device__ inline void masks_from_keys(const CellRec &r, uint32_t b, uint32_t e, const float *x,
                                       const float *y, const float *z, int w, int lane,
                                       uint64_t &mineL, uint64_t &mineR)
{
    constexpr int kPerWave = kMaskWords / (kBlock / 64);
    const float cut = r.cut;
    const float *key = axis_ptr(x, y, z, r.axis);
    const uint32_t sb = b + (uint32_t)(w * kPerWave * 64);
    float v[kPerWave];
#pragma unroll
    for (int j = 0; j < kPerWave; ++j) {
        const uint32_t p = sb + (uint32_t)(64 * j + lane);
        v[j] = p < e ? key[p] : 0.0f;
    }
    mineL = 0ull;
    mineR = 0ull;
#pragma unroll
    for (int j = 0; j < kPerWave; ++j) {
        const bool in = sb + (uint32_t)(64 * j + lane) < e;
        const uint64_t mL = __ballot(in && !(v[j] < cut));
        const uint64_t mR = __ballot(in && !(v[j] > cut));
        mineL = lane == j ? mL : mineL;
        mineR = lane == j ? mR : mineR;
    }
}

__global__ __launch_bounds__(kBlock) void k_hoare_masks(const CellRec *recs, int32_t nCells,
                                                        const uint32_t *first,
                                                        const uint32_t *chunk_cell,
                                                        const float *x, const float *y,
                                                        const float *z, uint64_t *masks,
                                                        uint64_t *chunk_cnt)
{
    constexpr int kW = kBlock / 64;
    constexpr int kPerWave = kMaskWords / kW;  // 16 words per wave
    __shared__ uint32_t s_cnt[2][kW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        uint32_t cell, b, e;
        chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e);
        const CellRec r = recs[cell];
        uint64_t mineL, mineR;
        masks_from_keys(r, b, e, x, y, z, w, lane, mineL, mineR);
        uint64_t *mq = masks + (size_t)q * (2 * kMaskWords) + w * kPerWave;
        if (lane < kPerWave) {
            mq[lane] = mineL;
            mq[kMaskWords + lane] = mineR;
        }
        const uint32_t nl = wave_sum((uint32_t)__popcll(mineL));
        const uint32_t nr = wave_sum((uint32_t)__popcll(mineR));
        if (lane == 0) {
            s_cnt[0][w] = nl;
            s_cnt[1][w] = nr;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t L = 0, R = 0;
            for (int u = 0; u < kW; ++u) {
                L += s_cnt[0][u];
                R += s_cnt[1][u];
            }
            chunk_cnt[q] = (uint64_t)L | ((uint64_t)R << 32);
        }
        __syncthreads();  // the next chunk reuses s_cnt
    }
}

// Batched form: a wave per two chunks (the work is a few dependent loads,
// so each wave keeps both chunks' loads in flight together).  A cell that
// went through the window round (passes > kHistK) finishes that round's
// masks: keys above the window are L-stoppers only, keys below R-stoppers
// only, and the window keys, packed in position order in the quarter's
// slots, are compared with the cut; lane j owns word j.  Other cells read
// their keys, 64 per load.
//
// The chunks' records and their two window-round mask words per lane are
// requested together; the cell records next, with each lane's first packed
// key.  Chunks of cells that are not searched (subtree cells, no axis) are
// skipped: nothing reads their masks, and their count slot is written as
// zero.
constexpr int kMasksChunks = 2;  // k_hoare_masks_window: chunks per wave
__global__ __launch_bounds__(kBlock) void k_hoare_masks_window(
    const CellRec *__restrict__ recs, int32_t nCells, const uint32_t *__restrict__ first,
    const uint4 *__restrict__ chunk_desc, uint32_t max_chunks, const float *x, const float *y,
    const float *z, const float *pack, uint64_t *masks, uint64_t *chunk_cnt)
{
    constexpr int kW = kBlock / 64;
    constexpr int kQuarter = kMaskWords / kW;  // words per packing quarter
    constexpr int kPer = kMasksChunks;         // chunks per wave, their loads interleaved
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t total = first[nCells];
    for (uint32_t qa = (blockIdx.x * kW + (uint32_t)w) * kPer;; qa += gridDim.x * kW * kPer) {
        // each chunk's record and its two window-round mask words per lane
        uint4 d[kPer];
        uint64_t Gw[kPer], Ww[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const uint32_t q = qa + (uint32_t)i;
            d[i] = chunk_record(chunk_desc, q, max_chunks);
            const uint64_t *mq = masks + (size_t)(q < max_chunks ? q : 0u) * (2 * kMaskWords);
            Gw[i] = mq[lane];
            Ww[i] = mq[kMaskWords + lane];
        }
        if (qa >= total) break;
        // then each chunk's cell record and the first packed key of each lane's word
        float cut[kPer], v0[kPer];
        bool win[kPer];
        uint32_t rank[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const uint32_t q = qa + (uint32_t)i;
            const uint32_t b = uniform(d[i].x), e = uniform(d[i].y), cell = uniform(d[i].z);
            const uint32_t axis = uniform(d[i].w);
            cut[i] = 0.0f;
            v0[i] = 0.0f;
            win[i] = false;
            rank[i] = 0u;
            if (q >= total || axis > 2u) continue;  // uniform
            const uint32_t len = e - b, p0 = 64u * (uint32_t)lane;  // chunk position of bit 0
            const uint64_t valid =
                p0 >= len ? 0ull : (len - p0 >= 64u ? ~0ull : (1ull << (len - p0)) - 1ull);
            // the window round's packed keys of this lane's word, ranked from its
            // in-window mask; the first is requested with the cell record (a
            // cell past that round uses it; for other cells the words are stale
            // but the rank stays below the quarter's valid positions, in bounds)
            const uint64_t W = Ww[i] & valid;
            const uint32_t cw = (uint32_t)__popcll(W);
            const uint32_t ex = wave_inclusive_scan(cw) - cw;
            const int q0 = lane & ~(kQuarter - 1);  // first word of this lane's quarter
            rank[i] = ex - (uint32_t)__shfl((int)ex, q0, 64);
            const float *pk = pack + b + 64u * (uint32_t)q0;
            v0[i] = W ? pk[rank[i]] : 0.0f;
            const CellRec &r = recs[cell];
            cut[i] = r.cut;
            win[i] = r.passes > kHistK;
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const uint32_t q = qa + (uint32_t)i;
            const uint32_t b = uniform(d[i].x), e = uniform(d[i].y);
            const uint32_t axis = uniform(d[i].w);
            if (q >= total) break;  // uniform
            if (axis > 2u) {        // uniform
                if (lane == 0) chunk_cnt[q] = 0ull;
                continue;
            }
            uint64_t *mq = masks + (size_t)q * (2 * kMaskWords);
            const uint32_t len = e - b, p0 = 64u * (uint32_t)lane;
            const uint64_t valid =
                p0 >= len ? 0ull : (len - p0 >= 64u ? ~0ull : (1ull << (len - p0)) - 1ull);
            const float c = cut[i];
            uint64_t L = 0ull, R = 0ull;
            if (win[i]) {  // uniform
                const uint64_t G = Gw[i] & valid;
                const uint64_t W = Ww[i] & valid;
                const int q0 = lane & ~(kQuarter - 1);
                const float *pk = pack + b + 64u * (uint32_t)q0;
                uint32_t rk = rank[i];
                uint64_t wd = W;
                bool first_bit = true;
                while (wd) {
                    const uint32_t bit = (uint32_t)__builtin_ctzll(wd);
                    wd &= wd - 1ull;
                    const float v = first_bit ? v0[i] : pk[rk];
                    ++rk;
                    first_bit = false;
                    L |= (!(v < c) ? 1ull : 0ull) << bit;
                    R |= (!(v > c) ? 1ull : 0ull) << bit;
                }
                L |= G;
                R |= valid & ~G & ~W;
            } else {
                const float *key = axis_ptr(x, y, z, (int)axis);
                for (int j0 = 0; j0 < kMaskWords; j0 += 16) {
                    float v[16];
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const uint32_t p = b + 64u * (uint32_t)(j0 + j) + (uint32_t)lane;
                        v[j] = p < e ? key[p] : 0.0f;
                    }
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const bool in = b + 64u * (uint32_t)(j0 + j) + (uint32_t)lane < e;
                        const uint64_t mL = __ballot(in && !(v[j] < c));
                        const uint64_t mR = __ballot(in && !(v[j] > c));
                        L = lane == j0 + j ? mL : L;
                        R = lane == j0 + j ? mR : R;
                    }
                }
            }
            mq[lane] = L;
            mq[kMaskWords + lane] = R;
            const uint32_t nl = wave_sum((uint32_t)__popcll(L));
            const uint32_t nr = wave_sum((uint32_t)__popcll(R));
            if (lane == 0) chunk_cnt[q] = (uint64_t)nl | ((uint64_t)nr << 32);
        }
    }
}

__device__ inline uint32_t lo32(uint64_t v) { return (uint32_t)v; }
__device__ inline uint32_t hi32(uint64_t v) { return (uint32_t)(v >> 32); }

// Last k in [lo, hi) with pred(k), given pred(lo) and pred true then false
// over [lo, hi): a 64-way search, one ballot per step.  Every lane calls
// pred on every step (with lo where its probe is out of range), so pred
// may shuffle.
template <class Pred>
__device__ inline uint32_t wave_last_true(uint32_t lo, uint32_t hi, Pred pred)
{
    const int lane = threadIdx.x & 63;
    while (hi - lo > 1u) {
        const uint32_t span = hi - lo;
        const bool dense = span <= 65u;
        auto probe = [&](uint32_t j) {
            return dense ? lo + 1u + j : lo + (uint32_t)(((uint64_t)(j + 1u) * span) / 65u);
        };
        const uint32_t k = probe((uint32_t)lane);
        const bool valid = k < hi;
        const bool pr = pred(valid ? k : lo) && valid;
        const uint64_t m = __ballot(pr);
        if (m == 0ull) {
            hi = probe(0u);
        } else {
            const uint32_t j = 63u - (uint32_t)__clzll((long long)m);
            const uint32_t kn = j < 63u ? probe(j + 1u) : hi;
            lo = probe(j);
            hi = kn < hi ? kn : hi;
        }
    }
    return lo;
}

// A chunk's mask of one side, word `lane` in each lane, with the exclusive
// prefix of the words' popcounts.
__device__ inline void load_mask(const uint64_t *masks, uint32_t q, int side, uint64_t &word,
                                 uint32_t &ex)
{
    const int lane = threadIdx.x & 63;
    word = masks[(size_t)q * (2 * kMaskWords) + (size_t)side * kMaskWords + (uint32_t)lane];
    const uint32_t c = (uint32_t)__popcll(word);
    ex = wave_inclusive_scan(c) - c;
}

// Chunk position (0..4095) of the set bit of rank r of a mask held as by
// load_mask.  Each lane asks for its own r (below the mask's popcount); all
// lanes call (shuffles).
__device__ inline uint32_t wave_select(uint64_t word, uint32_t ex, uint32_t r)
{
    int k = 0;  // the last word whose prefix is <= r holds the bit
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const uint32_t e = (uint32_t)__shfl((int)ex, k + s, 64);
        if (e <= r) k += s;
    }
    const uint32_t wl = (uint32_t)__shfl((int)(uint32_t)word, k, 64);
    const uint32_t wh = (uint32_t)__shfl((int)(uint32_t)(word >> 32), k, 64);
    uint32_t rr = r - (uint32_t)__shfl((int)ex, k, 64);
    uint64_t wd = ((uint64_t)wh << 32) | wl;
    uint32_t bit = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const uint32_t c = (uint32_t)__popcll(wd & ((1ull << s) - 1ull));
        if (rr >= c) {
            rr -= c;
            bit += (uint32_t)s;
            wd >>= s;
        }
    }
    return 64u * (uint32_t)k + bit;
}

// The chunk offsets of one cell: side 0 counts L-stoppers, side 1
// R-stoppers, both left to right; P(k) for k in [0, nch] (P(nch) = total).
struct CellOffsets {
    const uint64_t *pre;  // chunk_pre + first chunk of the cell
    uint64_t base;
    __device__ uint32_t at(int side, uint32_t k) const
    {
        const uint64_t p = pre[k] - base;
        return side ? hi32(p) : lo32(p);
    }
    // the chunk in [k0, nch) holding index a (P(k0) <= a < P(nch))
    __device__ uint32_t chunk_of(int side, uint32_t a, uint32_t k0, uint32_t nch) const
    {
        return wave_last_true(k0, nch, [&](uint32_t k) { return at(side, k) <= a; });
    }
};

// Cell-local position of index a of one side (a below the side's total).
__device__ inline uint32_t locate(const CellOffsets &co, const uint64_t *masks, uint32_t f,
                                  uint32_t nch, int side, uint32_t a)
{
    const uint32_t k = co.chunk_of(side, a, 0u, nch);
    uint64_t wd;
    uint32_t ex;
    load_mask(masks, f + k, side, wd, ex);
    return k * (uint32_t)kChunk + wave_select(wd, ex, a - co.at(side, k));
}

// One wave64 per cell.  With A(k) the L-stoppers before chunk k and B(k)
// the R-stoppers at or after it, let k* be the last chunk with
// A(k*) <= B(k*) (A grows, B shrinks).  Every t below
// max(A(k*), B(k*+1)) is swapped and none from min(B(k*), A(k*+1)) on, and
// for t in between both L_t and R_t lie in chunk k*: its masks decide K.
__global__ __launch_bounds__(kBlock) void k_hoare_split(const CellRec *recs, int32_t nCells,
                                                        const uint32_t *first,
                                                        const uint64_t *chunk_pre,
                                                        const uint64_t *masks,
                                                        SplitRec *splits, uint32_t *ranges,
                                                        int32_t *status, int32_t group)
{
    const int lane = threadIdx.x & 63;
    // a cell handled elsewhere (batched form's subtree kernel) gets an empty
    // split from its own lane
    auto active = [&](int32_t c) {
        if (!recs[c].flags) return true;
        splits[c] = SplitRec{0u, 0u, 0u, 0u};
        return false;
    };
    for_cell_groups(nCells, group, active, [&](int32_t c) {
    const CellRec r = recs[c];
    const uint32_t f = first[c], nch = first[c + 1] - f;
    const CellOffsets co{chunk_pre + f, chunk_pre[f]};
    const uint32_t TL = co.at(0, nch), TR = co.at(1, nch);
    const uint32_t n = r.end - r.begin;
    SplitRec s{TL, TR, 0u, 0u};
    const int64_t lid = left_child_id(r.id), rid = right_child_id(r.id);
    if (n == 0 || TL == 0 || r.axis < 0 || r.axis > 2 || rid >= r.capr) {
        // empty cell or no key >= cut: the reference's loop leaves the cell
        if (lane == 0) {
            report_error(status + r.st, r.lp & 63, kErrPartition, r.id);
            s.TL = 0u;  // the swap kernels skip the cell
            splits[c] = s;
        }
        return;
    }
    // A(nch) = TL > 0 = B(nch): k* < nch
    const uint32_t ks = wave_last_true(0u, nch, [&](uint32_t k) {
        return co.at(0, k) + co.at(1, k) <= TR;  // A(k) <= TR - R-stoppers before k
    });
    const uint32_t A0 = co.at(0, ks), A1 = co.at(0, ks + 1);
    const uint32_t B0 = TR - co.at(1, ks), B1 = TR - co.at(1, ks + 1);
    const uint32_t lo_t = A0 > B1 ? A0 : B1, hi_t = B0 < A1 ? B0 : A1;
    uint64_t wL, wR;
    uint32_t exL, exR;
    load_mask(masks, f + ks, 0, wL, exL);
    load_mask(masks, f + ks, 1, wR, exR);
    // u in [1, m] stands for t = lo_t + u - 1; u = 0 is true by definition
    const uint32_t m = hi_t - lo_t;
    const uint32_t K = lo_t + wave_last_true(0u, m + 1u, [&](uint32_t u) {
        const uint32_t t = lo_t + (u == 0u ? 0u : u - 1u);
        const uint32_t pl = wave_select(wL, exL, t - A0);
        const uint32_t pr = wave_select(wR, exR, B0 - 1u - t);
        return u == 0u || pl < pr;
    });
    uint32_t i;
    if (K == 0u) {
        i = locate(co, masks, f, nch, 0, 0u);
    } else {
        i = locate(co, masks, f, nch, 1, TR - K);  // R_{K-1}
        if (K < TL) {
            const uint32_t lk = locate(co, masks, f, nch, 0, K);
            i = lk < i ? lk : i;
        }
    }
    if (lane != 0) return;
    s.K = K;
    s.i = i;
    splits[c] = s;
    const int64_t gl = r.hb + lid, gr = r.hb + rid;
    ranges[2 * gl] = r.begin;
    ranges[2 * gl + 1] = r.begin + i;
    ranges[2 * gr] = r.begin + i;
    ranges[2 * gr + 1] = r.end;
    });
}

// A swap block owns the pairs whose stopper on the cell's sparser side (the
// larger of [0, i) and [i, n), which holds as many swapped stoppers as the
// other) lies in its chunk: at most 4096 pairs, whose partners on the
// denser side span about two chunks.
struct SwapUnit {
    int owner;        // side whose stoppers lie in the chunk (0: L, 1: R)
    uint32_t t0, t1;  // pairs [t0, t1)
};

__device__ inline SwapUnit swap_unit(const CellOffsets &co, const SplitRec &s, uint32_t n,
                                     uint32_t kc)
{
    SwapUnit u;
    u.owner = 2ull * s.i >= (uint64_t)n ? 0 : 1;
    if (u.owner == 0) {
        u.t0 = co.at(0, kc);
        u.t1 = co.at(0, kc + 1);
    } else {
        u.t0 = s.TR - co.at(1, kc + 1);
        u.t1 = s.TR - co.at(1, kc);
    }
    if (u.t1 > s.K) u.t1 = s.K;
    return u;
}

// The indices of one side that pairs [t0, t1) use: L-stopper t, R-stopper
// TR - 1 - t (R-stoppers indexed left to right).
__device__ inline void side_range(int side, const SwapUnit &u, uint32_t TR, uint32_t &a0,
                                  uint32_t &a1)
{
    a0 = side ? TR - u.t1 : u.t0;
    a1 = side ? TR - u.t0 : u.t1;
}

// Swaps m pairs: positions sl[j] and sr[j] (cell-local) for j < m, with
// kSwapBatch pairs per thread in flight (all loads, then all stores).
constexpr int kSwapBatch = 8;

// Particle arrays seen from a cell's first particle: x, y, z, then the
// payload arrays that move with them (ids, weights), whichever are on, in
// that order (`count` of them).  Narrow: raw buffer accesses with 32-bit
// byte offsets over one descriptor per array (one address register per
// access; the cell holds fewer than 2^30 particles).  Wide: plain 64-bit
// addressing.
template <bool kNarrow>
struct CellArrays;

template <>
struct CellArrays<true> {
    __amdgpu_buffer_rsrc_t a[5];
    int count;
    __device__ CellArrays(float *x, float *y, float *z, uint32_t *pid, uint32_t *w, uint32_t begin)
    {
        // raw buffer, 32-bit data format, no range limit below 4 GiB
        a[0] = __builtin_amdgcn_make_buffer_rsrc(x + begin, 0, -1, 0x00020000);
        a[1] = __builtin_amdgcn_make_buffer_rsrc(y + begin, 0, -1, 0x00020000);
        a[2] = __builtin_amdgcn_make_buffer_rsrc(z + begin, 0, -1, 0x00020000);
        uint32_t *p3 = pid ? pid : w, *p4 = pid ? w : nullptr;
        a[3] = __builtin_amdgcn_make_buffer_rsrc(p3 ? p3 + begin : nullptr, 0, -1, 0x00020000);
        a[4] = __builtin_amdgcn_make_buffer_rsrc(p4 ? p4 + begin : nullptr, 0, -1, 0x00020000);
        count = 3 + (pid ? 1 : 0) + (w ? 1 : 0);
    }
    using Off = uint32_t;
    __device__ uint32_t load(int arr, Off o) const
    {
        return __builtin_amdgcn_raw_buffer_load_b32(a[arr], (int)o, 0, 0);
    }
    __device__ void store(int arr, Off o, uint32_t v) const
    {
        __builtin_amdgcn_raw_buffer_store_b32(v, a[arr], (int)o, 0, 0);
    }
};

template <>
struct CellArrays<false> {
    char *a[5];
    int count;
    __device__ CellArrays(float *x, float *y, float *z, uint32_t *pid, uint32_t *w, uint32_t begin)
    {
        a[0] = reinterpret_cast<char *>(x + begin);
        a[1] = reinterpret_cast<char *>(y + begin);
        a[2] = reinterpret_cast<char *>(z + begin);
        uint32_t *p3 = pid ? pid : w, *p4 = pid ? w : nullptr;
        a[3] = reinterpret_cast<char *>(p3 ? p3 + begin : nullptr);
        a[4] = reinterpret_cast<char *>(p4 ? p4 + begin : nullptr);
        count = 3 + (pid ? 1 : 0) + (w ? 1 : 0);
    }
    using Off = uint64_t;
    __device__ uint32_t load(int arr, Off o) const
    {
        return *reinterpret_cast<const uint32_t *>(a[arr] + o);
    }
    __device__ void store(int arr, Off o, uint32_t v) const
    {
        *reinterpret_cast<uint32_t *>(a[arr] + o) = v;
    }
};

// kArrays: the most arrays this instantiation moves (3 without payload, up
// to 5 with ids and weights): the register arrays are sized by it.
template <bool kNarrow, int kArrays>
__device__ inline void swap_pairs(const uint32_t *sl, const uint32_t *sr, uint32_t m,
                                  const CellArrays<kNarrow> &ca)
{
    using Off = typename CellArrays<kNarrow>::Off;
    for (uint32_t j0 = 0; j0 < m; j0 += (uint32_t)(kBlock * kSwapBatch)) {
        const uint32_t j = j0 + threadIdx.x;
        Off pa[kSwapBatch], pb[kSwapBatch];
#pragma unroll
        for (int u = 0; u < kSwapBatch; ++u) {
            const uint32_t jj = j + (uint32_t)(u * kBlock);
            pa[u] = jj < m ? (Off)sl[jj] * (Off)4 : (Off)0;
            pb[u] = jj < m ? (Off)sr[jj] * (Off)4 : (Off)0;
        }
        const int na = ca.count;  // uniform, at most kArrays
        // defined on every path: otherwise the arrays are carried from batch
        // to batch in registers
        uint32_t va[kArrays][kSwapBatch] = {}, vb[kArrays][kSwapBatch] = {};
#pragma unroll
        for (int arr = 0; arr < kArrays; ++arr) {
            if (arr >= na) break;
#pragma unroll
            for (int u = 0; u < kSwapBatch; ++u) {
                if (j + (uint32_t)(u * kBlock) < m) {
                    va[arr][u] = ca.load(arr, pa[u]);
                    vb[arr][u] = ca.load(arr, pb[u]);
                }
            }
        }
#pragma unroll
        for (int arr = 0; arr < kArrays; ++arr) {
            if (arr >= na) break;
#pragma unroll
            for (int u = 0; u < kSwapBatch; ++u) {
                if (j + (uint32_t)(u * kBlock) < m) {
                    ca.store(arr, pa[u], vb[arr][u]);
                    ca.store(arr, pb[u], va[arr][u]);
                }
            }
        }
    }
}

// Positions of a unit's pairs: each thread takes 16 bits of the owner
// chunk's mask (first round only) and of the partner chunk's, one packed
// block scan ranks both, and every stopper whose index the unit uses goes
// to its slot (L index t -> t - t0, R index j -> TR - 1 - j - t0).  Later
// rounds continue on the partner side only, chunk by chunk.
// Each chunk's swap unit, planned once per level by a thread per chunk (so
// that a swap block starts from one load instead of the chain chunk ->
// cell -> split and offsets -> a search for the first partner chunk): the
// owner side and pairs [t0, t1), the cell's R-stopper total, first
// particle, first chunk and chunk count, and the first partner chunk
// (cell-relative).  t0o = ~0 marks a chunk with no pairs to swap.
struct SwapPlan {
    uint32_t t0o;  // t0 | owner << 31
    uint32_t t1;
    uint32_t TR;
    uint32_t begin;
    uint32_t f, nch, kp, pad;
};
static_assert(sizeof(SwapPlan) == 32, "two 16-byte loads");

__global__ __launch_bounds__(kBlock) void k_hoare_plan(const CellRec *__restrict__ recs,
                                                       int32_t nCells,
                                                       const uint32_t *__restrict__ first,
                                                       const uint32_t *__restrict__ chunk_cell,
                                                       const uint64_t *__restrict__ chunk_pre,
                                                       const SplitRec *__restrict__ splits,
                                                       SwapPlan *__restrict__ plan)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += gridDim.x * blockDim.x) {
        const uint32_t cell = chunk_cell[q];
        const SplitRec s = splits[cell];
        SwapPlan pl{~0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
        if (s.K != 0u) {  // else nothing to swap, or reported by k_hoare_split
            const CellRec &r = recs[cell];
            const uint32_t n = r.end - r.begin;
            const uint32_t f = first[cell], nch = first[cell + 1] - f, kc = q - f;
            const CellOffsets co{chunk_pre + f, chunk_pre[f]};
            const SwapUnit u = swap_unit(co, s, n, kc);
            if (u.t0 < u.t1) {
                const int sp = 1 - u.owner;
                uint32_t ap0, ap1;
                side_range(sp, u, s.TR, ap0, ap1);
                // the last chunk k with at(sp, k) <= ap0 (at(sp, 0) = 0)
                uint32_t lo = 0, hi = nch - 1u;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1u) >> 1;
                    if (co.at(sp, mid) <= ap0) lo = mid;
                    else hi = mid - 1u;
                }
                pl = SwapPlan{u.t0 | ((uint32_t)u.owner << 31), u.t1, s.TR, r.begin, f, nch, lo, 0u};
            }
        }
        uint4 *dst = reinterpret_cast<uint4 *>(plan + q);
        dst[0] = make_uint4(pl.t0o, pl.t1, pl.TR, pl.begin);
        dst[1] = make_uint4(pl.f, pl.nch, pl.kp, 0u);
    }
}

template <bool kNarrow, int kArrays>
__global__ __launch_bounds__(kBlock) void k_hoare_swap(const uint32_t *first,
                                                       int32_t nCells,
                                                       const uint64_t *chunk_pre,
                                                       const uint64_t *masks,
                                                       const SwapPlan *plan,
                                                       float *x, float *y, float *z,
                                                       uint32_t *pid, uint32_t *w, int rev)
{
    __shared__ uint32_t spos[2][kChunk];  // cell-local positions of L_t, R_t at t - t0
    __shared__ uint32_t sm[kBlock / 64];
    const uint32_t total = first[nCells];
    const int j = threadIdx.x;
    const int wdi = j >> 2, sh = (j & 3) * 16;  // this thread's 16 bits of a chunk mask
    for (uint32_t q0 = blockIdx.x; q0 < total; q0 += gridDim.x) {
        const uint32_t q = rev ? total - 1u - q0 : q0;  // the level's chunks from the last (rev)
        const uint4 p0 = reinterpret_cast<const uint4 *>(plan + q)[0];
        const uint4 p1 = reinterpret_cast<const uint4 *>(plan + q)[1];
        if (p0.x == ~0u) continue;  // uniform: no pairs in this chunk
        SwapUnit u;
        u.owner = (int)(p0.x >> 31);
        u.t0 = p0.x & 0x7fffffffu;
        u.t1 = p0.y;
        const uint32_t TR = p0.z, begin = p0.w, f = p1.x, nch = p1.y, kc = q - f;
        const CellOffsets co{chunk_pre + f, chunk_pre[f]};
        const int so = u.owner, sp = 1 - so;
        uint32_t ao0, ao1, ap0, ap1;
        side_range(so, u, TR, ao0, ao1);
        side_range(sp, u, TR, ap0, ap1);
        auto put = [&](int side, uint32_t k, uint32_t idx, uint32_t bits, uint32_t a0,
                       uint32_t a1) {
            if (idx >= a1 || idx + (uint32_t)__popc(bits) <= a0) return;
            const uint32_t pos0 = k * (uint32_t)kChunk + 16u * (uint32_t)j;
            uint32_t *dst = spos[side];
            while (bits) {
                const uint32_t b = (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1u;
                if (idx >= a0 && idx < a1) dst[side ? TR - 1u - idx - u.t0 : idx - u.t0] = pos0 + b;
                ++idx;
            }
        };
        uint32_t kp = p1.z;  // the chunk holding the unit's first partner index
        bool own = true;
        while (true) {  // uniform
            const uint32_t bo =
                own ? (uint32_t)(masks[(size_t)(f + kc) * (2 * kMaskWords) +
                                       (size_t)so * kMaskWords + (uint32_t)wdi] >> sh) & 0xffffu
                    : 0u;
            const uint32_t bp = (uint32_t)(masks[(size_t)(f + kp) * (2 * kMaskWords) +
                                                 (size_t)sp * kMaskWords + (uint32_t)wdi] >> sh) &
                                0xffffu;
            uint32_t tot;
            const uint32_t ex = block_exclusive_scan((uint32_t)__popc(bo) | ((uint32_t)__popc(bp) << 16),
                                                     sm, &tot);
            if (own) put(so, kc, co.at(so, kc) + (ex & 0xffffu), bo, ao0, ao1);
            put(sp, kp, co.at(sp, kp) + (ex >> 16), bp, ap0, ap1);
            own = false;
            const uint32_t next = co.at(sp, kp + 1u);
            if (next >= ap1) break;
            // next < ap1 <= P(nch): chunk kp + 1 exists; skip empty chunks by search
            kp = co.at(sp, kp + 2u) > next ? kp + 1u : co.chunk_of(sp, next, kp + 1u, nch);
        }
        __syncthreads();
        const CellArrays<kNarrow> ca(x, y, z, pid, w, begin);
        swap_pairs<kNarrow, kArrays>(spos[0], spos[1], u.t1 - u.t0, ca);
        __syncthreads();  // the next unit overwrites spos
    }
}

__global__ void k_hoare_final(const CellRec *recs, int32_t nCells, const SplitRec *splits,
                              float *x, float *y, float *z, uint32_t *pid, uint32_t *w)
{
    int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nCells) return;
    const CellRec r = recs[c];
    hoare_final_rows(r.begin, r.end, splits[c], x, y, z, pid, w);
}

int launch_hoare_final(Ctx &c, int32_t nCells)
{
    const unsigned gc = (unsigned)((nCells + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_hoare_final, dim3(gc > 0 ? gc : 1), dim3(kBlock), 0, c.stream, c.recs.p,
                       nCells, c.splits.p, c.x(), c.y(), c.z(), c.ids(), c.weights());
    ORB_HIP(hipGetLastError());
    return ORB_OK;
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
        if (r.passes > kHistK) {
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
        } else if (r.found >= 2) {
            // found by the first round: its rows and tie rows are this
            // chunk's counts around the cut
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

// Timing only: the pairs this level's swap exchanges, added to *acc.
__global__ void k_sum_swap_pairs(const CellRec *recs, const SplitRec *splits, int32_t nCells,
                                 unsigned long long *acc)
{
    unsigned long long k = 0;
    for (int32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nCells;
         c += gridDim.x * blockDim.x)
        if (!recs[c].flags) k += splits[c].K;
    k = wave_sum(k);
    if ((threadIdx.x & 63) == 0 && k) atomicAdd(acc, k);
}

int launch_hoare(Ctx &c, int32_t nCells, int32_t *status, const HistCounts *hc, int rev,
                 bool defer_final)
{
    if (!status) status = c.status.p;
    const unsigned gq = chunk_grid(c);
    if (hc && hc->pack)  // batched form: the window round left masks
        hipLaunchKernelGGL(k_hoare_masks_window, dim3((gq + 4 * kMasksChunks - 1) / (4 * kMasksChunks)),
                           dim3(kBlock), 0, c.stream,
                           c.recs.p, nCells, c.cell_first.p, c.chunk_desc.p,
                           (uint32_t)c.max_chunks, c.x(), c.y(), c.z(), hc->pack, c.masks.p,
                           c.chunk_cnt.p);
    else
        hipLaunchKernelGGL(k_hoare_masks, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                           c.cell_first.p, c.chunk_cell.p, c.x(), c.y(), c.z(), c.masks.p,
                           c.chunk_cnt.p);
    ORB_HIP(hipGetLastError());
    ORB_TRY(scan_exclusive_u64(c, c.chunk_cnt.p, c.chunk_pre.p, (size_t)c.max_chunks + 1,
                               c.cell_first.p + nCells));
    const int32_t cgroup = cell_group(nCells);
    hipLaunchKernelGGL(k_hoare_split, dim3(cell_group_blocks(nCells, cgroup)), dim3(kBlock), 0,
                       c.stream, c.recs.p, nCells, c.cell_first.p, c.chunk_pre.p, c.masks.p,
                       c.splits.p, c.ranges.p, status, cgroup);
    ORB_HIP(hipGetLastError());
    if (c.timing && c.swap_pairs.p) {
        const unsigned gs = (unsigned)((nCells + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_sum_swap_pairs, dim3(gs < 1024 ? gs : 1024), dim3(kBlock), 0, c.stream,
                           c.recs.p, c.splits.p, nCells, c.swap_pairs.p);
        ORB_HIP(hipGetLastError());
    }
    // byte offsets inside any cell fit 32 bits below 2^30 particles; the
    // payload-free instantiation keeps the fewest registers
    const bool narrow = c.stride < (int64_t(1) << 30);
    const int arrays = 3 + (c.ids() ? 1 : 0) + (c.weights() ? 1 : 0);
    auto swap = narrow ? (arrays == 3 ? k_hoare_swap<true, 3>
                          : arrays == 4 ? k_hoare_swap<true, 4> : k_hoare_swap<true, 5>)
                       : (arrays == 3 ? k_hoare_swap<false, 3>
                          : arrays == 4 ? k_hoare_swap<false, 4> : k_hoare_swap<false, 5>);
    SwapPlan *plan = reinterpret_cast<SwapPlan *>(c.swap_plan.p);
    hipLaunchKernelGGL(k_hoare_plan, dim3(grid_for((size_t)c.max_chunks, 16384)), dim3(kBlock), 0,
                       c.stream, c.recs.p, nCells, c.cell_first.p, c.chunk_cell.p, c.chunk_pre.p,
                       c.splits.p, plan);
    ORB_HIP(hipGetLastError());
    const int ts = timing_mark(c);
    hipLaunchKernelGGL(swap, dim3(gq), dim3(kBlock), 0, c.stream, c.cell_first.p, nCells,
                       c.chunk_pre.p, c.masks.p, plan, c.x(), c.y(), c.z(), c.ids(), c.weights(),
                       rev);
    ORB_HIP(hipGetLastError());
    ORB_TRY(timing_pair(c, ts, kTimeSwap));
    // the last row swaps: here, or (batched builds) in the next level's first
    // launch, which reads no particle
    return defer_final ? ORB_OK : launch_hoare_final(c, nCells);
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
                                              const uint32_t *wt, float *ox, float *oy, float *oz,
                                              uint32_t *opid, uint32_t *ow)
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
            if (ow) ow[d] = wt[p];
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
                                                           const uint32_t *w, float *ox, float *oy,
                                                           float *oz, uint32_t *opid, uint32_t *ow)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x) {
        k_stable_scatter_chunk(q, recs, nCells, first, chunk_cell, chunk_pre, x, y, z, pid, w, ox,
                               oy, oz, opid, ow);
        __syncthreads();  // the next chunk reuses this block's LDS
    }
}

// The level's ranges back from the other buffer (a level that does not hold
// every particle cannot flip buffers).
__device__ inline void k_copy_back_chunk(uint32_t q, const CellRec *recs, int32_t nCells,
                                         const uint32_t *first, const uint32_t *chunk_cell,
                                         const float *ox, const float *oy, const float *oz,
                                         const uint32_t *opid, const uint32_t *ow, float *x,
                                         float *y, float *z, uint32_t *pid, uint32_t *w)
{
    uint32_t cell, b, e;
    if (!chunk_range(q, first, nCells, chunk_cell, recs, cell, b, e)) return;
    for (uint32_t p = b + threadIdx.x; p < e; p += kBlock) {
        x[p] = ox[p];
        y[p] = oy[p];
        z[p] = oz[p];
        if (pid) pid[p] = opid[p];
        if (w) w[p] = ow[p];
    }
}

__global__ __launch_bounds__(kBlock) void k_copy_back(const CellRec *recs, int32_t nCells,
                                                      const uint32_t *first,
                                                      const uint32_t *chunk_cell, const float *ox,
                                                      const float *oy, const float *oz,
                                                      const uint32_t *opid, const uint32_t *ow,
                                                      float *x, float *y, float *z, uint32_t *pid,
                                                      uint32_t *w)
{
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x; q < total; q += gridDim.x)
        k_copy_back_chunk(q, recs, nCells, first, chunk_cell, ox, oy, oz, opid, ow, x, y, z, pid,
                          w);
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
        report_error(status + r.st, r.lp & 63, kErrPartition, r.id);
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
                       c.weights(), ox, oy, oz, c.other_ids(), c.other_weights());
    ORB_HIP(hipGetLastError());
    if (full_cover) {
        // every particle was written to the other buffer: it becomes current
        c.in_tmp = !c.in_tmp;
    } else {
        hipLaunchKernelGGL(k_copy_back, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                           c.cell_first.p, c.chunk_cell.p, ox, oy, oz, c.other_ids(),
                           c.other_weights(), c.x(), c.y(), c.z(), c.ids(), c.weights());
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
