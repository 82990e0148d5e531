// orb_static.hip -- the count-left service with no allocation.
//
// orb_count_left_static is orb_count_left / orb_count_left_gpu over arrays the
// caller already holds in device memory (particles, range map, cells, counts)
// for callers that may not allocate: code under stream capture, tight serving
// loops.  Its tables (cell records, chunk offsets, chunk table, per-chunk and
// per-cell counts) live in one file-scope device array, `scratch`, sized for
// the reference's cell limit (MAX_CELLS = 8096, src/constants.h:11) and 2^30
// particles: each GPU the library runs on holds one copy until the process
// ends, and every call on that GPU shares it (calls that overlap on one GPU
// race).  The launches are those of the plain count pass.
#include <mutex>

#include "orb_device.hpp"

namespace orb {

namespace {

constexpr int32_t kStaticMaxCells = 8096;
constexpr int64_t kStaticMaxKeys = int64_t(1) << 30;
constexpr uint32_t kStaticMaxChunks = (uint32_t)(kStaticMaxKeys / kChunk) + kStaticMaxCells + 1;
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// byte offsets of the tables inside the scratch array
constexpr size_t kOffRecs = 0;
constexpr size_t kOffNchunks = kOffRecs + align256(sizeof(CellRec) * kStaticMaxCells);
constexpr size_t kOffFirst = kOffNchunks + align256(sizeof(uint32_t) * (kStaticMaxCells + 1));
constexpr size_t kOffPartials = kOffFirst + align256(sizeof(uint32_t) * (kStaticMaxCells + 1));
constexpr size_t kOffCnt = kOffPartials + 256;  // 8 scan partials + the overflow word
constexpr size_t kOffStamps = kOffCnt + align256(sizeof(uint32_t) * 2 * kStaticMaxCells);
constexpr size_t kOffChunkCell = kOffStamps + align256(sizeof(CountStamp) * kStaticMaxCells);
constexpr size_t kOffChunkDesc = kOffChunkCell + align256(sizeof(uint32_t) * kStaticMaxChunks);
constexpr size_t kOffChunkLtLe = kOffChunkDesc + align256(sizeof(uint4) * kStaticMaxChunks);
constexpr size_t kStaticScratchBytes = kOffChunkLtLe + align256(sizeof(uint32_t) * (kStaticMaxChunks + 1));
static_assert(kStaticScratchBytes < (size_t(8) << 20), "scratch stays below 8 MiB");
constexpr int kOverflowWord = 8;  // index among the partials' words

}  // namespace

// The tables of orb_count_left_static (7.2 MB).  Host code never names it
// except to look up its address on each device (hipGetSymbolAddress).
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// This device's copy of the scratch array: looked up once per device, with
// the device current, and kept.
int device_scratch(unsigned char **out)
{
    int dev = 0;
    ORB_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return ORB_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        ORB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ORB_ERR_HIP;
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return ORB_OK;
}

}  // namespace

// Cell records from caller cells and the caller's range map.  Defensive where
// the context path checks on the host: an id outside [0, cap) is an empty
// cell, ranges are clamped into [0, n).
__global__ void k_static_recs(const orb_cell *cells, int32_t nCells, const uint32_t *ranges,
                              int64_t cap, uint32_t n, CellRec *recs, uint32_t *nchunks)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nCells) return;
    if (i == nCells) {
        nchunks[nCells] = 0u;
        return;
    }
    const orb_cell cl = cells[i];
    CellRec r;
    const bool known = cl.id >= 0 && (int64_t)cl.id < cap;
    uint32_t b = known ? ranges[2 * (size_t)cl.id] : 0u;
    uint32_t e = known ? ranges[2 * (size_t)cl.id + 1] : 0u;
    b = b < n ? b : n;
    e = e < b ? b : (e < n ? e : n);
    r.begin = b;
    r.end = e;
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
    r.capr = (int32_t)(cap < INT32_MAX ? cap : INT32_MAX);
    r.lp = 0;
    r.st = 0;
    r.lmax = 0;
    recs[i] = r;
    nchunks[i] = (e - b + (uint32_t)kChunk - 1u) / (uint32_t)kChunk;
}

// Keeps the chunk offsets within the table.  Disjoint ranges in [0, n) hold
// at most n / 4096 + nCells chunks, which fits; overlapping ones may not:
// then the offsets are cut at the table's end (every access stays inside it)
// and the overflow word is set, which turns every count into UINT32_MAX.
__global__ void k_static_clamp(uint32_t *first, int32_t nCells, uint32_t max_chunks,
                               uint32_t *overflow)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= nCells;
         i += gridDim.x * blockDim.x) {
        const uint32_t v = first[i];
        if (i == nCells) *overflow = v > max_chunks ? 1u : 0u;
        if (v > max_chunks) first[i] = max_chunks;
    }
}

// out[i] = the cell's count (< cut skipping found cells, or <= cut for all).
__global__ void k_static_out(const CellRec *recs, int32_t nCells, const uint32_t *cnt,
                             const uint32_t *overflow, int lt_skip_found, uint32_t *out)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nCells) return;
    if (lt_skip_found && recs[i].found) return;  // countLeft.cpp:19-21: left untouched
    out[i] = *overflow ? 0xffffffffu : cnt[2 * (size_t)i + (lt_skip_found ? 0 : 1)];
}

}  // namespace orb

using namespace orb;

extern "C" int orb_count_left_static(const float *x, const float *y, const float *z, int64_t n,
                                     const uint32_t *ranges, int64_t cap, const orb_cell *cells,
                                     int32_t nCells, int32_t cmp, uint32_t *out, void *stream)
{
    if (nCells < 0 || nCells > kStaticMaxCells || n < 0 || n > kStaticMaxKeys || cap < 0 ||
        (cmp != ORB_CMP_LT && cmp != ORB_CMP_LE))
        return ORB_ERR_ARG;
    if (nCells == 0) return 0;
    if (!cells || !out || !ranges || cap == 0 || (n > 0 && (!x || !y || !z))) return ORB_ERR_ARG;
    unsigned char *base = nullptr;
    ORB_TRY(device_scratch(&base));
    hipStream_t s = static_cast<hipStream_t>(stream);
    CellRec *recs = reinterpret_cast<CellRec *>(base + kOffRecs);
    uint32_t *nchunks = reinterpret_cast<uint32_t *>(base + kOffNchunks);
    uint32_t *first = reinterpret_cast<uint32_t *>(base + kOffFirst);
    uint32_t *partials = reinterpret_cast<uint32_t *>(base + kOffPartials);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(base + kOffCnt);
    CountStamp *stamps = reinterpret_cast<CountStamp *>(base + kOffStamps);
    uint32_t *chunk_cell = reinterpret_cast<uint32_t *>(base + kOffChunkCell);
    uint4 *chunk_desc = reinterpret_cast<uint4 *>(base + kOffChunkDesc);
    uint32_t *chunk_ltle = reinterpret_cast<uint32_t *>(base + kOffChunkLtLe);
    const unsigned gc = (unsigned)((nCells + 1 + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_static_recs, dim3(gc), dim3(kBlock), 0, s, cells, nCells, ranges, cap,
                       (uint32_t)n, recs, nchunks);
    ORB_HIP(hipGetLastError());
    // every entry is written before it is read: nothing needs clearing
    ORB_TRY(scan_exclusive_u32_into(s, nchunks, first, (size_t)nCells + 1, partials));
    hipLaunchKernelGGL(k_static_clamp, dim3(gc), dim3(kBlock), 0, s, first, nCells,
                       kStaticMaxChunks, partials + kOverflowWord);
    ORB_HIP(hipGetLastError());
    const int64_t chunks = n / kChunk + nCells + 1;
    const unsigned gq = (unsigned)(chunks < 65536 ? chunks : 65536);
    const bool lt = cmp == ORB_CMP_LT;
    ORB_TRY(launch_count_explicit(s, recs, nCells, first, chunk_cell, chunk_desc, kStaticMaxChunks,
                                  x, y, z, lt, chunk_ltle, cnt, stamps, gq));
    hipLaunchKernelGGL(k_static_out, dim3((unsigned)((nCells + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       s, recs, nCells, cnt, partials + kOverflowWord, lt ? 1 : 0, out);
    ORB_HIP(hipGetLastError());
    return (int)(nCells * sizeof(uint32_t));
}
