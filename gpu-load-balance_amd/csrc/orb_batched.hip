// orb_batched.hip -- the batched (device-driven) ORB build.
//
// The plain form asks the host after every bisection pass.  Here the whole
// build runs on the device with no host round trip:
//
//  * the cell heap lives in HBM; the level schedule (which heap ids a level
//    bisects) depends only on d, so the host computes it once;
//  * every cell of a level is bisected in the same launches.  One streaming
//    pass over a cell's keys bins each key against the 255 candidate cuts of
//    the next 8 bisection steps (the midpoints the serial loop could visit,
//    computed with the same float arithmetic).  The candidates are sorted, so
//    prefix sums of the 256-bin histogram give count(x < c) (or <= c) at
//    every candidate, and a replay kernel walks the reference's termination
//    rule over them: 8 passes of src/orbit.cpp:149-231 per HBM read, bit-exact
//    by construction.  Four rounds cover the 32-pass cap;
//  * a cell of at most 8832 keys (7680 with ids, 5120 with weights) is
//    handed to a subtree kernel: one 1024-thread block loads its x, y, z
//    into LDS once, runs every remaining level of the cell's subtree there
//    (the same 8-pass histogram rounds and replay, exact Hoare by ballot
//    scans) and writes x, y, z back once;
//  * the partition of the remaining (large) cells uses the same exact Hoare
//    kernels as the plain form.
#include "orb_device.hpp"

namespace orb {

constexpr int kRounds = (ORB_MAX_PASSES + kHistK - 1) / kHistK;  // 4
// Leaves under a subtree root (the subtree kernel's key capacity is SubCap).
constexpr int kSmallLeaves = 64;
constexpr int kReduceRows = 64;                // chunk histograms per reduce block

struct BatchedState {
    int32_t d = 0, levels_mode = -1;
    int64_t cap = 0;
    int32_t D = 0;
    bool complete = true;              // every level lists all its splittable cells
    std::vector<int32_t> level_n;      // [D+1] cells per level (index l)
    std::vector<uint8_t> level_cover;  // [D+1] the level's cells hold every particle
    std::vector<int64_t> level_off;    // [D+1] offset into level_ids
    std::vector<int32_t> level_min_leaf;  // [D+1] fewest leaves under a cell of the level
    DevBuf<LevelEntry> level_ids;      // all levels' cells, concatenated
    DevBuf<orb_cell> heap;             // [cap]
    DevBuf<uint8_t> exists;            // [cap]
    DevBuf<uint8_t> sub;               // [cap] cell handled inside a subtree kernel
    DevBuf<uint16_t> chunk_hist;       // [max_chunks][256]
    DevBuf<uint32_t> cell_hist;        // [max_level_cells][256]
    DevBuf<uint64_t> chunk_hist_w;     // weighted builds: [max_chunks][256] weight sums
    DevBuf<uint64_t> cell_hist_w;      // weighted builds: [max_level_cells][256]
    bool hist_w_clean = false;         // cell_hist_w all zero (as cell_hist)
    DevBuf<uint64_t> trace;            // [cap][32]
    DevBuf<int32_t> level_passes;      // [64]
    PinBuf<int32_t> hpasses;           // [64]
    // captured build: valid while its signature (graph_signature) holds
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> gsig;
    // cell_hist rows are added to by k_b_hist_reduce and consumed and
    // re-zeroed by k_b_replay in the same round, so the buffer is all zero
    // between builds; it is cleared after allocation and after a failure
    bool hist_clean = false;
};

// The previous level's due row swaps as a kernel argument (nCells = 0:
// none).
struct PendingFinalDev {
    const LevelEntry *entries;
    int32_t nCells;
    const SplitRec *splits;
    float *x, *y, *z;
    uint32_t *pid, *w;
};

// ---------------------------------------------------------------------------
// schedule (host): ids per level, data-independent
// ---------------------------------------------------------------------------
int schedule_levels(int32_t d, int32_t levels_mode, int64_t cap,
                    std::vector<std::vector<int32_t>> &levels, bool *complete,
                    std::vector<int32_t> *min_leaf)
{
    // existence and nLeafCells follow from d alone (left = ceil(n/2),
    // right = n - left), so the schedule needs no particle data
    const int32_t D = n_levels(d);
    std::vector<uint8_t> ex((size_t)cap, 0);
    std::vector<orb_cell> hp((size_t)cap);
    std::memset(hp.data(), 0, sizeof(orb_cell) * hp.size());
    hp[0].nLeafCells = d;
    ex[0] = 1;
    std::vector<int32_t> ids;
    levels.assign((size_t)D + 1, {});
    if (min_leaf) min_leaf->assign((size_t)D + 2, 0);
    *complete = true;
    for (int32_t l = 1; l <= D; ++l) {
        int32_t bad = -1;
        if (!level_cells(d, levels_mode, l, ex, hp, ids, &bad)) return ORB_ERR_UB;
        if (levels_mode == ORB_LEVELS_REFERENCE && l >= D) break;
        // the subtree kernel needs every cell of depth l-1 with >= 2 leaves
        int64_t a = (int64_t(1) << (l - 1)) - 1, e = (int64_t(1) << l) - 2;
        int32_t expect = 0;
        for (int64_t id = a; id <= e; ++id)
            if (ex[(size_t)id] && hp[(size_t)id].nLeafCells >= 2) expect++;
        if (expect != (int32_t)ids.size()) *complete = false;
        for (int32_t id : ids) {
            int32_t nl = hp[(size_t)id].nLeafCells;
            if (min_leaf && ((*min_leaf)[(size_t)l] == 0 || nl < (*min_leaf)[(size_t)l]))
                (*min_leaf)[(size_t)l] = nl;
            int32_t left = (nl + 1) / 2;
            int32_t lid = left_child_id(id), rid = right_child_id(id);
            if (rid >= cap) return ORB_ERR_ARG;
            hp[(size_t)lid].nLeafCells = left;
            hp[(size_t)rid].nLeafCells = nl - left;
            ex[(size_t)lid] = ex[(size_t)rid] = 1;
        }
        levels[(size_t)l] = ids;
    }
    return ORB_OK;
}

static int make_schedule(Ctx &c, BatchedState &b, int32_t levels_mode, bool *complete)
{
    b.d = c.d;
    b.levels_mode = levels_mode;
    b.cap = c.cap;
    b.D = n_levels(c.d);
    std::vector<std::vector<int32_t>> levels;
    ORB_TRY(schedule_levels(c.d, levels_mode, c.cap, levels, complete, &b.level_min_leaf));
    const int32_t lmax = levels_mode == ORB_LEVELS_REFERENCE ? b.D - 1 : b.D;
    std::vector<LevelEntry> all;
    b.level_n.assign((size_t)b.D + 2, 0);
    b.level_off.assign((size_t)b.D + 2, 0);
    b.level_cover.assign((size_t)b.D + 2, 0);
    for (int32_t l = 1; l < (int32_t)levels.size(); ++l) {
        b.level_off[(size_t)l] = (int64_t)all.size();
        b.level_n[(size_t)l] = (int32_t)levels[(size_t)l].size();
        // every particle sits in a cell of depth l-1 while each earlier level
        // split all cells of its depth
        const int32_t expect = l == 1 ? 1 : 2 * b.level_n[(size_t)l - 1];
        b.level_cover[(size_t)l] = (l == 1 || b.level_cover[(size_t)l - 1]) &&
                                   b.level_n[(size_t)l] == expect;
        for (int32_t id : levels[(size_t)l])
            all.push_back(LevelEntry{id, 0, (int32_t)c.cap, l, 0, lmax, *complete ? 1 : 0, 0});
    }
    ORB_TRY(b.level_ids.alloc(all.size() + 1));
    if (!all.empty())
        ORB_HIP(hipMemcpy(b.level_ids.p, all.data(), sizeof(LevelEntry) * all.size(),
                          hipMemcpyHostToDevice));
    return ORB_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// Fresh build state: root cell, empty heap and flags, range(0) = [0, n),
// cleared status and pass counters (no memset or host copy: graph-safe).
__global__ void k_b_init(orb_cell *heap, uint8_t *exists, uint8_t *sub, int64_t cap,
                         orb_cell root, uint64_t *trace, int64_t trace_n, int32_t *level_passes,
                         uint32_t *ranges, uint32_t n, int32_t *status)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < cap; k += stride) {
        ranges[2 * k] = 0u;
        ranges[2 * k + 1] = k == 0 ? n : 0u;
        store_root_or_zero(&heap[k], root, k == 0);
        exists[k] = k == 0 ? 1 : 0;
        sub[k] = 0;
    }
    for (int64_t k = i; k < trace_n; k += stride) trace[k] = ~0ull;
    if (i < 64) level_passes[i] = 0;
    if (i < 2) status[i] = 0;
}

// Level record of entry i from the device heap (small cells are marked
// for the subtree kernel, cells inside a finished subtree are skipped);
// returns its chunk count.
__device__ inline uint32_t level_rec(const LevelEntry *entries, int32_t i, const orb_cell *heap,
                                     const uint32_t *ranges, const uint8_t *sub, int use_subtree,
                                     uint32_t sub_max, CellRec *recs)
{
    const LevelEntry en = entries[i];
    const int32_t id = en.id;
    const int64_t g = (int64_t)en.hb + id;
    const orb_cell cl = heap[g];
    CellRec r;
    r.begin = ranges[2 * g];
    r.end = ranges[2 * g + 1];
    if (r.end < r.begin) r.end = r.begin;
    r.cut = cell_get_cut(cl);
    r.axis = cl.cutAxis;
    r.found = 0;
    r.nLeaf = cl.nLeafCells;
    r.margin_l = cl.cutMarginLeft;
    r.margin_r = cl.cutMarginRight;
    r.id = id;
    r.passes = 0;
    r.hb = en.hb;
    r.capr = en.capr;
    r.lp = en.lp;
    r.st = en.st;
    r.lmax = en.lmax;
    const uint32_t n = r.end - r.begin;
    if (sub[g]) r.flags = 1;
    else if (use_subtree && en.sub_ok && n <= sub_max && cl.nLeafCells <= kSmallLeaves)
        r.flags = 2;
    else r.flags = 0;
    recs[i] = r;
    return r.flags ? 0u : (n + (uint32_t)kChunk - 1u) / (uint32_t)kChunk;
}

// The previous level's last Hoare row swaps, from blocks [fb0, gridDim.x) of
// a level's first launch: cell c's rows from its range in the map and its
// split (the cell records are being rewritten by the other blocks).
__device__ inline void level_finals(uint32_t fb0, const PendingFinalDev &fin,
                                    const uint32_t *ranges)
{
    const int32_t c = (int32_t)((blockIdx.x - fb0) * blockDim.x + threadIdx.x);
    if (c >= fin.nCells) return;
    const LevelEntry en = fin.entries[c];
    const int64_t g = (int64_t)en.hb + en.id;
    hoare_final_rows(ranges[2 * g], ranges[2 * g + 1], fin.splits[c], fin.x, fin.y, fin.z, fin.pid,
                     fin.w);
}

__global__ void k_b_recs(const LevelEntry *entries, int32_t nCells, const orb_cell *heap,
                         const uint32_t *ranges, const uint8_t *sub, int use_subtree,
                         uint32_t sub_max, CellRec *recs, uint32_t *nchunks, uint32_t rec_blocks,
                         PendingFinalDev fin)
{
    if (blockIdx.x >= rec_blocks) {  // uniform
        level_finals(rec_blocks, fin, ranges);
        return;
    }
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nCells) return;
    nchunks[i] = i == nCells ? 0u
                             : level_rec(entries, i, heap, ranges, sub, use_subtree, sub_max, recs);
}

// A level of fewer than kRecsScan cells in one block of kRecsScan threads,
// a cell per thread: the cell records and the exclusive scan of their chunk
// counts into first[] (first[nCells] = the level's chunks), in place of a
// record launch and a scan launch.
constexpr int kRecsScan = 1024;
__global__ __launch_bounds__(kRecsScan) void k_b_recs_scan(
    const LevelEntry *__restrict__ entries, int32_t nCells, const orb_cell *__restrict__ heap,
    const uint32_t *__restrict__ ranges, const uint8_t *__restrict__ sub, int use_subtree,
    uint32_t sub_max, CellRec *__restrict__ recs, uint32_t *__restrict__ first,
    PendingFinalDev fin)
{
    if (blockIdx.x > 0) {  // uniform: the previous level's row swaps
        level_finals(1u, fin, ranges);
        return;
    }
    constexpr int kW = kRecsScan / 64;
    __shared__ uint32_t wsum[kW];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t v =
        t < nCells ? level_rec(entries, t, heap, ranges, sub, use_subtree, sub_max, recs) : 0u;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t pre = 0;
    for (int i = 0; i < w; ++i) pre += wsum[i];
    if (t <= nCells) first[t] = pre + inc - v;  // t = nCells: the level's chunk count
}

// in-order index t in [1, 255] of the depth-8 candidate tree on [lo, hi]
__device__ inline float candidate(int t, float lo, float hi)
{
    int node = kBins / 2, step = kBins / 4;
    while (node != t) {
        float m = cell_cut(lo, hi);
        if (t < node) {
            hi = m;
            node -= step;
        } else {
            lo = m;
            node += step;
        }
        step >>= 1;
    }
    return cell_cut(lo, hi);
}

// Bin of key v among the sorted candidates cand[1..255] (cand[0] = -inf,
// cand[256] = +inf): the number of candidates <= v (LE = 0) or < v (LE = 1).
// Keys outside the margins [lo, hi] never reach here.  The candidates sit
// close to lo + j*(hi-lo)/256, so the arithmetic estimate is checked against
// its two neighbours (one pair of LDS reads); only an estimate that misses
// (rounding at a bin edge, or a window a few ulps wide) walks, and the walk
// alone makes the result exact.
template <int LE>
__device__ inline int bin_of(const float *cand, float v, float lo, float scale, int &tie)
{
    const float ef = fminf(fmaxf((v - lo) * scale, 0.0f), (float)(kBins - 1));
    int e = (int)ef;
    const float c0 = cand[e], c1 = cand[e + 1];
    const bool ok = LE ? (c0 < v && v <= c1) : (c0 <= v && v < c1);
    if (__builtin_expect(!ok, 0)) {
        if (LE) {
            while (e > 0 && cand[e] >= v) --e;
            while (e < kBins - 1 && cand[e + 1] < v) ++e;
        } else {
            while (e > 0 && cand[e] > v) --e;
            while (e < kBins - 1 && cand[e + 1] <= v) ++e;
        }
    }
    // the candidate v equals, if any: cand[e] (<) or cand[e + 1] (<=)
    tie = LE ? (cand[e + 1] == v ? e + 1 : 0) : (cand[e] == v ? e : 0);
    return e;
}

// How far (in bins) the estimate (v - lo) * scale must stay from both edges
// of its bin for that bin to be exact without the table.  With
// M = max(|lo|, |hi|): a candidate is at most 8 float midpoints deep, each
// midpoint of values in [lo, hi] off the exact one by at most 2^-24 M (the
// sum's rounding) plus 2^-150 (a halving into the subnormals), so candidate
// j sits within 2^-21 M + 2^-147 of lo + j (hi - lo) / 256, i.e. within
// (2^-21 M + 2^-147) * scale (1 + 3 * 2^-24) bins of j.  The estimate itself
// carries four roundings (v - lo, hi - lo, 256 / width, the product), at
// most 2^-14 bins anywhere in [0, 256].  The returned bound doubles both
// terms; M >= 2^126 (sums that could overflow) returns 1: never trusted.
__device__ inline float bin_tolerance(float lo, float hi, float scale)
{
    const float m = fmaxf(fabsf(lo), fabsf(hi));
    return m < 0x1p126f ? 0x1p-13f + (m + 0x1p-120f) * (0x1p-20f * scale) : 1.0f;
}

// Bin e of key v in [lo, hi] when the estimate lies more than tol from both
// edges of [e, e + 1) (lim = 0.5 - tol: the fraction's distance from 0.5 is
// below lim): candidate e is then strictly below v and candidate e + 1
// strictly above, so e is the bin for both comparisons and v ties no
// candidate.  False when the table has to decide (NaN estimates included).
// An estimate at or past 256 needs no test of its own: v <= hi keeps it
// within 2^-14 of 256, too close to that edge to pass.
__device__ inline bool bin_estimate(float v, float lo, float scale, float lim, int &e)
{
    const float ef = (v - lo) * scale;
    e = (int)ef;
    return fabsf(__builtin_amdgcn_fractf(ef) - 0.5f) < lim;
}

// The candidate table of a cell into LDS, with the -inf / +inf sentinels
// at 0 and kBins: each thread loads its entry (cand_entry, issued before the
// chunk's keys so that waiting for it does not wait for them), then stores it.
__device__ inline float cand_entry(const float *cell_cand, uint32_t cell)
{
    const int t = threadIdx.x;
    return t ? cell_cand[(size_t)cell * kBins + t] : -INFINITY;
}
__device__ inline void store_cands(float *cand, float entry)
{
    const int t = threadIdx.x;
    cand[t] = entry;
    if (t == 0) cand[kBins] = INFINITY;
}

// The keys of a chunk (record dd) that wave w reads, 16 loads of 64
// consecutive keys, through a range-checked descriptor: loads past the
// chunk's end return 0.  A chunk that is not live (past the level's count,
// or of a cell that is not searched) gets an empty range: its loads return
// 0 and touch no memory.
template <int kPerWave>
__device__ inline void issue_chunk_loads(const float *x, const float *y, const float *z, bool live,
                                         uint4 dd, uint32_t sb, int lane, float (&vv)[kPerWave])
{
    const uint32_t ax = uniform(dd.w), b = uniform(dd.x), e = uniform(dd.y);
    const float *kp = axis_ptr(x, y, z, live ? (int)ax : 0) + (live ? b : 0u);
    const __amdgpu_buffer_rsrc_t kr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(kp), 0, live ? (int)((e - b) * 4u) : 0, 0x00020000);
    // the whole offset in the per-lane operand: the range check covers it,
    // not the scalar offset
#pragma unroll
    for (int j = 0; j < kPerWave; ++j)
        vv[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
            kr, (int)((sb + (uint32_t)lane) * 4u + 256u * (uint32_t)j), 0, 0));
}

// The first round's candidate table of every searched cell of the level
// (the 255 cuts of the next 8 bisection steps from the cell's margins): once
// per cell here instead of once per chunk in the round.  A wave per cell
// (for_cell_groups), four entries per lane.
// The launch also fills the level's chunk table (k_fill_chunks' work): the
// first `fill_blocks` blocks take the chunks, the others the cells.
__global__ __launch_bounds__(kBlock) void k_b_tables(const uint32_t *__restrict__ first,
                                                     int32_t nCells,
                                                     const CellRec *__restrict__ recs,
                                                     uint32_t *__restrict__ chunk_cell,
                                                     uint4 *__restrict__ chunk_desc,
                                                     uint32_t max_chunks,
                                                     float *__restrict__ cell_cand, int32_t group,
                                                     uint32_t fill_blocks)
{
    if (blockIdx.x < fill_blocks) {  // uniform
        uint32_t total = first[nCells];
        if (total > max_chunks) total = max_chunks;
        for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total;
             q += fill_blocks * blockDim.x)
            fill_chunk(q, first, nCells, recs, chunk_cell, chunk_desc);
        return;
    }
    const int lane = threadIdx.x & 63;
    // a cell whose chunks are read: searched, not found, not in a subtree
    auto active = [&](int32_t c) {
        const CellRec &r = recs[c];
        return !r.found && !r.flags && r.axis >= 0 && r.axis <= 2;
    };
    for_cell_groups(nCells, group, blockIdx.x - fill_blocks, gridDim.x - fill_blocks, active,
                    [&](int32_t c) {
        const float l = recs[c].margin_l, h = recs[c].margin_r;
        float4 v;
        v.x = lane ? candidate(4 * lane, l, h) : 0.0f;  // entry 0 is never read
        v.y = candidate(4 * lane + 1, l, h);
        v.z = candidate(4 * lane + 2, l, h);
        v.w = candidate(4 * lane + 3, l, h);
        reinterpret_cast<float4 *>(cell_cand + (size_t)c * kBins)[lane] = v;
    });
}

// First histogram round: one block per chunk bins every key of the chunk
// over the cell's 255 candidate cuts and writes the chunk's 256-bin row and,
// for the stable partition's stopper counts, its tie row.  Wave w takes the quarter [1024w, 1024w + 1024) of the
// chunk, 64 consecutive keys per load.  A key's bin is the estimate
// (v - lo) * 256 / (hi - lo) clamped to [-0.5, 255.5] whenever that estimate
// lies more than the tolerance from a bin edge (a clamped estimate lies a
// half bin outside the margins, past every candidate on its side: bin 0 or
// 255); otherwise the candidate table decides.  One chunk in flight per
// block: eight blocks per CU (64 VGPRs) hide the loads better than a block
// that keeps its next chunk's keys in registers.
template <int LE>
__global__ __launch_bounds__(kBlock) void k_b_hist(
    const CellRec *__restrict__ recs, int32_t nCells, const uint32_t *__restrict__ first,
    const uint4 *__restrict__ chunk_desc, uint32_t max_chunks, const float *__restrict__ x,
    const float *__restrict__ y, const float *__restrict__ z, const float *__restrict__ cell_cand,
    uint16_t *__restrict__ chunk_hist, uint16_t *__restrict__ chunk_tie, int rev)
{
    constexpr int kW = kBlock / 64;
    constexpr int kPerWave = kMaskWords / kW;  // 16 loads of 64 keys
    __shared__ float cand[kBins + 1];
    __shared__ uint32_t hist[kW][kBins];
    __shared__ uint32_t ties[kBins];  // keys equal to candidate j (rare)
    const int t = threadIdx.x, lane = t & 63, w = (int)uniform((uint32_t)t >> 6);
    const uint32_t sb = (uint32_t)(w * kPerWave * 64);  // chunk position of the quarter
    const uint32_t total = first[nCells];
    for (uint32_t q0 = blockIdx.x;; q0 += gridDim.x) {
        const uint32_t q = rev && q0 < total ? total - 1u - q0 : q0;  // from the last (rev)
        const uint4 d = chunk_record(chunk_desc, q, max_chunks);
        const uint32_t b = uniform(d.x), e = uniform(d.y), cell = uniform(d.z);
        const uint32_t axis = uniform(d.w);
        keep_values(b, e, cell, axis);  // the whole record, in flight with total
        if (q >= total) break;
        if (axis > 2u) continue;  // uniform: the cell is not searched
        const float ce = cand_entry(cell_cand, cell);  // issued before the keys
        float v[kPerWave];
        issue_chunk_loads(x, y, z, true, d, sb, lane, v);
        const float lo = uniform(recs[cell].margin_l), hi = uniform(recs[cell].margin_r);
        store_cands(cand, ce);
#pragma unroll
        for (int u = 0; u < kW; ++u) hist[u][t] = 0u;
        ties[t] = 0u;
        __syncthreads();
        const float width = hi - lo;
        const float scale = width > 0.0f ? (float)kBins / width : 0.0f;
        // clamping is exact while the estimate stays within a quarter bin of
        // the truth; past that every key goes to the table
        const float lim = 0.5f - bin_tolerance(lo, hi, scale);
        const float limx = lim > 0.25f ? lim : -1.0f;
        const uint32_t len = e - b;
        uint32_t *h = hist[w];
        auto scan = [&](auto full) __attribute__((always_inline)) {
            static_for<0, kPerWave>([&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                const uint32_t p0 = sb + 64u * (uint32_t)j;  // chunk position of lane 0
                const bool ok = decltype(full)::value || p0 + (uint32_t)lane < len;
                const float vj = v[j];
                const float ef = __builtin_amdgcn_fmed3f((vj - lo) * scale, -0.5f, 255.5f);
                int bin = (int)ef;
                const bool slow = !(fabsf(__builtin_amdgcn_fractf(ef) - 0.5f) < limx);
                if (__ballot(ok && slow)) {  // wave-uniform; rare: near a bin edge
                    if (ok && slow) {
                        int tj;
                        bin = bin_of<LE>(cand, vj, lo, scale, tj);
                        if (chunk_tie && tj) atomicAdd(&ties[tj], 1u);
                    }
                }
                if (ok) atomicAdd(&h[bin], 1u);
            });
        };
        if (len == (uint32_t)kChunk) scan(std::true_type{});  // uniform
        else scan(std::false_type{});
        __syncthreads();
        const uint32_t sum = hist[0][t] + hist[1][t] + hist[2][t] + hist[3][t];  // <= 4096
        chunk_hist[(size_t)q * kBins + t] = (uint16_t)sum;
        if (chunk_tie) chunk_tie[(size_t)q * kBins + t] = (uint16_t)ties[t];
        __syncthreads();  // the next chunk rewrites cand, hist and ties
    }
}

// Second histogram round.  Wave w takes the quarter [1024w, 1024w + 1024)
// of the chunk, 16 keys per lane (64 consecutive keys per load).  Keys below
// lo and above hi are counted in registers, the few in the window [lo, hi]
// are binned, and each wave packs its window keys, in position order, into
// its quarter's slots of `pack` (comp[8q + w] = how many; comp[8q + 4] = the
// chunk's keys below lo).  Later windows lie inside this one, so later
// rounds need only these keys.  The wave also leaves two of the chunk's
// bitmasks for the Hoare partition (bit j of word k = chunk position
// 64k + j): keys above hi (masks[q][0]) and keys in the window (masks[q][1]).
// The cell's final cut lies in [lo, hi], so these decide every stopper bit
// but those of the packed keys.
// Eight waves per SIMD (the compiler would keep the load offsets in scalar
// registers and stop at seven).
template <int LE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_b_hist_window(
    const CellRec *__restrict__ recs, int32_t nCells, const uint32_t *__restrict__ first,
    const uint4 *__restrict__ chunk_desc, uint32_t max_chunks, const float *__restrict__ x,
    const float *__restrict__ y, const float *__restrict__ z, const float *__restrict__ cell_cand,
    uint16_t *__restrict__ chunk_hist, float *__restrict__ pack, uint32_t *__restrict__ comp,
    uint64_t *__restrict__ masks, int rev)
{
    constexpr int kW = kBlock / 64;
    constexpr int kPerWave = kMaskWords / kW;  // 16 loads of 64 keys
    __shared__ float cand[kBins + 1];
    __shared__ uint32_t hist[kBins];  // window keys are few: one histogram
    __shared__ uint32_t sm[4];
    __shared__ float wkeys[kW][kChunk / kW];  // each wave's window keys, packed
    __shared__ uint64_t words[kW][2][kPerWave];  // each wave's mask words, staged
    const int t = threadIdx.x, lane = t & 63, w = (int)uniform((uint32_t)t >> 6);
    const uint64_t lanes_below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t total = first[nCells];
    for (uint32_t q0 = blockIdx.x;; q0 += gridDim.x) {
        const uint32_t q = rev && q0 < total ? total - 1u - q0 : q0;  // from the last (rev)
        const uint4 d = chunk_record(chunk_desc, q, max_chunks);
        const uint32_t b = uniform(d.x), e = uniform(d.y), cell = uniform(d.z);
        const uint32_t axis = uniform(d.w);
        keep_values(b, e, cell, axis);  // the whole record, in flight with total
        if (q >= total) break;
        if (axis > 2u) continue;  // uniform: the cell is not searched
        if (recs[cell].found) continue;  // uniform: the first round found the cut
        // the chunk's keys through a range-checked descriptor: loads past the
        // chunk's end return 0 and are masked out below, with no branches
        const float *kp = axis_ptr(x, y, z, (int)axis) + b;
        const float ce = cand_entry(cell_cand, cell);
        const __amdgpu_buffer_rsrc_t kr =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(kp), 0, (int)((e - b) * 4u), 0x00020000);
        const uint32_t sb = (uint32_t)(w * kPerWave * 64);  // chunk position of the quarter
        float v[kPerWave];
        // the whole offset in the per-lane operand: the range check covers
        // it, not the scalar offset
#pragma unroll
        for (int j = 0; j < kPerWave; ++j)
            v[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                kr, (int)((sb + (uint32_t)lane) * 4u + 256u * (uint32_t)j), 0, 0));
        const float lo = uniform(recs[cell].margin_l), hi = uniform(recs[cell].margin_r);
        store_cands(cand, ce);
        hist[t] = 0u;
        __syncthreads();
        const float width = hi - lo;
        const float scale = width > 0.0f ? (float)kBins / width : 0.0f;
        const uint32_t len = e - b;
        uint32_t below = 0, packed = 0;  // wave-uniform
        float *dst = pack + b + sb;
        float *wk = wkeys[w];
        auto scan = [&](auto full) __attribute__((always_inline)) {
            static_for<0, kPerWave>([&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                uint64_t valid = ~0ull;
                if constexpr (!decltype(full)::value) {
                    const uint32_t p0 = sb + 64u * (uint32_t)j;  // chunk position of lane 0
                    valid = p0 >= len ? 0ull
                                      : (len - p0 >= 64u ? ~0ull : (1ull << (len - p0)) - 1ull);
                }
                const float vj = v[j];
                const uint64_t mB = __ballot(vj < lo) & valid;
                const uint64_t mG = __ballot(vj > hi) & valid;
                const uint64_t mW = valid & ~mB & ~mG;
                below += (uint32_t)__popcll(mB);
                // staged by one lane: a cross-lane write into a register would
                // be lost if the compiler moved that register under a partial
                // exec mask (the binning below is divergent)
                if (lane == 0) {
                    words[w][0][j] = mG;
                    words[w][1][j] = mW;
                }
                if (mW) {  // wave-uniform: most loads have no key in the window
                    if ((mW >> lane) & 1ull) {
                        const uint32_t slot = packed + (uint32_t)__popcll(mW & lanes_below);
                        dst[slot] = vj;
                        wk[slot] = vj;
                    }
                    packed += (uint32_t)__popcll(mW);
                }
            });
        };
        if (len == (uint32_t)kChunk) scan(std::true_type{});  // uniform
        else scan(std::false_type{});
        // the quarter's keys that are neither below nor in the window
        const uint32_t mine = len > sb ? (len - sb < 1024u ? len - sb : 1024u) : 0u;
        const uint32_t above = mine - below - packed;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const uint64_t mineG = lane < kPerWave ? words[w][0][lane] : 0ull;
        const uint64_t mineW = lane < kPerWave ? words[w][1][lane] : 0ull;
        for (uint32_t k = (uint32_t)lane; k < packed; k += 64) {
            const float vk = wk[k];
            int tj;  // ties: the stopper counts of cells past this round use the packed keys
            atomicAdd(&hist[bin_of<LE>(cand, vk, lo, scale, tj)], 1u);
        }
        uint64_t *mq = masks + (size_t)q * (2 * kMaskWords) + w * kPerWave;
        if (lane < kPerWave) {
            mq[lane] = mineG;
            mq[kMaskWords + lane] = mineW;
        }
        if (lane == 0) {
            comp[8 * (size_t)q + w] = packed;
            sm[w] = below | (above << 16);  // each <= 1024
        }
        __syncthreads();
        const uint32_t ba = sm[0] + sm[1] + sm[2] + sm[3];
        if (t == 0) comp[8 * (size_t)q + 4] = ba & 0xffffu;
        uint32_t sum = hist[t];
        if (t == 0) sum += ba & 0xffffu;
        if (t == kBins - 1) sum += ba >> 16;
        chunk_hist[(size_t)q * kBins + t] = (uint16_t)sum;
        __syncthreads();  // the next chunk rewrites cand, hist and sm
    }
}

// Sums chunk histograms into cell histograms: block g takes chunks
// [g*64, g*64+64), thread t owns bin t; all 64 row loads are in flight at
// once, and one atomic per (block, cell, bin) flushes a run.
__global__ __launch_bounds__(kBlock) void k_b_hist_reduce(const CellRec *recs, int32_t nCells,
                                                          const uint32_t *first,
                                                          const uint32_t *chunk_cell,
                                                          const uint16_t *chunk_hist,
                                                          uint32_t *cell_hist)
{
    __shared__ int32_t scell[kReduceRows];  // cell of each row, -1 if not live
    const uint32_t total = first[nCells];
    for (uint32_t q0 = blockIdx.x * (uint32_t)kReduceRows; q0 < total;
         q0 += gridDim.x * (uint32_t)kReduceRows) {
    const uint32_t nrows = total - q0 < (uint32_t)kReduceRows ? total - q0 : (uint32_t)kReduceRows;
    const int t = threadIdx.x;
    if (t < kReduceRows) {
        int32_t cl = -1;
        if ((uint32_t)t < nrows) {
            uint32_t cell = chunk_cell[q0 + t];
            const CellRec &r = recs[cell];
            if (!(r.found || r.flags)) cl = (int32_t)cell;
        }
        scell[t] = cl;
    }
    __syncthreads();
    int32_t cur = -1;
    uint32_t acc = 0;
    uint32_t v[kReduceRows];  // every row's load in flight at once
#pragma unroll
    for (int u = 0; u < kReduceRows; ++u)
        v[u] = ((uint32_t)u < nrows && scell[u] >= 0) ? chunk_hist[(size_t)(q0 + u) * kBins + t] : 0u;
#pragma unroll
    for (int u = 0; u < kReduceRows; ++u) {
        const int32_t cl = (uint32_t)u < nrows ? scell[u] : cur;
        if (cl != cur) {
            if (cur >= 0 && acc) atomicAdd(&cell_hist[(size_t)cur * kBins + t], acc);
            cur = cl;
            acc = 0;
        }
        acc += v[u];
    }
    if (cur >= 0 && acc) atomicAdd(&cell_hist[(size_t)cur * kBins + t], acc);
    __syncthreads();  // scell is rewritten for the next tile
    }
}

__device__ inline void record_level_passes(int32_t *level_passes, int32_t l, int32_t passes)
{
    // a plain read filters almost every atomic once the level maximum is in
    if (passes > *((volatile int32_t *)&level_passes[l])) atomicMax(&level_passes[l], passes);
}

// A level's pass count is the most passes any of its cells took, so a cell
// records its count once, when it finishes.  The cells a wave (or block)
// finishes share one running maximum, recorded per run of cells with the
// same counter (one build's level) and at the end: thousands of cells of a
// level finishing in one round would otherwise queue thousands of atomics
// on one word.
struct PassMax {
    int32_t lp = -1, mx = 0;
    __device__ void add(int32_t *level_passes, int32_t l, int32_t p)
    {
        if (l != lp) {
            flush(level_passes);
            lp = l;
            mx = p;
        } else if (p > mx) {
            mx = p;
        }
    }
    __device__ void flush(int32_t *level_passes)
    {
        if (lp >= 0) record_level_passes(level_passes, lp, mx);
        lp = -1;
    }
};

// End of a 4-wave block whose lanes 0 each hold a PassMax: the waves' last
// runs are combined in LDS and recorded by one thread (one atomic per block
// where its waves share a counter).  Every thread of the block calls it.
__device__ inline void flush_block(PassMax &pm, int32_t *level_passes)
{
    __shared__ int32_t s_lp[kBlock / 64], s_mx[kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        s_lp[w] = pm.lp;
        s_mx[w] = pm.mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        PassMax b;
        for (int i = 0; i < kBlock / 64; ++i)
            if (s_lp[i] >= 0) b.add(level_passes, s_lp[i], s_mx[i]);
        b.flush(level_passes);
    }
}

// One histogram round of the bisection controller (orbit.cpp:149-231),
// replayed by a whole wave.  P[j] (j in [1, 255]) is the count below
// candidate cut j of the round that starts from margins [lo, hi].  Counts
// grow with j and so does the controller's difference, so the passes it
// makes are the binary-search path towards b, the first node whose
// difference is > 0 (go left at a node >= b, right otherwise); the round
// stops at the first node of that path within +-2 of the target.  b comes
// from one ballot count, the path from b alone, so the 8 counts are read in
// parallel instead of one dependent read per pass.
struct RoundOut {
    float L, R;
    int32_t passes;
    bool found;
    int32_t node;  // the candidate the round stopped at (when found)
};

// The controller's difference for a left count (uint32) or a left weight
// (uint64, weighted builds): weights grow with the candidate as counts do.
__device__ inline int64_t round_diff(uint32_t left, float target)
{
    return difference_from_target(left, target);
}
__device__ inline int64_t round_diff(uint64_t left, float target)
{
    return difference_from_target_w(left, target);
}

// The path of one round from b, the first candidate whose difference is
// > 0 (nle + 1), with count_at(j) = P[j] for the path's candidates.
template <class T, class CountAt>
__device__ inline RoundOut replay_path(uint32_t nle, CountAt count_at, float lo, float hi,
                                       float target, int32_t passes0, int lane,
                                       uint64_t *trace_row)
{
    const int b = 1 + (int)nle;
    const int steps = kHistK < ORB_MAX_PASSES - passes0 ? kHistK : ORB_MAX_PASSES - passes0;
    int nodes[kHistK];
    {
        int node = kBins / 2, step = kBins / 4;
#pragma unroll
        for (int d = 0; d < kHistK; ++d) {
            nodes[d] = node;
            node = node >= b ? node - step : node + step;
            step >>= 1;
        }
    }
    T cnt[kHistK];
#pragma unroll
    for (int d = 0; d < kHistK; ++d) cnt[d] = count_at(nodes[d]);
    int s = steps;
#pragma unroll
    for (int d = kHistK - 1; d >= 0; --d) {
        const int64_t df = round_diff(cnt[d], target);
        if (d < steps && df < 3 && df > -3) s = d;
    }
    const bool found = s < steps;
    const int moves = found ? s : steps;
    float L = lo, R = hi;
#pragma unroll
    for (int d = 0; d < kHistK; ++d) {
        if (d < moves) {
            const float cut = cell_cut(L, R);
            if (nodes[d] >= b) R = cut;
            else L = cut;
        }
    }
    const int used = found ? s + 1 : steps;
    if (trace_row && lane < used) {
        T v = 0;
#pragma unroll
        for (int d = 0; d < kHistK; ++d)
            if (d == lane) v = cnt[d];
        trace_row[passes0 + lane] = v;
    }
    int32_t stop = 0;
#pragma unroll
    for (int d = 0; d < kHistK; ++d)
        if (d == s) stop = nodes[d];
    return RoundOut{L, R, passes0 + used, found, found ? stop : 0};
}

template <class T>
__device__ inline RoundOut replay_round(const T *P, float lo, float hi, float target,
                                        int32_t passes0, int lane, uint64_t *trace_row)
{
    uint32_t nle = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = 4 * lane + 1 + q;
        const bool le0 = j < kBins && round_diff(P[j], target) <= 0;
        nle += (uint32_t)__popcll(__ballot(le0));
    }
    return replay_path<T>(nle, [&](int j) { return P[j]; }, lo, hi, target, passes0, lane,
                          trace_row);
}

// A wave-uniform value of lane src.
__device__ inline uint32_t lane_value(uint32_t v, int src)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
}
__device__ inline uint64_t lane_value(uint64_t v, int src)
{
    return (uint64_t)lane_value((uint32_t)v, src) | ((uint64_t)lane_value((uint32_t)(v >> 32), src) << 32);
}

// replay_round with the round's prefix sums in registers, not in LDS: lane l
// holds P[4l + 1 + q] in pv[q].  The path's counts are read across lanes
// (uniform indices), so no store, fence or load stands between the scan and
// the replay.
template <class T>
__device__ inline RoundOut replay_round_lanes(const T (&pv)[4], float lo, float hi, float target,
                                              int32_t passes0, int lane, uint64_t *trace_row)
{
    uint32_t nle = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int j = 4 * lane + 1 + q;
        const bool le0 = j < kBins && round_diff(pv[q], target) <= 0;
        nle += (uint32_t)__popcll(__ballot(le0));
    }
    auto count_at = [&](int j) {  // j in [1, 255], the same in every lane
        const int q = (j - 1) & 3;
        const T v = q == 0 ? pv[0] : (q == 1 ? pv[1] : (q == 2 ? pv[2] : pv[3]));
        return lane_value(v, (j - 1) >> 2);
    };
    return replay_path<T>(nle, count_at, lo, hi, target, passes0, lane, trace_row);
}

// The record's found field after a round that found its cut at candidate
// ro.node of [lo, hi]: 1 + node when that candidate differs from its
// neighbours (the stopper counts can then come from the round's histograms;
// equal candidates -- a window a few ulps wide -- would split the keys equal
// to the cut over several tie bins), else 1 (found, the keys are read); 0
// when not found.
__device__ inline int32_t found_encoding(const RoundOut &ro, float lo, float hi)
{
    if (!ro.found) return 0;
    const int j = ro.node;
    const float cj = candidate(j, lo, hi);
    const float cp = j > 1 ? candidate(j - 1, lo, hi) : -INFINITY;
    const float cn = j < kBins - 1 ? candidate(j + 1, lo, hi) : INFINITY;
    return (cj == cp || cj == cn) ? 1 : 1 + j;
}

// Lane 0 of the replaying wave: the round's margins into the cell's record
// and heap cell; a finished cell (cut found, or the pass cap with the last
// midpoint) records its passes (into pm when given) and is split at once
// (orbit.cpp:235-250).
__device__ inline void finish_round(CellRec r, int32_t c, const RoundOut &ro, int32_t found_enc,
                                    CellRec *recs, orb_cell *heap, int32_t *level_passes,
                                    uint8_t *exists, int32_t *status, PassMax *pm = nullptr)
{
    const float L = ro.L, R = ro.R;
    const int32_t passes = ro.passes;
    const bool found = ro.found;
    r.margin_l = L;
    r.margin_r = R;
    r.cut = cell_cut(L, R);
    r.passes = passes;
    r.found = found_enc;  // the stopper counts read this round's histograms
    recs[c] = r;
    orb_cell &hc = heap[(int64_t)r.hb + r.id];
    hc.cutMarginLeft = L;
    hc.cutMarginRight = R;
    hc.foundCut = found;
    if (!found && passes < ORB_MAX_PASSES) return;  // a later round records it
    if (pm) pm->add(level_passes, r.lp, passes);
    else record_level_passes(level_passes, r.lp, passes);
    const orb_cell p = hc;
    orb_cell lc, rc;
    uint32_t *wl = reinterpret_cast<uint32_t *>(&lc);
    uint32_t *wr = reinterpret_cast<uint32_t *>(&rc);
#pragma unroll
    for (int j = 0; j < (int)(sizeof(orb_cell) / 4); ++j) wl[j] = wr[j] = 0u;
    const int64_t hb = r.hb;
    if (right_child_id(r.id) >= r.capr || !cell_split(p, lc, rc)) {
        report_error(status + r.st, r.lp & 63, kErrSplit, r.id);
        return;
    }
    heap[hb + lc.id] = lc;
    heap[hb + rc.id] = rc;
    exists[hb + lc.id] = 1;
    exists[hb + rc.id] = 1;
}


// Rounds 3 and 4 gather a searching cell's keys into up to 64 lists (see
// k_b_select): list r covers a contiguous range of the cell's chunks.
constexpr int kSelChunks = 4;  // chunks per wave: a lane per chunk quarter
constexpr int kSelLists = 64;
struct SelLists {
    uint32_t f, g0, S, R;  // first chunk, first group, groups per list, lists
    __device__ SelLists(uint32_t first_chunk, uint32_t nch)
    {
        f = first_chunk;
        g0 = f / kSelChunks;
        const uint32_t ng = nch ? (f + nch - 1u) / kSelChunks - g0 + 1u : 0u;
        S = (ng + kSelLists - 1u) / kSelLists;
        R = S ? (ng + S - 1u) / S : 0u;
    }
    __device__ uint32_t list_of_group(uint32_t g) const { return (g - g0) / S; }
    __device__ uint32_t start_chunk(uint32_t r) const
    {
        const uint32_t q = (g0 + r * S) * kSelChunks;
        return q > f ? q : f;
    }
};

// One wave per cell: prefix sums of the cell histogram in registers, then
// the wave replays up to 8 passes of the bisection controller.
__global__ __launch_bounds__(kBlock) void k_b_replay(CellRec *recs, int32_t nCells,
                                                     uint32_t *cell_hist, orb_cell *heap,
                                                     uint64_t *trace, int32_t *level_passes,
                                                     uint8_t *exists, int32_t *status,
                                                     float *next_cand, const uint32_t *first,
                                                     unsigned long long *sel_clear, int32_t group)
{
    const int lane = threadIdx.x & 63;
    auto active = [&](int32_t c) { return !recs[c].found && !recs[c].flags; };
    PassMax pm;  // lane 0
    for_cell_groups(nCells, group, active, [&](int32_t c) {
    const CellRec r = recs[c];
    uint4 *row = reinterpret_cast<uint4 *>(cell_hist + (size_t)c * kBins);
    const uint4 h4 = row[lane];
    row[lane] = make_uint4(0u, 0u, 0u, 0u);  // the next round starts from zero
    const uint32_t s0 = h4.x, s1 = s0 + h4.y, s2 = s1 + h4.z, s3 = s2 + h4.w;
    const uint32_t base = wave_inclusive_scan(s3) - s3;
    // P[4 lane + 1 + q] (P[j] = keys in bins below j)
    const uint32_t pv[4] = {base + s0, base + s1, base + s2, base + s3};
    const uint32_t n = r.end - r.begin;
    const float target = difference_target(n, r.nLeaf);
    const RoundOut ro = replay_round_lanes(pv, r.margin_l, r.margin_r, target, r.passes, lane,
                                           trace ? trace + ((size_t)r.hb + r.id) * ORB_MAX_PASSES
                                                 : nullptr);
    if (next_cand && !ro.found) {  // the next round's candidate table (uniform)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 4 * lane + i;
            next_cand[(size_t)c * kBins + j] = j ? candidate(j, ro.L, ro.R) : 0.0f;
        }
    }
    if (sel_clear) {  // rounds 3 and 4 gather from zero
        const uint32_t f = first[c];
        const SelLists sl(f, first[c + 1] - f);
        if ((uint32_t)lane < sl.R) sel_clear[sl.start_chunk((uint32_t)lane)] = 0ull;
    }
    if (lane != 0) return;
    finish_round(r, c, ro, found_encoding(ro, r.margin_l, r.margin_r), recs, heap, level_passes,
                 exists, status, &pm);
    });
    flush_block(pm, level_passes);
}

// Rounds 3 and 4 read only the keys that the window round packed and that
// lie inside the margins after round 2 (every later candidate lies inside
// them).  k_b_select gathers them, per searching cell, into up to 64 lists
// in `sel`, each over a contiguous range of the cell's chunks (whole
// kSelChunks-chunk groups, one wave's share): the list fills positions from
// its range's first particle on, which it cannot outgrow, and its counter
// (kept at the range's first chunk) takes the atomics of few waves even in
// a cell of the whole level.

// Each run of one cell in a wave's chunks counts its packed keys inside the
// cell's margins and below them (the count below also takes the keys below
// the window round's margins, which that round counted per chunk), then
// reserves slots in its list with one 64-bit atomic (keys | below << 32)
// and writes the keys there.  A wave takes 4 chunks, a lane each quarter,
// and reads the quarters' packed keys laid end to end, four loads per lane
// in flight; a wave's keys are usually fewer than 256 and stay in registers
// between counting and writing (a dense cluster packs a good part of a
// chunk: few chunks per wave keep such waves short).
constexpr int kSelBatch = 4;
__global__ __launch_bounds__(kBlock) void k_b_select(const CellRec *__restrict__ recs,
                                                     int32_t nCells,
                                                     const uint32_t *__restrict__ first,
                                                     const uint4 *__restrict__ chunk_desc,
                                                     uint32_t max_chunks,
                                                     const float *__restrict__ ckeys,
                                                     const uint32_t *__restrict__ comp,
                                                     float *__restrict__ sel,
                                                     unsigned long long *__restrict__ sel_lists)
{
    constexpr int kW = kBlock / 64;
    constexpr int kQ = 4 * kSelChunks;  // quarters (lanes) per wave
    static_assert(kQ <= 64 && (kQ & (kQ - 1)) == 0, "a lane per chunk quarter");
    __shared__ uint32_t s_below[kW][64], s_cnt[kW][64];  // per run of one cell
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, sh = lane & 3;
    const uint32_t wave = blockIdx.x * (uint32_t)kW + (uint32_t)w;
    const uint32_t nwaves = gridDim.x * (uint32_t)kW;
    const uint32_t total = first[nCells];
    for (uint32_t q0 = wave * kSelChunks; q0 < total; q0 += nwaves * kSelChunks) {
        const uint32_t q = q0 + (uint32_t)(lane >> 2);
        const bool inq = lane < kQ && q < total;
        const uint32_t qr = inq ? q : 0u;  // in-bounds record and counts
        const uint4 d = chunk_record(chunk_desc, qr, max_chunks);
        const uint32_t nc0 = comp[8 * (size_t)qr + sh], nb0 = comp[8 * (size_t)qr + 4];
        const uint32_t cell = inq ? d.z : 0xffffffffu;  // runs end at the table's end
        bool live = inq && d.w <= 2u;
        float lo = 0.0f, hi = 0.0f;
        uint32_t begin = 0, f = 0, nch = 0;
        if (live) {
            const CellRec &r = recs[cell];
            f = first[cell];
            nch = first[cell + 1] - f;
            live = !r.found;
            lo = r.margin_l;
            hi = r.margin_r;
            begin = r.begin;
        }
        uint32_t off = 0, cnt = 0;
        if (inq) wave_share(d.x, d.y, sh, off, cnt);
        const uint32_t nc = live ? nc0 : 0u;
        const uint32_t kbase = d.x + off;  // the quarter's first packed key
        const uint32_t prev = (uint32_t)__shfl_up((int)cell, 1, 64);
        const uint64_t heads = __ballot(lane == 0 || prev != cell);
        const bool head = (heads >> lane) & 1ull;
        const int run = __popcll(heads & ((2ull << lane) - 1ull)) - 1;
        s_below[w][lane] = 0u;
        s_cnt[w][lane] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (live && sh == 0 && nb0) atomicAdd(&s_below[w][run], nb0);
        const uint32_t incl = wave_inclusive_scan(nc), ex = incl - nc;
        const uint32_t T = (uint32_t)__shfl((int)incl, 63, 64);
        // key t of the wave: quarter j = the last with ex_j <= t (it holds t)
        float v[kSelBatch], lj[kSelBatch], hj[kSelBatch];
        int rj[kSelBatch];
        auto load_batch = [&](uint32_t t0) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < kSelBatch; ++u) {
                const uint32_t t = t0 + 64u * (uint32_t)u + (uint32_t)lane;
                int j = 0;  // (lanes from kQ on hold no keys: ex = T there)
#pragma unroll
                for (int st = kQ / 2; st > 0; st >>= 1)
                    if ((uint32_t)__shfl((int)ex, j + st, 64) <= t) j += st;
                const uint32_t at =
                    (uint32_t)__shfl((int)kbase, j, 64) + (t - (uint32_t)__shfl((int)ex, j, 64));
                lj[u] = __shfl(lo, j, 64);
                hj[u] = __shfl(hi, j, 64);
                rj[u] = __shfl(run, j, 64);
                v[u] = t < T ? ckeys[at] : 0.0f;
            }
        };
        constexpr uint32_t kStep = 64u * kSelBatch;
        for (uint32_t t0 = 0; t0 < T; t0 += kStep) {
            load_batch(t0);
#pragma unroll
            for (int u = 0; u < kSelBatch; ++u) {
                if (t0 + 64u * (uint32_t)u + (uint32_t)lane < T) {
                    if (v[u] < lj[u]) atomicAdd(&s_below[w][rj[u]], 1u);
                    else if (!(v[u] > hj[u])) atomicAdd(&s_cnt[w][rj[u]], 1u);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        uint32_t slot = 0;
        if (head && live) {
            const uint32_t ns = s_cnt[w][run], nb = s_below[w][run];
            const SelLists sl(f, nch);
            const uint32_t qs = sl.start_chunk(sl.list_of_group(q0 / kSelChunks));
            const unsigned long long old =
                ns || nb ? atomicAdd(&sel_lists[qs], (unsigned long long)ns |
                                                         ((unsigned long long)nb << 32))
                         : 0ull;
            slot = begin + (qs - f) * (uint32_t)kChunk + (uint32_t)old;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (head) s_cnt[w][run] = slot;  // the run's next free slot in sel
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t t0 = 0; t0 < T; t0 += kStep) {
            if (T > kStep) load_batch(t0);  // uniform: one batch is still in registers
#pragma unroll
            for (int u = 0; u < kSelBatch; ++u) {
                if (t0 + 64u * (uint32_t)u + (uint32_t)lane < T && !(v[u] < lj[u]) &&
                    !(v[u] > hj[u]))
                    sel[atomicAdd(&s_cnt[w][rj[u]], 1u)] = v[u];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();  // the run tables are cleared for the next chunks
    }
}

// Rounds 3 and 4 of each searching cell from its gathered keys, a wave per
// cell: lane r reads list r's count (its counts below add up to the cell's
// count below the margins), the keys of all lists are read laid end to end
// and binned against the round's candidates in LDS (keys below the margins
// add to bin 0, as in the chunk rounds; keys above them are never read by
// the replay), and the round is replayed.  A cell that neither found its
// cut nor reached the pass cap runs round 4 over the same keys with the
// narrowed margins (the keys that now fall below them join the count
// below).  The record, heap cell and split are written as after any round.
constexpr int kFinBatch = 8;  // gathered keys per lane in flight
template <int LE>
__global__ __launch_bounds__(kBlock) void k_b_finish(CellRec *recs, int32_t nCells,
                                                     const uint32_t *__restrict__ first,
                                                     const float *__restrict__ sel,
                                                     const unsigned long long *__restrict__ sel_lists,
                                                     orb_cell *heap, uint64_t *trace,
                                                     int32_t *level_passes, uint8_t *exists,
                                                     int32_t *status, int32_t group)
{
    __shared__ float cand[4][kBins + 1];
    __shared__ uint32_t P[4][kBins + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    float *cw = cand[w];
    uint32_t *pw = P[w];
    auto active = [&](int32_t c) { return !recs[c].found && !recs[c].flags; };
    PassMax pm;  // lane 0
    for_cell_groups(nCells, group, active, [&](int32_t c) {
    CellRec r = recs[c];
    const uint32_t f = first[c], nch = first[c + 1] - f;
    const SelLists sl(f, nch);
    // list `lane`: its keys from position kpos on (at most the positions of
    // its chunk range, whatever the counter holds)
    uint32_t mr = 0, kpos = 0;
    uint32_t below = 0;
    if ((uint32_t)lane < sl.R) {
        const uint32_t qs = sl.start_chunk((uint32_t)lane);
        const uint32_t qn = (uint32_t)lane + 1u < sl.R ? sl.start_chunk((uint32_t)lane + 1u) : f + nch;
        const unsigned long long e = sel_lists[qs];
        kpos = r.begin + (qs - f) * (uint32_t)kChunk;
        const uint32_t span = (qn < f + nch ? r.begin + (qn - f) * (uint32_t)kChunk : r.end) - kpos;
        mr = (uint32_t)e < span ? (uint32_t)e : span;
        below = (uint32_t)(e >> 32);
    }
    below = wave_sum(below);
    const uint32_t lincl = wave_inclusive_scan(mr), lex = lincl - mr;
    const uint32_t m = (uint32_t)__shfl((int)lincl, 63, 64);
    const float target = difference_target(r.end - r.begin, r.nLeaf);
    uint64_t *trow = trace ? trace + ((size_t)r.hb + r.id) * ORB_MAX_PASSES : nullptr;
    for (int round = 0;; ++round) {
        const float lo = r.margin_l, hi = r.margin_r;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();  // the wave's last replay has read the tables
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int j = 4 * lane + i;
            cw[j] = j == 0 ? -INFINITY : candidate(j, lo, hi);
            pw[j] = 0u;
        }
        if (lane == 0) cw[kBins] = INFINITY;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const float width = hi - lo;
        const float scale = width > 0.0f ? (float)kBins / width : 0.0f;
        uint32_t bl = 0;
        for (uint32_t k0 = 0; k0 < m; k0 += 64u * kFinBatch) {
            float v[kFinBatch];  // all of a batch's loads in flight
#pragma unroll
            for (int u = 0; u < kFinBatch; ++u) {
                // key k of the cell: list j = the last with lex_j <= k
                const uint32_t k = k0 + 64u * (uint32_t)u + (uint32_t)lane;
                int j = 0;
#pragma unroll
                for (int st = 32; st > 0; st >>= 1)
                    if ((uint32_t)__shfl((int)lex, j + st, 64) <= k) j += st;
                const uint32_t at =
                    (uint32_t)__shfl((int)kpos, j, 64) + (k - (uint32_t)__shfl((int)lex, j, 64));
                v[u] = k < m ? sel[at] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kFinBatch; ++u) {
                if (k0 + 64u * (uint32_t)u + (uint32_t)lane >= m) continue;
                if (v[u] < lo) ++bl;
                else if (!(v[u] > hi)) {
                    int tj;
                    atomicAdd(&pw[bin_of<LE>(cw, v[u], lo, scale, tj)], 1u);
                }
            }
        }
        below += wave_sum(bl);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        // P[j] = the keys below the margins + bins < j
        const uint32_t s0 = pw[4 * lane], s1 = s0 + pw[4 * lane + 1];
        const uint32_t s2 = s1 + pw[4 * lane + 2], s3 = s2 + pw[4 * lane + 3];
        const uint32_t base = below + wave_inclusive_scan(s3) - s3;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        pw[4 * lane + 1] = base + s0;
        pw[4 * lane + 2] = base + s1;
        pw[4 * lane + 3] = base + s2;
        pw[4 * lane + 4] = base + s3;
        if (lane == 0) pw[0] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        const RoundOut ro = replay_round(pw, lo, hi, target, r.passes, lane, trow);
        if (round == 1 || ro.found || ro.passes >= ORB_MAX_PASSES) {  // uniform
            if (lane == 0)
                finish_round(r, c, ro, found_encoding(ro, lo, hi), recs, heap, level_passes,
                             exists, status, &pm);
            break;
        }
        r.margin_l = ro.L;
        r.margin_r = ro.R;
        r.passes = ro.passes;
    }
    });
    flush_block(pm, level_passes);
}

// The same two rounds with a block per cell, for levels of few cells (whose
// cells hold the most gathered keys): the four waves bin a quarter of the
// keys each into one LDS histogram, wave 0 prefix-sums and replays, and the
// round's outcome reaches the other waves through LDS.
constexpr int kFinBlockCells = 1024;  // levels up to this many cells
template <int LE>
__global__ __launch_bounds__(kBlock) void k_b_finish_block(CellRec *recs, int32_t nCells,
                                                           const uint32_t *__restrict__ first,
                                                           const float *__restrict__ sel,
                                                           const unsigned long long *__restrict__ sel_lists,
                                                           orb_cell *heap, uint64_t *trace,
                                                           int32_t *level_passes, uint8_t *exists,
                                                           int32_t *status)
{
    static_assert(kBins == kBlock, "a candidate and a bin per thread");
    __shared__ float cand[kBins + 1];
    __shared__ uint32_t hist[kBins], P[kBins + 1];
    __shared__ uint32_t sm[4];
    __shared__ float s_lr[2];
    __shared__ int32_t s_pf[2];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    PassMax pm;  // thread 0
    for (int32_t c = blockIdx.x; c < nCells; c += gridDim.x) {
        CellRec r = recs[c];
        if (r.found || r.flags) continue;  // uniform
        const uint32_t f = first[c], nch = first[c + 1] - f;
        const SelLists sl(f, nch);
        // every wave holds the list table (list `lane`: count, first position)
        uint32_t mr = 0, kpos = 0, bw = 0;
        if ((uint32_t)lane < sl.R) {
            const uint32_t qs = sl.start_chunk((uint32_t)lane);
            const uint32_t qn =
                (uint32_t)lane + 1u < sl.R ? sl.start_chunk((uint32_t)lane + 1u) : f + nch;
            const unsigned long long e = sel_lists[qs];
            kpos = r.begin + (qs - f) * (uint32_t)kChunk;
            const uint32_t span =
                (qn < f + nch ? r.begin + (qn - f) * (uint32_t)kChunk : r.end) - kpos;
            mr = (uint32_t)e < span ? (uint32_t)e : span;
            bw = (uint32_t)(e >> 32);
        }
        uint32_t below = wave_sum(bw);
        const uint32_t lincl = wave_inclusive_scan(mr), lex = lincl - mr;
        const uint32_t m = (uint32_t)__shfl((int)lincl, 63, 64);
        const float target = difference_target(r.end - r.begin, r.nLeaf);
        uint64_t *trow = trace ? trace + ((size_t)r.hb + r.id) * ORB_MAX_PASSES : nullptr;
        for (int round = 0;; ++round) {
            const float lo = r.margin_l, hi = r.margin_r;
            __syncthreads();  // the last round (or cell) has read the tables
            cand[t] = t ? candidate(t, lo, hi) : -INFINITY;
            hist[t] = 0u;
            if (t == 0) cand[kBins] = INFINITY;
            __syncthreads();
            const float width = hi - lo;
            const float scale = width > 0.0f ? (float)kBins / width : 0.0f;
            uint32_t bl = 0;
            for (uint32_t k0 = 0; k0 < m; k0 += (uint32_t)kBlock * kFinBatch) {
                float v[kFinBatch];  // all of a batch's loads in flight
#pragma unroll
                for (int u = 0; u < kFinBatch; ++u) {
                    // key k of the cell: list j = the last with lex_j <= k
                    const uint32_t k = k0 + (uint32_t)(kBlock * u + t);
                    int j = 0;
#pragma unroll
                    for (int st = 32; st > 0; st >>= 1)
                        if ((uint32_t)__shfl((int)lex, j + st, 64) <= k) j += st;
                    const uint32_t at = (uint32_t)__shfl((int)kpos, j, 64) +
                                        (k - (uint32_t)__shfl((int)lex, j, 64));
                    v[u] = k < m ? sel[at] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < kFinBatch; ++u) {
                    if (k0 + (uint32_t)(kBlock * u + t) >= m) continue;
                    if (v[u] < lo) ++bl;
                    else if (!(v[u] > hi)) {
                        int tj;
                        atomicAdd(&hist[bin_of<LE>(cand, v[u], lo, scale, tj)], 1u);
                    }
                }
            }
            below += block_sum(bl, sm);  // its barriers also complete the binning
            if (w == 0) {
                // P[j] = the keys below the margins + bins < j
                const uint32_t s0 = hist[4 * lane], s1 = s0 + hist[4 * lane + 1];
                const uint32_t s2 = s1 + hist[4 * lane + 2], s3 = s2 + hist[4 * lane + 3];
                const uint32_t base = below + wave_inclusive_scan(s3) - s3;
                P[4 * lane + 1] = base + s0;
                P[4 * lane + 2] = base + s1;
                P[4 * lane + 3] = base + s2;
                P[4 * lane + 4] = base + s3;
                if (lane == 0) P[0] = 0u;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
                __builtin_amdgcn_wave_barrier();
                const RoundOut ro = replay_round(P, lo, hi, target, r.passes, lane, trow);
                if (lane == 0) {
                    s_lr[0] = ro.L;
                    s_lr[1] = ro.R;
                    s_pf[0] = ro.passes;
                    s_pf[1] = (ro.found ? 1 : 0) | (ro.node << 1);
                }
            }
            __syncthreads();
            RoundOut ro;
            ro.L = s_lr[0];
            ro.R = s_lr[1];
            ro.passes = s_pf[0];
            ro.found = (s_pf[1] & 1) != 0;
            ro.node = s_pf[1] >> 1;
            if (round == 1 || ro.found || ro.passes >= ORB_MAX_PASSES) {  // uniform
                if (t == 0)
                    finish_round(r, c, ro, found_encoding(ro, lo, hi), recs, heap, level_passes,
                                 exists, status, &pm);
                break;
            }
            r.margin_l = ro.L;
            r.margin_r = ro.R;
            r.passes = ro.passes;
        }
    }
    if (t == 0) pm.flush(level_passes);
}

// ---------------------------------------------------------------------------
// weighted rounds (workload weights w(p_i), documentation/notes.tex:259-271,
// 326-336): the bins hold weight sums (64-bit) instead of counts, and the
// controller compares the weight left of each candidate with the cell's
// weight times the left share.  Every round bins every key of a searching
// cell (no window packing), so a row's total is the chunk's weight and a
// cell histogram's total the cell's weight.
// ---------------------------------------------------------------------------
template <int LE>
__global__ __launch_bounds__(kBlock) void k_bw_hist(
    const CellRec *__restrict__ recs, int32_t nCells, const uint32_t *__restrict__ first,
    const uint4 *__restrict__ chunk_desc, uint32_t max_chunks, const float *__restrict__ x,
    const float *__restrict__ y, const float *__restrict__ z, const uint32_t *__restrict__ wgt,
    uint64_t *__restrict__ chunk_hist_w)
{
    __shared__ float cand[kBins + 1];
    __shared__ unsigned long long hist[kBins];
    const int t = threadIdx.x;
    const uint32_t total = first[nCells];
    for (uint32_t q = blockIdx.x;; q += gridDim.x) {
        const uint4 d = chunk_record(chunk_desc, q, max_chunks);
        const uint32_t b = uniform(d.x), e = uniform(d.y), cell = uniform(d.z);
        const uint32_t axis = uniform(d.w);
        keep_values(b, e, cell, axis);
        if (q >= total) break;
        if (axis > 2u) continue;  // uniform: the cell is not searched
        const CellRec &r = recs[cell];
        if (r.found) continue;    // uniform: an earlier round found the cut
        const float lo = uniform(r.margin_l), hi = uniform(r.margin_r);
        store_cands(cand, t ? candidate(t, lo, hi) : -INFINITY);
        hist[t] = 0ull;
        __syncthreads();
        const float width = hi - lo;
        const float scale = width > 0.0f ? (float)kBins / width : 0.0f;
        const float lim = 0.5f - bin_tolerance(lo, hi, scale);
        const float *key = axis_ptr(x, y, z, (int)axis);
        for (uint32_t p = b + (uint32_t)t; p < e; p += kBlock) {
            const float v = key[p];
            const unsigned long long wv = wgt[p];
            int bin;
            if (v < lo) {
                bin = 0;                  // below every candidate
            } else if (v > hi) {
                bin = kBins - 1;          // above every candidate
            } else if (!bin_estimate(v, lo, scale, lim, bin)) {
                int tj;
                bin = bin_of<LE>(cand, v, lo, scale, tj);
            }
            atomicAdd(&hist[bin], wv);
        }
        __syncthreads();
        chunk_hist_w[(size_t)q * kBins + t] = hist[t];
        __syncthreads();  // the next chunk rewrites cand and hist
    }
}

// Chunk rows into cell rows (64-bit): block g takes chunks [64g, 64g + 64),
// thread t bin t, one atomic per (block, cell, bin).
__global__ __launch_bounds__(kBlock) void k_bw_reduce(const CellRec *recs, int32_t nCells,
                                                      const uint32_t *first,
                                                      const uint32_t *chunk_cell,
                                                      const uint64_t *chunk_hist_w,
                                                      uint64_t *cell_hist_w)
{
    __shared__ int32_t scell[kReduceRows];
    const uint32_t total = first[nCells];
    const int t = threadIdx.x;
    for (uint32_t q0 = blockIdx.x * (uint32_t)kReduceRows; q0 < total;
         q0 += gridDim.x * (uint32_t)kReduceRows) {
        const uint32_t nrows =
            total - q0 < (uint32_t)kReduceRows ? total - q0 : (uint32_t)kReduceRows;
        if (t < kReduceRows) {
            int32_t cl = -1;
            if ((uint32_t)t < nrows) {
                const uint32_t cell = chunk_cell[q0 + t];
                const CellRec &r = recs[cell];
                if (!(r.found || r.flags) && r.axis >= 0 && r.axis <= 2) cl = (int32_t)cell;
            }
            scell[t] = cl;
        }
        __syncthreads();
        int32_t cur = -1;
        unsigned long long acc = 0;
        for (uint32_t u = 0; u < nrows; ++u) {
            const int32_t cl = scell[u];
            if (cl != cur) {
                if (cur >= 0 && acc)
                    atomicAdd(reinterpret_cast<unsigned long long *>(cell_hist_w) +
                                  (size_t)cur * kBins + t,
                              acc);
                cur = cl;
                acc = 0;
            }
            if (cl >= 0) acc += chunk_hist_w[(size_t)(q0 + u) * kBins + t];
        }
        if (cur >= 0 && acc)
            atomicAdd(reinterpret_cast<unsigned long long *>(cell_hist_w) + (size_t)cur * kBins + t,
                      acc);
        __syncthreads();  // scell is rewritten for the next tile
    }
}

// One wave per cell: prefix sums of the cell's weight histogram in registers,
// the cell's weight (the row's total), then up to 8 passes of the controller.
__global__ __launch_bounds__(kBlock) void k_bw_replay(CellRec *recs, int32_t nCells,
                                                      uint64_t *cell_hist_w, orb_cell *heap,
                                                      uint64_t *trace, int32_t *level_passes,
                                                      uint8_t *exists, int32_t *status)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t c = blockIdx.x * 4 + w;
    if (c >= nCells) return;  // uniform per wave
    CellRec r = recs[c];
    if (r.found || r.flags || r.axis < 0 || r.axis > 2) return;
    uint64_t *row = cell_hist_w + (size_t)c * kBins;
    uint64_t h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        h[i] = row[4 * lane + i];
        row[4 * lane + i] = 0ull;  // the next round starts from zero
    }
    const uint64_t s0 = h[0], s1 = s0 + h[1], s2 = s1 + h[2], s3 = s2 + h[3];
    const uint64_t base = wave_inclusive_scan(s3) - s3;
    // P[4 lane + 1 + q] (P[j] = weight in bins below j)
    const uint64_t pv[4] = {base + s0, base + s1, base + s2, base + s3};
    const uint64_t weight = lane_value(pv[3], 63);  // every key of the cell was binned
    const float target = difference_target_w(weight, r.nLeaf);
    const RoundOut ro = replay_round_lanes(pv, r.margin_l, r.margin_r, target, r.passes, lane,
                                           trace ? trace + ((size_t)r.hb + r.id) * ORB_MAX_PASSES
                                                 : nullptr);
    if (lane != 0) return;
    // found: the partition reads the keys (no stopper counts from the bins)
    finish_round(r, c, ro, ro.found ? 1 : 0, recs, heap, level_passes, exists, status);
}

// ---------------------------------------------------------------------------
// LDS-resident subtree kernel
// ---------------------------------------------------------------------------
struct SubCell {
    orb_cell cell;
    uint32_t lb, le;   // local range in the block's LDS arrays
    int32_t valid;
    int32_t pad;
};

// Word j of a cell record held the same in every lane (0 past its end), so
// that lane j can write it.
constexpr int kCellWords = (int)(sizeof(orb_cell) / 4);
__device__ inline uint32_t cell_word(const orb_cell &c, int j)
{
    const uint32_t *p = reinterpret_cast<const uint32_t *>(&c);
    uint32_t v = 0u;
#pragma unroll
    for (int q = 0; q < kCellWords; ++q) v = j == q ? p[q] : v;
    return v;
}

// An LDS address as the buffer-to-LDS loads take it (the LDS address space).
typedef __attribute__((address_space(3))) void LdsVoid;
__device__ inline LdsVoid *lds_ptr(void *p)
{
    return (LdsVoid *)p;
}

__device__ inline void lds_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Keys of [pb, pe) in LDS below the cut (lt) or at most the cut (le), summed
// over the wave: one ballot + popcount per 64 keys, four LDS reads in flight.
__device__ inline uint32_t wave_count_below(const float *key, uint32_t pb, uint32_t pe, float cut,
                                            int le_cmp, int lane)
{
    uint32_t cnt = 0;
    for (uint32_t base = pb; base < pe; base += 256u) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t p = base + (uint32_t)(64 * u + lane);
            v[u] = p < pe ? key[p] : INFINITY;  // never below a finite cut
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool below = le_cmp ? (v[u] <= cut) : (v[u] < cut);
            cnt += (uint32_t)__popcll(__ballot(below));
        }
    }
    return cnt;
}

// The weight of the keys of [pb, pe) below the cut (weighted build), summed
// over the wave in 64 bits.
__device__ inline uint64_t wave_weight_below(const float *key, const uint32_t *wgt, uint32_t pb,
                                             uint32_t pe, float cut, int le_cmp, int lane)
{
    uint64_t s = 0;
    for (uint32_t p = pb + (uint32_t)lane; p < pe; p += 64u) {
        const float v = key[p];
        const bool below = le_cmp ? (v <= cut) : (v < cut);
        s += below ? (uint64_t)wgt[p] : 0ull;
    }
    return wave_sum(s);
}

// The total weight of [pb, pe), summed over the wave.
__device__ inline uint64_t wave_weight_total(const uint32_t *wgt, uint32_t pb, uint32_t pe, int lane)
{
    uint64_t s = 0;
    for (uint32_t p = pb + (uint32_t)lane; p < pe; p += 64u) s += wgt[p];
    return wave_sum(s);
}

// The left count (unweighted) or left weight (weighted) of [pb, pe).
template <bool W>
__device__ inline uint64_t wave_left(const float *key, const uint32_t *wgt, uint32_t pb, uint32_t pe,
                                     float cut, int le_cmp, int lane)
{
    if constexpr (W) return wave_weight_below(key, wgt, pb, pe, cut, le_cmp, lane);
    else return wave_count_below(key, pb, pe, cut, le_cmp, lane);
}

// The termination rule's difference for a left count or weight.
template <bool W>
__device__ inline int64_t left_difference(uint64_t left, float target)
{
    if constexpr (W) return difference_from_target_w(left, target);
    else return difference_from_target((uint32_t)left, target);
}

// Exact bin of key v in [lo, hi] among the 255 candidate cuts of the next 8
// bisection steps, without a table: the walk down the candidate tree
// computes each midpoint with the cut arithmetic (cell.h:74-76) and counts
// the candidates at or below v (LE = 0) or below v (LE = 1).  The
// candidates are sorted, so a node that qualifies takes its whole left
// subtree with it.
template <int LE>
__device__ inline int bin_walk(float v, float lo, float hi)
{
    int e = 0;
#pragma unroll
    for (int d = 0; d < kHistK; ++d) {
        const float m = cell_cut(lo, hi);
        const bool right = LE ? (m < v) : (m <= v);
        if (right) {
            e += 1 << (kHistK - 1 - d);
            lo = m;
        } else {
            hi = m;
        }
    }
    return e;
}

// The LDS capacity of the subtree kernel (keys of one subtree root) per
// payload: x, y, z (12 B), the two 16-bit stopper lists (4 B), with ids a
// 16-bit local index (2 B), with weights the weight (4 B) and 64-bit
// histogram slots; 155-163 KB of gfx950's 160 KiB LDS in each variant.  A
// level's cells sit within a few keys of n / 2^(l-1), so the plain
// capacity stays a little above 8192 for a level of 8192-key cells to
// qualify whole.
template <bool IDS, bool W>
struct SubCap {
    static constexpr int value = W ? 5120 : (IDS ? 7680 : 8832);
};
constexpr int kSubBlock = 1024;                 // threads of the subtree kernel
constexpr int kSubWaves = kSubBlock / 64;       // 16 histogram slots

static int32_t subtree_capacity(const Ctx &c)
{
    if (c.weighted) return SubCap<true, true>::value;
    return c.track_ids ? SubCap<true, false>::value : SubCap<false, false>::value;
}

// Every remaining level of the subtree under a small cell, in LDS: the cell's
// x, y, z (and ids' local index, weights) are loaded once, every remaining
// level is built there, and x, y, z are written back once.
//
// A depth's cells are taken up to 16 at a time (a slot each); a slot has
// 16 / (cells) waves when the depth has fewer than 16 cells.  The bisection
// controller (orbit.cpp:149-231) runs in rounds of 8 passes, as the chunk
// path's histogram rounds do: the slot's waves bin the cell's keys over the
// 255 candidate cuts of its margins into a 256-bin LDS histogram (two 16-bit
// bins per word; 64-bit bins of weight sums with W), and the slot's first
// wave prefix-sums it in place and replays up to 8 passes (replay_round),
// exact by construction.  The Hoare partition (partition.cpp:28-60) is the
// stopper closed form: stopper lists by ballot scans, the crossing K by a
// 64-way search, K swaps, the final swap.  Every barrier is reached by all
// waves the same number of times.
// IDS: particle ids move with the particles; a 16-bit local index rides along
// in LDS and the ids are permuted once at the write-back.  W: workload
// weights ride along in LDS and the histograms sum weights.
template <bool IDS, bool W>
__global__ __launch_bounds__(kSubBlock) void k_b_subtree(const CellRec *recs, int32_t nCells,
                                                         orb_cell *heap, uint8_t *exists,
                                                         uint8_t *sub, uint32_t *ranges, float *x,
                                                         float *y, float *z, int le_cmp,
                                                         uint64_t *trace, int32_t *level_passes,
                                                         int32_t *status, uint32_t *pid,
                                                         uint32_t *wgt, int32_t group)
{
    constexpr int kCap = SubCap<IDS, W>::value;
    using HistWord = typename std::conditional<W, unsigned long long, uint32_t>::type;
    constexpr int kHistWords = W ? kBins : kBins / 2;  // W: a bin per word; else two
    __shared__ float sx[kCap], sy[kCap], sz[kCap];
    __shared__ uint32_t sw[W ? kCap : 1];
    // the two stopper lists (16-bit), reused as the cell's 32-bit ids at the end
    __shared__ uint32_t spos_raw[kCap];
    __shared__ uint16_t sidx[IDS ? kCap : 1];
    uint16_t *const slpos = reinterpret_cast<uint16_t *>(spos_raw);
    uint16_t *const srpos = slpos + kCap;
    // during a histogram round the stopper lists are free: each wave queues
    // its slow keys there (and with W their weights in the second half)
    constexpr int kQueue = ((W ? kCap / 2 : kCap) / kSubWaves) & ~63;
    static_assert(kQueue >= 64 && kSubWaves * kQueue <= (W ? kCap / 2 : kCap), "queue fits");
    __shared__ HistWord shist[kSubWaves][kHistWords];
    __shared__ SubCell scells[2][kSmallLeaves];  // a depth's cells and the next's
    __shared__ float sL[kSubWaves], sR[kSubWaves];
    __shared__ int32_t sPasses[kSubWaves], sFound[kSubWaves];
    __shared__ uint64_t sWt[kSubWaves];
    __shared__ uint32_t sTL[kSubWaves], sTR[kSubWaves], sK[kSubWaves], sI[kSubWaves];
    __shared__ int32_t s_ncur, s_nvalid, s_err;
    // the block's pass maxima per level of the build slot in s_lp0, flushed
    // with one atomicMax per level when the slot changes and at the end
    __shared__ int32_t s_lpmax[64];
    __shared__ int32_t s_lp0;
    __shared__ uint64_t s_gmask;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint64_t lanes_below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    if (t < 64) s_lpmax[t] = 0;
    if (t == 0) s_lp0 = -1;
    // every wave histogram starts clear and is left clear by the round that
    // used it
    for (int i = t; i < kSubWaves * kHistWords; i += kSubBlock) (&shist[0][0])[i] = 0;
    __syncthreads();
    // groups of `group` consecutive cells (a power of two <= 64): wave 0 reads
    // the group's flags in one load and the block takes the subtree roots
    // among them, so a level with few roots (or none) costs a few loads per
    // block, not one per cell
    for (int32_t g0 = blockIdx.x * group; g0 < nCells; g0 += gridDim.x * group) {
    if (w == 0) {
        const int32_t cc = g0 + lane;
        const bool root = lane < group && cc < nCells && recs[cc].flags == 2;
        const uint64_t m = __ballot(root);
        if (lane == 0) s_gmask = m;
    }
    __syncthreads();
    uint64_t gm = s_gmask;  // uniform
    __syncthreads();        // read before wave 0 rewrites it
    while (gm) {
    const int32_t c = g0 + __builtin_ctzll(gm);
    gm &= gm - 1ull;
    const CellRec rr = recs[c];
    if ((rr.lp - (rr.lp & 63)) != s_lp0) {
        if (t < 64 && s_lp0 >= 0 && s_lpmax[t] > 0) atomicMax(&level_passes[s_lp0 + t], s_lpmax[t]);
        __syncthreads();
        if (t < 64) s_lpmax[t] = 0;
        if (t == 0) s_lp0 = rr.lp - (rr.lp & 63);
        __syncthreads();
    }
    const uint32_t begin = rr.begin, n = rr.end - rr.begin;
    const int64_t hb = rr.hb;
    const int32_t level = rr.lp & 63, lmax = rr.lmax;
    int32_t *st = status + rr.st;
    if (n > (uint32_t)kCap) {
        // the level records hand over at most kCap keys; refuse anything else
        if (t == 0) report_error(st, level, kErrSplit, rr.id);
        continue;
    }
    // the root cell's 13 words, one per lane of wave 0 (no struct copy)
    uint32_t root_w = 0u;
    if (t < kCellWords) root_w = reinterpret_cast<const uint32_t *>(heap + hb + rr.id)[t];
    {
        // the cell's arrays straight from memory into LDS, through no
        // registers: wave w takes the 64-key pieces that start at
        // 64 w + 1024 u below n, one buffer load to LDS per piece and array,
        // all in flight together.  The range check is the cell's byte count:
        // keys past n land as 0 in slots below kCap (a piece starts below
        // n <= kCap, and kCap is a multiple of 64).
        static_assert(kCap % 64 == 0, "whole 64-key pieces");
        const int nb = (int)(n * 4u);
        const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(x + begin, 0, nb, 0x00020000);
        const __amdgpu_buffer_rsrc_t ry = __builtin_amdgcn_make_buffer_rsrc(y + begin, 0, nb, 0x00020000);
        const __amdgpu_buffer_rsrc_t rz = __builtin_amdgcn_make_buffer_rsrc(z + begin, 0, nb, 0x00020000);
        for (uint32_t k0 = 64u * (uint32_t)w; k0 < n; k0 += (uint32_t)kSubBlock) {
            const int vo = (int)(4u * (k0 + (uint32_t)lane));
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rx, lds_ptr(sx + k0), 4, vo, 0, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ry, lds_ptr(sy + k0), 4, vo, 0, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rz, lds_ptr(sz + k0), 4, vo, 0, 0, 0);
            if constexpr (W) {
                const __amdgpu_buffer_rsrc_t rw =
                    __builtin_amdgcn_make_buffer_rsrc(wgt + begin, 0, nb, 0x00020000);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rw, lds_ptr(sw + k0), 4, vo, 0, 0, 0);
            }
        }
        if (IDS)
            for (uint32_t k = (uint32_t)t; k < n; k += (uint32_t)kSubBlock) sidx[k] = (uint16_t)k;
        // the loads have written LDS before the barrier below lets any wave read it
        __builtin_amdgcn_s_waitcnt(0);
    }
    int cb = 0;  // scells[cb]: this depth's cells
    if (t < kCellWords) reinterpret_cast<uint32_t *>(&scells[0][0].cell)[t] = root_w;
    if (t == 0) {
        scells[0][0].lb = 0;
        scells[0][0].le = n;
        scells[0][0].valid = 1;
        s_ncur = 1;
        s_nvalid = 1;
        s_err = 0;
    }
    __syncthreads();
    for (int32_t l = level; l <= lmax; ++l) {
        const int32_t ncur = s_ncur;
        if (ncur == 0 || s_nvalid == 0 || s_err) break;
        if (2 * ncur > kSmallLeaves) {
            // cannot happen while the root has <= kSmallLeaves leaves; refuse
            // rather than leave the subtree half built
            if (t == 0) report_error(st, l, kErrSplit, rr.id);
            break;
        }
        const int nslots = ncur >= kSubWaves ? kSubWaves : ncur;   // ncur is a power of two
        const int g = kSubWaves / nslots;                           // waves per cell
        const int nbatch = ncur >= kSubWaves ? ncur / kSubWaves : 1;
        const int slot = w / g, part = w % g, w0 = slot * g;
        for (int bt = 0; bt < nbatch; ++bt) {
            const int32_t k = bt * nslots + slot;
            const SubCell sc = scells[cb][k];
            const bool valid = sc.valid != 0;
            orb_cell cl = sc.cell;
            const uint32_t lb = sc.lb, le = sc.le, m = le - lb;
            const float *key = cl.cutAxis == 0 ? sx : (cl.cutAxis == 1 ? sy : sz);
            if (part == 0 && lane == 0) {
                scells[cb ^ 1][2 * k].valid = 0;
                scells[cb ^ 1][2 * k + 1].valid = 0;
            }
            // this wave's part of the cell
            const uint32_t seg = ((m + (uint32_t)g - 1u) / (uint32_t)g + 63u) & ~63u;
            uint32_t pb = lb + (uint32_t)part * seg, pe = pb + seg;
            if (pb > le) pb = le;
            if (pe > le) pe = le;
            float target;
            if constexpr (W) {
                // the cell's weight (count.cpp:14-17 with weights): the g waves
                // of the cell add their parts through LDS
                const uint64_t pw = valid ? wave_weight_total(sw, pb, pe, lane) : 0ull;
                if (lane == 0) sWt[w] = pw;
                __syncthreads();
                const uint64_t tot = wave_sum(lane < g ? (uint64_t)sWt[w0 + lane] : 0ull);
                target = difference_target_w(tot, cl.nLeafCells);
            } else {
                target = difference_target(m, cl.nLeafCells);
            }
            uint64_t *trow = trace ? trace + ((size_t)hb + cl.id) * ORB_MAX_PASSES : nullptr;
            float L = cl.cutMarginLeft, R = cl.cutMarginRight;
            int32_t passes = 0;
            bool found = false;
            // the slot's g waves bin their parts into the slot's one histogram
            // (its first wave's); that wave replays from it and clears it.  Two
            // barriers per round; amask (the same in every wave) holds the
            // batch's slots still searching.
            HistWord *const myhist = shist[w0];
            uint32_t amask = (uint32_t)__ballot(lane < nslots && scells[cb][bt * nslots + lane].valid);
            while (amask) {
                const bool act = ((amask >> slot) & 1u) != 0u;
                if (act) {
                    const float width = R - L;
                    const float scale = width > 0.0f ? (float)kBins / width : 0.0f;
                    const float lim = 0.5f - bin_tolerance(L, R, scale);
                    // keys outside [L, R] (most of them after the first
                    // round) are counted in registers and added to bins 0 and
                    // 255 once per wave: as atomics they would all hit the
                    // same two LDS words
                    uint32_t nbelow = 0, nabove = 0;     // wave-uniform counts
                    uint64_t wbelow = 0ull, wabove = 0ull;  // W: this lane's weights
                    // keys too close to a bin edge for the estimate (about 1%
                    // in a first round) queue in the wave's part of the free
                    // stopper-list space and are walked 64 at a time, not one
                    // divergent walk per load that holds one
                    float *const qk = reinterpret_cast<float *>(spos_raw) + w * kQueue;
                    uint32_t *const qw = spos_raw + (W ? kCap / 2 : 0) + w * kQueue;  // W: weights
                    uint32_t qn = 0;  // wave-uniform
                    auto drain = [&]() {
                        lds_wave_sync();  // the queue's stores before its loads
                        for (uint32_t k = (uint32_t)lane; k < qn; k += 64u) {
                            const float vv = qk[k];
                            const int b = le_cmp ? bin_walk<1>(vv, L, R) : bin_walk<0>(vv, L, R);
                            if constexpr (W) atomicAdd(&myhist[b], (unsigned long long)qw[k]);
                            else atomicAdd(&myhist[b >> 1], (b & 1) ? 0x10000u : 1u);
                        }
                        lds_wave_sync();  // its loads before the next stores
                        qn = 0;
                    };
                    // 256 keys per call (64 per load); a full group (every
                    // group but the part's last) needs no bounds checks
                    auto group = [&](uint32_t base, auto full) __attribute__((always_inline)) {
                        constexpr bool kFull = decltype(full)::value;
                        float v[4];
                        uint32_t wv[W ? 4 : 1];
                        bool in[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {  // four LDS reads in flight
                            const uint32_t p = base + (uint32_t)(64 * u + lane);
                            in[u] = kFull || p < pe;
                            v[u] = in[u] ? key[p] : 0.0f;
                            if constexpr (W) wv[u] = in[u] ? sw[p] : 0u;
                        }
                        // the four bins first, without branches, then the
                        // queueing of slow keys, then the atomics of the other
                        // keys in the window: no key waits for the previous
                        // key's atomic
                        int bin[4];
                        bool slow[4], fast[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float vv = v[u];
                            const bool below = vv < L, above = vv > R;  // past every candidate
                            // bin_estimate's test; the clamp only keeps the
                            // conversion of keys outside [L, R] in range
                            const float ef = (vv - L) * scale;
                            bin[u] = (int)fminf(fmaxf(ef, 0.0f), 256.0f);
                            const bool est = fabsf(__builtin_amdgcn_fractf(ef) - 0.5f) < lim;
                            const bool inwin = in[u] && !below && !above;
                            slow[u] = inwin && !est;
                            fast[u] = inwin && est;
                            if constexpr (W) {
                                wbelow += in[u] && below ? (uint64_t)wv[u] : 0ull;
                                wabove += in[u] && above ? (uint64_t)wv[u] : 0ull;
                            } else {
                                nbelow += (uint32_t)__popcll(__ballot(in[u] && below));
                                nabove += (uint32_t)__popcll(__ballot(in[u] && above));
                            }
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const uint64_t sm = __ballot(slow[u]);
                            if (sm) {  // wave-uniform
                                if (qn + 64u > (uint32_t)kQueue) drain();
                                if (slow[u]) {
                                    const uint32_t at = qn + (uint32_t)__popcll(sm & lanes_below);
                                    qk[at] = v[u];
                                    if constexpr (W) qw[at] = wv[u];
                                }
                                qn += (uint32_t)__popcll(sm);
                            }
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (!fast[u]) continue;
                            if constexpr (W) {
                                atomicAdd(&myhist[bin[u]], (unsigned long long)wv[u]);
                            } else {
                                // bin b: half b & 1 of word b >> 1
                                atomicAdd(&myhist[bin[u] >> 1], 1u << ((bin[u] & 1) << 4));
                            }
                        }
                    };
                    uint32_t base = pb;
                    for (; base + 256u <= pe; base += 256u) group(base, std::true_type{});
                    if (base < pe) group(base, std::false_type{});
                    if (qn) drain();
                    if constexpr (W) {
                        const uint64_t sb = wave_sum(wbelow), sa = wave_sum(wabove);
                        if (lane == 0) {
                            if (sb) atomicAdd(&myhist[0], (unsigned long long)sb);
                            if (sa) atomicAdd(&myhist[kBins - 1], (unsigned long long)sa);
                        }
                    } else if (lane == 0) {
                        // bin 0: the low half of word 0; bin 255: the high half of the last word
                        if (nbelow) atomicAdd(&myhist[0], nbelow);
                        if (nabove) atomicAdd(&myhist[kHistWords - 1], nabove << 16);
                    }
                }
                __syncthreads();  // every histogram of the round is complete
                if (act && part == 0) {
                    // the slot's histogram read and cleared, its prefix sums
                    // in registers, then up to 8 passes of the controller on them
                    RoundOut ro;
                    if constexpr (W) {
                        uint64_t h[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) h[i] = myhist[4 * lane + i];
#pragma unroll
                        for (int i = 0; i < 4; ++i) myhist[4 * lane + i] = 0ull;  // this lane's bins
                        const uint64_t s0 = h[0], s1 = s0 + h[1], s2 = s1 + h[2], s3 = s2 + h[3];
                        const uint64_t base = wave_inclusive_scan(s3) - s3;
                        // P[4 lane + 1 + q] (P[j] = weight in bins below j)
                        const uint64_t pv[4] = {base + s0, base + s1, base + s2, base + s3};
                        ro = replay_round_lanes(pv, L, R, target, passes, lane, trow);
                    } else {
                        // two 16-bit bins per word: counts <= kCap
                        const uint2 hw = reinterpret_cast<const uint2 *>(myhist)[lane];
                        reinterpret_cast<uint2 *>(myhist)[lane] = make_uint2(0u, 0u);  // this lane's bins
                        const uint32_t w01 = hw.x, w23 = hw.y;
                        const uint32_t s0 = w01 & 0xffffu, s1 = s0 + (w01 >> 16);
                        const uint32_t s2 = s1 + (w23 & 0xffffu), s3 = s2 + (w23 >> 16);
                        const uint32_t base = wave_inclusive_scan(s3) - s3;  // <= kCap
                        // P[4 lane + 1 + q] (P[j] = keys in bins below j)
                        const uint32_t pv[4] = {base + s0, base + s1, base + s2, base + s3};
                        ro = replay_round_lanes(pv, L, R, target, passes, lane, trow);
                    }
                    if (lane == 0) {
                        sL[slot] = ro.L;
                        sR[slot] = ro.R;
                        sPasses[slot] = ro.passes;
                        sFound[slot] = ro.found ? 1 : 0;
                    }
                }
                __syncthreads();  // the slots' new margins out, every histogram clear
                {
                    // slots still searching: one read per lane, not one per slot
                    const bool on = ((amask >> lane) & 1u) != 0u && lane < kSubWaves;
                    const bool more = on && !sFound[lane] && sPasses[lane] < ORB_MAX_PASSES;
                    amask = (uint32_t)__ballot(more);
                }
                if (act) {
                    L = sL[slot];
                    R = sR[slot];
                    passes = sPasses[slot];
                    found = sFound[slot] != 0;
                }
            }
            // ---- partition: stopper lists (each wave one contiguous part)
            const float cut = cell_cut(L, R);
            uint32_t tl = 0, tr = 0;
            // 256 keys per call; a full group (every group but the part's
            // last) needs no bounds checks
            auto count_group = [&](uint32_t base, auto full) __attribute__((always_inline)) {
                constexpr bool kFull = decltype(full)::value;
                float v[4];
                bool in[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t p = base + (uint32_t)(64 * u + lane);
                    in[u] = kFull || p < pe;
                    v[u] = in[u] ? key[p] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    tl += (uint32_t)__popcll(__ballot(in[u] && v[u] >= cut));
                    tr += (uint32_t)__popcll(__ballot(in[u] && v[u] <= cut));
                }
            };
            if (valid) {
                uint32_t base = pb;
                for (; base + 256u <= pe; base += 256u) count_group(base, std::true_type{});
                if (base < pe) count_group(base, std::false_type{});
            }
            if (lane == 0) {
                sTL[w] = tl;
                sTR[w] = tr;
            }
            __syncthreads();
            // the slot's waves' stopper counts: totals and this wave's offsets,
            // by a scan across lanes (lane q holds wave w0 + q)
            uint32_t oL, oR, TL, TR;
            {
                const uint32_t a = lane < g ? sTL[w0 + lane] : 0u;
                const uint32_t b2 = lane < g ? sTR[w0 + lane] : 0u;
                const uint32_t ia = wave_inclusive_scan(a), ib = wave_inclusive_scan(b2);
                TL = (uint32_t)__shfl((int)ia, 63);
                TR = (uint32_t)__shfl((int)ib, 63);
                oL = (uint32_t)__shfl((int)(ia - a), part);
                oR = (uint32_t)__shfl((int)(ib - b2), part);
            }
            uint16_t *const Lp = slpos + lb;
            uint16_t *const Rp = srpos + lb;
            auto list_group = [&](uint32_t base, auto full) __attribute__((always_inline)) {
                constexpr bool kFull = decltype(full)::value;
                float v[4];
                bool in[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t p = base + (uint32_t)(64 * u + lane);
                    in[u] = kFull || p < pe;
                    v[u] = in[u] ? key[p] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t p = base + (uint32_t)(64 * u + lane);
                    const bool isL = in[u] && v[u] >= cut, isR = in[u] && v[u] <= cut;
                    const uint64_t mL = __ballot(isL), mR = __ballot(isR);
                    if (isL) Lp[oL + __popcll(mL & lanes_below)] = (uint16_t)(p - lb);
                    if (isR) Rp[oR + __popcll(mR & lanes_below)] = (uint16_t)(p - lb);
                    oL += (uint32_t)__popcll(mL);
                    oR += (uint32_t)__popcll(mR);
                }
            };
            if (valid) {
                uint32_t base = pb;
                for (; base + 256u <= pe; base += 256u) list_group(base, std::true_type{});
                if (base < pe) list_group(base, std::false_type{});
            }
            __syncthreads();
            const bool bad = valid && (m == 0 || TL == 0);
            if (valid && !bad && part == 0) {
                // K = #{t : L[t] < R[TR-1-t]} (true then false in t)
                uint32_t lo = 0, hi = TL < TR ? TL : TR;
                while (hi - lo > 64u) {
                    const uint32_t span = hi - lo;
                    const uint32_t tp = lo + (uint32_t)(((lane + 1) * span) / 65u);
                    const uint64_t mk = __ballot(Lp[tp] < Rp[TR - 1 - tp]);
                    if (mk == 0ull) {
                        hi = lo + span / 65u;
                    } else {
                        const int j = 63 - __clzll((long long)mk);
                        const uint32_t tj = lo + (uint32_t)(((j + 1) * span) / 65u);
                        const uint32_t tn = j < 63 ? lo + (uint32_t)(((j + 2) * span) / 65u) : hi;
                        lo = tj + 1;
                        hi = tn;
                    }
                }
                {
                    const uint32_t tp = lo + (uint32_t)lane;
                    lo += (uint32_t)__popcll(__ballot(tp < hi && Lp[tp] < Rp[TR - 1 - tp]));
                }
                const uint32_t K = lo;
                uint32_t i;
                if (K == 0) {
                    i = Lp[0];
                } else {
                    i = Rp[TR - K];
                    if (K < TL && Lp[K] < i) i = Lp[K];
                }
                if (lane == 0) {
                    sK[slot] = K;
                    sI[slot] = i;
                }
            }
            __syncthreads();
            if (valid && !bad) {
                const uint32_t K = sK[slot];
                for (uint32_t tt = (uint32_t)(part * 64 + lane); tt < K; tt += (uint32_t)(g * 64)) {
                    const uint32_t a = lb + Lp[tt], b2 = lb + Rp[TR - 1 - tt];
                    const float ax = sx[a], ay = sy[a], az = sz[a];
                    sx[a] = sx[b2];
                    sy[a] = sy[b2];
                    sz[a] = sz[b2];
                    sx[b2] = ax;
                    sy[b2] = ay;
                    sz[b2] = az;
                    if constexpr (W) {
                        const uint32_t wa = sw[a];
                        sw[a] = sw[b2];
                        sw[b2] = wa;
                    }
                    if (IDS) {
                        const uint16_t ia = sidx[a];
                        sidx[a] = sidx[b2];
                        sidx[b2] = ia;
                    }
                }
            }
            __syncthreads();
            if (valid && part == 0) {
                // the slot's first wave, all lanes with the same values: the
                // cell and its children are written a word per lane
                cl.cutMarginLeft = L;
                cl.cutMarginRight = R;
                cl.foundCut = found;
                const uint32_t cw = cell_word(cl, lane);
                if (lane < kCellWords) reinterpret_cast<uint32_t *>(heap + hb + cl.id)[lane] = cw;
                if (lane == 0) atomicMax(&s_lpmax[l], passes);  // LDS: the slots' waves finish together

                orb_cell lc, rc;
                uint32_t *wl = reinterpret_cast<uint32_t *>(&lc);
                uint32_t *wr = reinterpret_cast<uint32_t *>(&rc);
#pragma unroll
                for (int j = 0; j < kCellWords; ++j) wl[j] = wr[j] = 0u;
                const uint32_t i = sI[slot];
                // the reference splits a level's cells before it partitions any
                const bool split_ok = right_child_id(cl.id) < rr.capr && cell_split(cl, lc, rc);
                if (!split_ok || bad) {
                    if (lane == 0) {
                        s_err = 1;
                        report_error(st, l, split_ok ? kErrPartition : kErrSplit, cl.id);
                    }
                } else {
                    if (lb + i != le - 1 && lane < 5) {
                        // Hoare's last step, an array per lane
                        const uint32_t a = lb + i, b2 = le - 1;
                        if (lane < 3) {
                            float *arr = lane == 0 ? sx : (lane == 1 ? sy : sz);
                            const float va = arr[a];
                            arr[a] = arr[b2];
                            arr[b2] = va;
                        } else if (lane == 3) {
                            if constexpr (W) {
                                const uint32_t wa = sw[a];
                                sw[a] = sw[b2];
                                sw[b2] = wa;
                            }
                        } else if (IDS) {
                            const uint16_t ia = sidx[a];
                            sidx[a] = sidx[b2];
                            sidx[b2] = ia;
                        }
                    }
                    const int64_t gl = hb + lc.id, gr = hb + rc.id;
                    // lanes [0, 13): the left child's words, [32, 45): the right's
                    const int half = lane >> 5, wj = lane & 31;
                    const uint32_t kw = cell_word(half ? rc : lc, wj);
                    SubCell *nx = scells[cb ^ 1];
                    if (wj < kCellWords) {
                        reinterpret_cast<uint32_t *>(heap + (half ? gr : gl))[wj] = kw;
                        reinterpret_cast<uint32_t *>(&nx[2 * k + half].cell)[wj] = kw;
                    }
                    if (lane == 0) {
                        exists[gl] = 1;
                        exists[gr] = 1;
                        sub[gl] = 1;
                        sub[gr] = 1;
                        ranges[2 * gl] = begin + lb;
                        ranges[2 * gl + 1] = begin + lb + i;
                        ranges[2 * gr] = begin + lb + i;
                        ranges[2 * gr + 1] = begin + le;
                        nx[2 * k].lb = lb;
                        nx[2 * k].le = lb + i;
                        nx[2 * k].valid = lc.nLeafCells >= 2;
                        nx[2 * k + 1].lb = lb + i;
                        nx[2 * k + 1].le = le;
                        nx[2 * k + 1].valid = rc.nLeafCells >= 2;
                    }
                }
            }
            __syncthreads();  // the next batch reuses the position lists and the slots
        }
        const int32_t nn = 2 * ncur;  // <= kSmallLeaves = 64 (checked above)
        if (w == 0) {
            const uint64_t vm = __ballot(lane < nn && scells[cb ^ 1][lane].valid);
            if (lane == 0) {
                s_ncur = nn;
                s_nvalid = (int32_t)__popcll(vm);
            }
        }
        cb ^= 1;
        __syncthreads();
    }
    for (uint32_t k = (uint32_t)t; k < n; k += kSubBlock) {
        x[begin + k] = sx[k];
        y[begin + k] = sy[k];
        z[begin + k] = sz[k];
        if constexpr (W) wgt[begin + k] = sw[k];
    }
    if (IDS) {
        // the stopper lists are free: stage the cell's ids, then permute them
        for (uint32_t k = (uint32_t)t; k < n; k += kSubBlock) spos_raw[k] = pid[begin + k];
        __syncthreads();
        for (uint32_t k = (uint32_t)t; k < n; k += kSubBlock) pid[begin + k] = spos_raw[sidx[k]];
    }
    __syncthreads();  // the next cell reuses the LDS arrays
    }
    }
    if (t < 64 && s_lp0 >= 0 && s_lpmax[t] > 0) atomicMax(&level_passes[s_lp0 + t], s_lpmax[t]);
}

// ---------------------------------------------------------------------------
// orchestration
// ---------------------------------------------------------------------------
static int ensure_state(Ctx &c, int32_t levels_mode, bool *complete)
{
    if (!c.batched) c.batched = new (std::nothrow) BatchedState;
    if (!c.batched) return ORB_ERR_NOMEM;
    BatchedState &b = *c.batched;
    if (b.levels_mode != levels_mode || b.d != c.d) ORB_TRY(make_schedule(c, b, levels_mode, complete));
    else *complete = b.complete;
    b.complete = *complete;
    ORB_TRY(b.heap.alloc((size_t)c.cap));
    ORB_TRY(b.exists.alloc((size_t)c.cap));
    ORB_TRY(b.sub.alloc((size_t)c.cap));
    ORB_TRY(b.chunk_hist.alloc(kHistRegions * (size_t)(c.max_chunks + 1) * kBins));
    // a fresh (or grown, hence reallocated) buffer is cleared before use
    if (b.cell_hist.n < (size_t)c.max_level_cells * kBins) b.hist_clean = false;
    ORB_TRY(b.cell_hist.alloc((size_t)c.max_level_cells * kBins));
    ORB_TRY(b.level_passes.alloc(64));
    ORB_TRY(b.hpasses.alloc(64));
    if (c.trace) ORB_TRY(b.trace.alloc((size_t)c.cap * ORB_MAX_PASSES));
    if (c.weighted) {
        ORB_TRY(b.chunk_hist_w.alloc((size_t)(c.max_chunks + 1) * kBins));
        if (b.cell_hist_w.n < (size_t)c.max_level_cells * kBins) b.hist_w_clean = false;
        ORB_TRY(b.cell_hist_w.alloc((size_t)c.max_level_cells * kBins));
    }
    if (!c.weighted) ORB_TRY(c.sel.alloc((size_t)c.stride));  // rounds 3 and 4
    // every buffer the launches touch exists before a capture starts
    ORB_TRY(preallocate_scratch(c));
    return ORB_OK;
}

int32_t subtree_max_keys(const Ctx &c)
{
    const int32_t cap = subtree_capacity(c);
    return c.sub_max_keys < 0 ? cap : (c.sub_max_keys > cap ? cap : c.sub_max_keys);
}

static void drop_graph(BatchedState &b)
{
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    b.gsig.clear();
}

// What a captured build holds fixed: the settings that choose its launches,
// the launch shapes, and the address of every device buffer it touches (a
// buffer freed and allocated again -- orb_enable_ids off and on, a level
// table grown by a service call with more cells -- has a new address, and a
// replay through the old one would write freed memory).
static std::vector<uint64_t> graph_signature(const Ctx &c, const BatchedState &b,
                                             int32_t levels_mode)
{
    std::vector<uint64_t> s = {(uint64_t)levels_mode, (uint64_t)c.cmp, (uint64_t)c.part,
                               (uint64_t)c.trace, (uint64_t)c.track_ids, (uint64_t)c.weighted,
                               (uint64_t)c.n,
                               (uint64_t)c.d, (uint64_t)c.max_chunks,
                               (uint64_t)c.max_level_cells, (uint64_t)(int64_t)c.sub_max_keys};
    const void *bufs[] = {c.pos.p, c.tmp.p, c.pid.p, c.pid_tmp.p, c.ranges.p, c.recs.p,
                          c.cell_nchunks.p, c.cell_first.p, c.chunk_cell.p, c.chunk_desc.p,
                          c.chunk_ltle.p, c.chunk_cnt.p, c.chunk_pre.p, c.scan_partials.p,
                          c.packed.p, c.masks.p, c.cell_cand.p, c.splits.p, c.chunk_comp.p,
                          c.sel.p, c.sel_lists.p, c.swap_plan.p, c.status.p, c.stamps.p, c.cnt.p, b.level_ids.p, b.heap.p, b.exists.p,
                          b.sub.p, b.chunk_hist.p, b.cell_hist.p, b.trace.p, b.level_passes.p,
                          c.wt.p, c.wt_tmp.p, b.chunk_hist_w.p, b.cell_hist_w.p};
    for (const void *p : bufs) s.push_back((uint64_t)(uintptr_t)p);
    s.push_back((uint64_t)c.tune_chunk_blocks.size());
    for (int32_t v : c.tune_chunk_blocks) s.push_back((uint64_t)(int64_t)v);
    s.push_back((uint64_t)c.tune_sub_blocks.size());
    for (int32_t v : c.tune_sub_blocks) s.push_back((uint64_t)(int64_t)v);
    return s;
}

int enqueue_level(Ctx &c, const LevelEntry *entries, int32_t nCells, orb_cell *heap,
                  uint8_t *exists, uint8_t *sub, uint16_t *chunk_hist, uint32_t *cell_hist,
                  uint64_t *trace, int32_t *level_passes, int32_t *status, bool any_subtree,
                  bool full_cover, int32_t level, uint64_t *chunk_hist_w, uint64_t *cell_hist_w,
                  const PendingFinal *fin_in, bool defer_final)
{
    if (c.weighted && (!chunk_hist_w || !cell_hist_w)) return ORB_ERR_STATE;
    if (!c.weighted && (c.sel.n < (size_t)c.stride || c.sel_lists.n < (size_t)c.max_chunks + 1))
        return ORB_ERR_STATE;  // rounds 3 and 4 gather into these
    const int32_t sub_max = subtree_max_keys(c);
    const bool use_subtree = any_subtree && c.part == ORB_PART_HOARE && sub_max > 0;
    const int32_t cgroup = cell_group(nCells);
    // Chunk order: the first round walks the level's chunks from the last, so
    // it starts with the chunks the previous level's swap wrote last (still in
    // the memory-side cache), and the window round walks them from the first,
    // starting with the chunks the first round read last.  (Measured at C2:
    // -0.26 ms per build against both forward; reversing the swap on
    // alternate levels as well gained less.)
    const int rev = 1;
    const int le = c.cmp == ORB_CMP_LE ? 1 : 0;
    c.cur_level = level;  // chunk_grid() and the subtree grid read the level's shapes
    struct LevelScope {
        Ctx &c;
        ~LevelScope() { c.cur_level = -1; }
    } scope{c};
    // the level's first launch also runs the previous level's last row swaps
    // (extra blocks: the records it rewrites are not what they read)
    PendingFinalDev fin{nullptr, 0, c.splits.p, c.x(), c.y(), c.z(), c.ids(), c.weights()};
    if (fin_in && fin_in->nCells > 0) {
        fin.entries = fin_in->entries;
        fin.nCells = fin_in->nCells;
    }
    if (nCells < kRecsScan) {
        const unsigned gf = (unsigned)((fin.nCells + kRecsScan - 1) / kRecsScan);
        hipLaunchKernelGGL(k_b_recs_scan, dim3(1 + gf), dim3(kRecsScan), 0, c.stream, entries, nCells,
                           heap, c.ranges.p, sub, use_subtree ? 1 : 0, (uint32_t)sub_max, c.recs.p,
                           c.cell_first.p, fin);
        ORB_HIP(hipGetLastError());
    } else {
        unsigned g1 = (unsigned)((nCells + 1 + kBlock - 1) / kBlock);
        const unsigned gf = (unsigned)((fin.nCells + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_b_recs, dim3(g1 + gf), dim3(kBlock), 0, c.stream, entries, nCells, heap,
                           c.ranges.p, sub, use_subtree ? 1 : 0, (uint32_t)sub_max, c.recs.p,
                           c.cell_nchunks.p, g1, fin);
        ORB_HIP(hipGetLastError());
        ORB_TRY(scan_exclusive_u32(c, c.cell_nchunks.p, c.cell_first.p, (size_t)nCells + 1));
    }
    {
        // the chunk table, and the first round's candidate tables (not used
        // by weighted builds)
        const int tb = timing_mark(c);
        const unsigned fill_blocks =
            (unsigned)std::min<int64_t>(16384, std::max<int64_t>(1, (c.max_chunks + kBlock - 1) / kBlock));
        const unsigned cand_blocks = c.weighted ? 0u : cell_group_blocks(nCells, cgroup);
        hipLaunchKernelGGL(k_b_tables, dim3(fill_blocks + cand_blocks), dim3(kBlock), 0, c.stream,
                           c.cell_first.p, nCells, c.recs.p, c.chunk_cell.p, c.chunk_desc.p,
                           (uint32_t)c.max_chunks, c.cell_cand.p, cgroup, fill_blocks);
        ORB_HIP(hipGetLastError());
        ORB_TRY(timing_pair(c, tb, kTimeCount));
    }
    // tie rows: only the stable partition's stopper counts read them
    uint16_t *chunk_tie =
        c.part == ORB_PART_STABLE_LE ? chunk_hist + (size_t)(c.max_chunks + 1) * kBins : nullptr;
    uint16_t *later_hist = chunk_hist + 2 * (size_t)(c.max_chunks + 1) * kBins;
    if (use_subtree) {
        // one block fits a CU (LDS): four blocks per CU by default, each
        // taking groups of consecutive cells; a group holds about
        // nCells / 1024 cells (1 to 64), so a level of subtree roots spreads
        // over every CU and a level with none is read in few loads
        int32_t sb = 1024;
        if (level >= 0 && level < (int32_t)c.tune_sub_blocks.size() &&
            c.tune_sub_blocks[(size_t)level] > 0)
            sb = c.tune_sub_blocks[(size_t)level];
        int32_t group = 1;
        while (group < 64 && (int64_t)group * 2 * 1024 <= (int64_t)nCells) group *= 2;
        const int32_t ngroups = (nCells + group - 1) / group;
        const unsigned gs = (unsigned)(ngroups < sb ? ngroups : sb);
        auto subtree = c.track_ids ? (c.weighted ? k_b_subtree<true, true> : k_b_subtree<true, false>)
                                   : (c.weighted ? k_b_subtree<false, true> : k_b_subtree<false, false>);
        hipLaunchKernelGGL(subtree, dim3(gs), dim3(kSubBlock), 0, c.stream, c.recs.p, nCells, heap,
                           exists, sub, c.ranges.p, c.x(), c.y(), c.z(), le, trace, level_passes,
                           status, c.ids(), c.weights(), group);
        ORB_HIP(hipGetLastError());
    }
    const unsigned gq = chunk_grid(c);
    const unsigned gr = (unsigned)((gq + kReduceRows - 1) / kReduceRows);  // a block per 64-row tile
    const unsigned gw = (unsigned)((nCells + 3) / 4);
    if (c.weighted) {
        // weight-sum rounds over every key, then the partition from the keys
        for (int r = 0; r < kRounds; ++r) {
            const int tb = timing_mark(c);
            if (le)
                hipLaunchKernelGGL(k_bw_hist<1>, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p,
                                   nCells, c.cell_first.p, c.chunk_desc.p, (uint32_t)c.max_chunks,
                                   c.x(), c.y(), c.z(), c.weights(), chunk_hist_w);
            else
                hipLaunchKernelGGL(k_bw_hist<0>, dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p,
                                   nCells, c.cell_first.p, c.chunk_desc.p, (uint32_t)c.max_chunks,
                                   c.x(), c.y(), c.z(), c.weights(), chunk_hist_w);
            ORB_HIP(hipGetLastError());
            ORB_TRY(timing_pair(c, tb, kTimeCount));
            hipLaunchKernelGGL(k_bw_reduce, dim3(gr), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                               c.cell_first.p, c.chunk_cell.p, chunk_hist_w, cell_hist_w);
            hipLaunchKernelGGL(k_bw_replay, dim3(gw), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                               cell_hist_w, heap, trace, level_passes, exists, status);
            ORB_HIP(hipGetLastError());
        }
        const int tp = timing_mark(c);
        if (c.part == ORB_PART_HOARE) {
            ORB_TRY(launch_hoare(c, nCells, status, nullptr, 0, defer_final));
        } else {
            ORB_TRY(invalidate_stamps(c, nCells));  // the stopper counts read the keys
            ORB_TRY(launch_stable_le(c, nCells, status, full_cover, nullptr));
        }
        ORB_TRY(timing_pair(c, tp, kTimePartition));
        return ORB_OK;
    }
    for (int r = 0; r < 2; ++r) {
        const int tb = timing_mark(c);
        // rounds 1 and 2 over every key of the searching cells; the second
        // also packs the keys inside its margins, per chunk quarter
#define ORB_HIST(L)                                                                              \
    hipLaunchKernelGGL((k_b_hist<L>), dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p, nCells,     \
                       c.cell_first.p, c.chunk_desc.p, (uint32_t)c.max_chunks, c.x(), c.y(), c.z(), \
                       c.cell_cand.p, chunk_hist, chunk_tie, rev)
#define ORB_HIST_WINDOW(L)                                                                       \
    hipLaunchKernelGGL((k_b_hist_window<L>), dim3(gq), dim3(kBlock), 0, c.stream, c.recs.p,      \
                       nCells, c.cell_first.p, c.chunk_desc.p, (uint32_t)c.max_chunks, c.x(), c.y(), \
                       c.z(), c.cell_cand.p, later_hist, c.packed.p, c.chunk_comp.p, c.masks.p, !rev)
        if (le) {
            if (r == 0) ORB_HIST(1);
            else ORB_HIST_WINDOW(1);
        } else {
            if (r == 0) ORB_HIST(0);
            else ORB_HIST_WINDOW(0);
        }
#undef ORB_HIST
#undef ORB_HIST_WINDOW
        ORB_HIP(hipGetLastError());
        ORB_TRY(timing_pair(c, tb, kTimeCount));
        hipLaunchKernelGGL(k_b_hist_reduce, dim3(gr), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                           c.cell_first.p, c.chunk_cell.p, r == 0 ? chunk_hist : later_hist,
                           cell_hist);
        hipLaunchKernelGGL(k_b_replay, dim3(cell_group_blocks(nCells, cgroup)), dim3(kBlock), 0,
                           c.stream, c.recs.p, nCells, cell_hist, heap, trace, level_passes, exists,
                           status, r == 0 ? c.cell_cand.p : nullptr, c.cell_first.p,
                           r == 1 ? reinterpret_cast<unsigned long long *>(c.sel_lists.p) : nullptr,
                           cgroup);
        ORB_HIP(hipGetLastError());
    }
    {
        // rounds 3 and 4 from the packed keys inside the margins, gathered
        // per cell once
        const int tb = timing_mark(c);
        const unsigned gs = (gq + 4 * kSelChunks - 1) / (4 * kSelChunks);
        unsigned long long *lists = reinterpret_cast<unsigned long long *>(c.sel_lists.p);
        hipLaunchKernelGGL(k_b_select, dim3(gs), dim3(kBlock), 0, c.stream, c.recs.p, nCells,
                           c.cell_first.p, c.chunk_desc.p, (uint32_t)c.max_chunks, c.packed.p,
                           c.chunk_comp.p, c.sel.p, lists);
        ORB_HIP(hipGetLastError());
        if (nCells <= kFinBlockCells) {  // a block per cell
            auto finish = le ? k_b_finish_block<1> : k_b_finish_block<0>;
            hipLaunchKernelGGL(finish, dim3((unsigned)std::max<int32_t>(nCells, 1)), dim3(kBlock), 0,
                               c.stream, c.recs.p, nCells, c.cell_first.p, c.sel.p, lists, heap,
                               trace, level_passes, exists, status);
        } else {
            auto finish = le ? k_b_finish<1> : k_b_finish<0>;
            hipLaunchKernelGGL(finish, dim3(cell_group_blocks(nCells, cgroup)), dim3(kBlock), 0,
                               c.stream, c.recs.p, nCells, c.cell_first.p, c.sel.p, lists, heap,
                               trace, level_passes, exists, status, cgroup);
        }
        ORB_HIP(hipGetLastError());
        ORB_TRY(timing_pair(c, tb, kTimeCount));
    }
    const int tp = timing_mark(c);
    HistCounts hc;
    hc.hist = chunk_hist;
    hc.tie = chunk_tie;
    hc.le = le;
    hc.pack = c.packed.p;
    hc.comp = c.chunk_comp.p;
    if (c.part == ORB_PART_HOARE) ORB_TRY(launch_hoare(c, nCells, status, &hc, 0, defer_final));
    else ORB_TRY(launch_stable_le(c, nCells, status, full_cover, &hc));
    ORB_TRY(timing_pair(c, tp, kTimePartition));
    return ORB_OK;
}

// All launches of one batched build, in stream order; no allocation, no
// host synchronisation, no memset: the sequence can be captured.
static int enqueue_build(Ctx &c, int32_t levels_mode)
{
    BatchedState &b = *c.batched;
    uint64_t *trace = c.trace ? b.trace.p : nullptr;
    const int64_t trace_n = c.trace ? c.cap * ORB_MAX_PASSES : 0;
    const orb_cell root = cell_root(c.d);
    hipLaunchKernelGGL(k_b_init, dim3(1024), dim3(kBlock), 0, c.stream, b.heap.p, b.exists.p,
                       b.sub.p, (int64_t)c.cap, root, trace, trace_n, b.level_passes.p, c.ranges.p,
                       (uint32_t)c.n, c.status.p);
    ORB_HIP(hipGetLastError());
    const int32_t lmax = levels_mode == ORB_LEVELS_REFERENCE ? b.D - 1 : b.D;
    // chunks each level expects, for the chunk kernels' grids: cells split
    // by count, so a level's cell sizes follow from n and d (orb_level_shape,
    // within the termination rule's +-2 a split); below a level whose every
    // cell goes to the subtree kernel nothing is left.  Weighted builds split
    // by weight: unknown.
    c.level_chunk_hint.assign((size_t)lmax + 2, -1);
    if (!c.weighted) {
        const int32_t sub_max = subtree_max_keys(c);
        bool done = false;
        for (int32_t l = 1; l <= lmax; ++l) {
            int64_t lo = 0, hi = 0;
            const int cells = orb_level_shape(c.n, c.d, l, &lo, &hi);
            if (cells < 0) break;
            const int64_t max_leaf = ((int64_t)c.d + (int64_t(1) << (l - 1)) - 1) >> (l - 1);
            const bool all_sub = b.complete && c.part == ORB_PART_HOARE && sub_max > 0 &&
                                 hi + 4 * l <= sub_max && max_leaf <= kSmallLeaves;
            c.level_chunk_hint[(size_t)l] = (done || all_sub) ? 0 : c.n / kChunk + cells + 1;
            done = done || all_sub;
        }
    }
    // a level's last Hoare row swaps run in the next level's first launch;
    // the last level's at the end
    PendingFinal fin;
    for (int32_t l = 1; l <= lmax; ++l) {
        const int32_t nCells = b.level_n[(size_t)l];
        if (nCells == 0) continue;
        const int tlevel = timing_mark(c);
        // the subtree kernel runs where a cell may be small enough for it:
        // cells split by count, so one with nLeaf leaves holds about
        // n * nLeaf / d keys; where even the smallest expected cell is several
        // times the subtree's limit, the launch (and its flags) are left out
        // -- only a speed choice, every cell is built either way
        const int64_t few = c.n * (int64_t)b.level_min_leaf[(size_t)l] / (c.d > 0 ? c.d : 1);
        const bool may_sub = few <= 4 * (int64_t)subtree_max_keys(c);
        const LevelEntry *entries = b.level_ids.p + b.level_off[(size_t)l];
        ORB_TRY(enqueue_level(c, entries, nCells, b.heap.p, b.exists.p, b.sub.p, b.chunk_hist.p,
                              b.cell_hist.p, trace, b.level_passes.p, c.status.p,
                              b.complete && may_sub, b.level_cover[(size_t)l] != 0, l,
                              b.chunk_hist_w.p, b.cell_hist_w.p, &fin, true));
        fin.entries = entries;
        fin.nCells = c.part == ORB_PART_HOARE ? nCells : 0;
        ORB_TRY(timing_pair(c, tlevel, kTimeLevel + l));
    }
    if (fin.nCells > 0) ORB_TRY(launch_hoare_final(c, fin.nCells));
    return particles_home(c);  // in-stream: the capture ends where it began
}

int build_batched(Ctx &c, int32_t levels_mode, orb_build_info *info)
{
    bool complete = true;
    ORB_TRY(ensure_state(c, levels_mode, &complete));
    BatchedState &b = *c.batched;
    timing_reset(c);
    c.table_ids.clear();

    if (!b.hist_clean) {
        ORB_TRY(launch_zero_u32(c.stream, b.cell_hist.p, (size_t)c.max_level_cells * kBins));
        b.hist_clean = true;
    }
    if (c.weighted && !b.hist_w_clean) {
        ORB_TRY(launch_zero_u64(c.stream, b.cell_hist_w.p, (size_t)c.max_level_cells * kBins));
        b.hist_w_clean = true;
    }
    c.trace_weighted = c.weighted;
    ORB_TRY(particles_home(c));
    const bool graph = c.use_graph && !c.timing;
    if (graph) {
        std::vector<uint64_t> sig = graph_signature(c, b, levels_mode);
        if (!b.exec || sig != b.gsig) {
            drop_graph(b);
            ORB_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_build(c, levels_mode);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c.stream, &g);
            if (rc < 0) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            ORB_HIP(e);
            b.graph = g;
            ORB_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
            b.gsig = std::move(sig);
        }
        ORB_HIP(hipGraphLaunch(b.exec, c.stream));
    } else {
        ORB_TRY(enqueue_build(c, levels_mode));
    }
    c.table_ids.clear();

    // status and pass counts come back now; the tree stays in HBM until a
    // getter asks for it (batched_pull_results)
    ORB_HIP(hipMemcpyAsync(c.hstatus.p, c.status.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c.stream));
    ORB_HIP(hipMemcpyAsync(b.hpasses.p, b.level_passes.p, 64 * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    c.results_on_device = true;
    ORB_TRY(timing_collect(c));
    const int32_t lmax = levels_mode == ORB_LEVELS_REFERENCE ? b.D - 1 : b.D;
    for (int32_t l = 1; l <= lmax; ++l) {
        info->nlevels = l;
        if (b.level_n[(size_t)l] == 0) continue;
        info->level_passes[l - 1] = b.hpasses.p[l];
        info->total_passes += b.hpasses.p[l];
    }
    if (status_error(c.hstatus.p, &info->error_cell)) {
        b.hist_clean = false;  // an aborted subtree may leave rows behind
        b.hist_w_clean = false;
        return ORB_ERR_UB;
    }
    return ORB_OK;
}

int batched_pull_results(Ctx &c)
{
    if (!c.batched || !c.results_on_device) return ORB_OK;
    BatchedState &b = *c.batched;
    c.heap.resize((size_t)c.cap);
    c.exists.resize((size_t)c.cap);
    ORB_HIP(hipMemcpyAsync(c.heap.data(), b.heap.p, sizeof(orb_cell) * (size_t)c.cap,
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipMemcpyAsync(c.exists.data(), b.exists.p, (size_t)c.cap, hipMemcpyDeviceToHost,
                           c.stream));
    if (c.trace && b.trace.p) {
        c.trace_counts.resize((size_t)c.cap * ORB_MAX_PASSES);
        ORB_HIP(hipMemcpyAsync(c.trace_counts.data(), b.trace.p,
                               sizeof(uint64_t) * c.trace_counts.size(), hipMemcpyDeviceToHost,
                               c.stream));
    }
    ORB_HIP(hipStreamSynchronize(c.stream));
    c.results_on_device = false;
    return ORB_OK;
}

void batched_drop_graph(Ctx &c)
{
    if (c.batched) drop_graph(*c.batched);
}

void batched_release(Ctx &c)
{
    if (!c.batched) return;
    BatchedState &b = *c.batched;
    drop_graph(b);
    b.level_ids.release();
    b.heap.release();
    b.exists.release();
    b.sub.release();
    b.chunk_hist.release();
    b.cell_hist.release();
    b.chunk_hist_w.release();
    b.cell_hist_w.release();
    b.trace.release();
    b.level_passes.release();
    b.hpasses.release();
    delete c.batched;
    c.batched = nullptr;
}

}  // namespace orb
