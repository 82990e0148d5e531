// orb_internal.hpp -- internals shared by the HIP translation units of
// liborb_hip.so: the context, error plumbing, and the Cell rules compiled for
// host and device.
//
// All translation units are compiled with -ffp-contract=off: the cut midpoint
// (src/cell.h:74-76) and the termination rule (src/orbit.cpp:204-208) are
// float32 expressions whose results change if a multiply and a subtract are
// fused.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "orb.h"

namespace orb {

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;                 // threads per block (4 wave64s)
constexpr int kPerThread = 16;              // keys per thread per chunk
constexpr int kChunk = kBlock * kPerThread; // 4096 keys = 16 KiB per chunk
constexpr int kHistK = 8;                   // bisection steps per histogram round
constexpr int kBins = 1 << kHistK;          // 256 bins (255 candidate cuts)
constexpr int kMaskWords = kChunk / 64;     // 64-bit stopper-mask words per chunk and side
static_assert(kBins == kBlock, "one thread per histogram bin");
constexpr uint32_t kChunkIdle = 0xffffffffu;  // chunk record axis: no search this level
// batched form, per-chunk 256-bin rows: the first round's bins, the first
// round's ties (both read again by the partition), the later rounds' bins
constexpr int kHistRegions = 3;

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
void set_last_hip_error(hipError_t e);

#define ORB_HIP(call)                                                   \
    do {                                                                \
        hipError_t orb_e_ = (call);                                     \
        if (orb_e_ != hipSuccess) {                                     \
            ::orb::set_last_hip_error(orb_e_);                          \
            return ORB_ERR_HIP;                                         \
        }                                                               \
    } while (0)

#define ORB_TRY(expr)                                                   \
    do {                                                                \
        int orb_rc_ = (expr);                                           \
        if (orb_rc_ < 0) return orb_rc_;                                \
    } while (0)

// ---------------------------------------------------------------------------
// undefined-input report
// ---------------------------------------------------------------------------
// A build's status pair (two int32 words, 8-byte aligned) holds one 64-bit
// word: 0 while no cell was undefined, else the complement of the key of the
// first undefined cell in the reference's order.  orbit.cpp:102-275 walks the
// levels in order; within a level it splits every cell (orbit.cpp:235-250)
// before it partitions any (orbit.cpp:252-274); within each of those loops
// it takes the cells in heap-id order (the level slices are id ranges).  So
// the key is (level, split before partition, heap id), and every site that
// finds an undefined cell takes the maximum of the complemented words: the
// result does not depend on which check lands first.
enum { kErrSplit = 0, kErrPartition = 1 };
__host__ __device__ inline uint64_t error_word(int32_t level, int32_t kind, int32_t id)
{
    const uint64_t key = ((uint64_t)(uint32_t)level << 40) | ((uint64_t)(uint32_t)kind << 32) |
                         (uint64_t)(uint32_t)id;
    return ~key;
}
// The word of an input the library refuses (a non-finite coordinate in a
// device request): above every cell's word, so it is what the build reports.
constexpr uint64_t kErrInputWord = ~0ull;

// Host side: true if the pair records an undefined cell; *id gets its heap id.
inline bool status_error(const int32_t *pair, int32_t *id)
{
    const uint64_t w = (uint64_t)(uint32_t)pair[0] | ((uint64_t)(uint32_t)pair[1] << 32);
    if (w == 0) return false;
    if (id) *id = w == kErrInputWord ? -1 : (int32_t)(uint32_t)(~w);
    return true;
}

// ---------------------------------------------------------------------------
// Cell rules (src/cell.h), host and device
// ---------------------------------------------------------------------------
__host__ __device__ inline float cell_cut(float left, float right)
{
    // cell.h:74-76: (cutMarginRight + cutMarginLeft) / 2.0 -> float sum,
    // exact double halving, narrowed back to float.
    float s = right + left;
    return (float)((double)s / 2.0);
}

__host__ __device__ inline float cell_get_cut(const orb_cell &c)
{
    return cell_cut(c.cutMarginLeft, c.cutMarginRight);
}

__host__ __device__ inline int32_t left_child_id(int32_t id) { return (id + 1) * 2 - 1; }
__host__ __device__ inline int32_t right_child_id(int32_t id) { return (id + 1) * 2; }

__host__ __device__ inline int32_t difference_rule(uint32_t count_left, uint32_t count,
                                                   int32_t nLeafCells)
{
    // orbit.cpp:204-205.  ratio: double quotient narrowed to float; the count
    // converts to float; product and difference are separate float roundings;
    // the conversion to int truncates toward zero.
    float ratio = (float)(ceil(nLeafCells / 2.0) / (double)nLeafCells);
#if defined(__HIP_DEVICE_COMPILE__)
    float prod = __fmul_rn((float)count, ratio);
    float diff = __fsub_rn((float)count_left, prod);
#else
    float prod = (float)count * ratio;
    float diff = (float)count_left - prod;
#endif
    return (int32_t)diff;
}

// The same rule split in two: the product (float)count * ratio depends only on
// the cell, so a loop over passes computes it once (identical roundings).
__host__ __device__ inline float difference_target(uint32_t count, int32_t nLeafCells)
{
    float ratio = (float)(ceil(nLeafCells / 2.0) / (double)nLeafCells);
#if defined(__HIP_DEVICE_COMPILE__)
    return __fmul_rn((float)count, ratio);
#else
    return (float)count * ratio;
#endif
}

__host__ __device__ inline int32_t difference_from_target(uint32_t count_left, float target)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (int32_t)__fsub_rn((float)count_left, target);
#else
    return (int32_t)((float)count_left - target);
#endif
}

// uint64 -> float32, round to nearest even, by integer operations alone (the
// same bits on the host and the device): the top 24 significant bits, then
// the round bit and the sticky bits below it.  For values below 2^32 it is
// the hardware conversion of the uint32 value (a correctly rounded result is
// unique).
__host__ __device__ inline float u64_to_f32(uint64_t v)
{
    if (v == 0) return 0.0f;
    const int msb = 63 - __builtin_clzll(v);  // value in [2^msb, 2^(msb+1))
    uint64_t m;               // 24-bit significand
    if (msb <= 23) {
        m = v << (23 - msb);
    } else {
        const int sh = msb - 23;
        m = v >> sh;
        const uint64_t rest = v & ((1ull << sh) - 1ull), half = 1ull << (sh - 1);
        if (rest > half || (rest == half && (m & 1ull))) ++m;
    }
    int e = msb;
    if (m == (1ull << 24)) {  // rounded up to the next power of two
        m >>= 1;
        ++e;
    }
    const uint32_t bits = ((uint32_t)(e + 127) << 23) | ((uint32_t)m & 0x7fffffu);
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

// The termination rule of orbit.cpp:204-208 on weight sums (workload
// weights, documentation/notes.tex:259-271, 326-336): the weight left of the
// cut against the cell's weight times the left share, in the same float
// roundings; the truncation keeps 64 bits.  With every weight 1 the sums
// are the counts and the result is difference_rule's.
__host__ __device__ inline float difference_target_w(uint64_t weight, int32_t nLeafCells)
{
    float ratio = (float)(ceil(nLeafCells / 2.0) / (double)nLeafCells);
#if defined(__HIP_DEVICE_COMPILE__)
    return __fmul_rn(u64_to_f32(weight), ratio);
#else
    return u64_to_f32(weight) * ratio;
#endif
}

__host__ __device__ inline int64_t difference_from_target_w(uint64_t weight_left, float target)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (int64_t)__fsub_rn(u64_to_f32(weight_left), target);
#else
    return (int64_t)(u64_to_f32(weight_left) - target);
#endif
}

__host__ __device__ inline void cell_make(orb_cell &c, int32_t id, int32_t nLeaf,
                                          const float lower[3], const float upper[3])
{
    // cell.h:19-45
    c.id = id;
    c.nLeafCells = nLeaf;
    c.prevCutAxis = -1;
    c.cutAxis = -1;
    c.foundCut = false;
    c.cutMarginLeft = 0.0f;
    c.cutMarginRight = 0.0f;
    for (int d = 0; d < 3; ++d) {
        c.lower[d] = lower[d];
        c.upper[d] = upper[d];
    }
}

__host__ __device__ inline int32_t cell_cut_axis(const orb_cell &c)
{
    // cell.h:102-121: longest upper-lower, strict '>', first max, else -1
    int maxD = -1;
    float maxSize = 0.0f;
    for (int d = 0; d < 3; d++) {
        float size = c.upper[d] - c.lower[d];
        if (size > maxSize) {
            maxSize = size;
            maxD = d;
        }
    }
    return maxD;
}

// a[k] for k in [0, 2] by selects (no indexed private memory on the GPU)
__host__ __device__ inline float get3(const float a[3], int k)
{
    return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]);
}

__host__ __device__ inline void set3(float a[3], int k, float v)
{
    a[0] = k == 0 ? v : a[0];
    a[1] = k == 1 ? v : a[1];
    a[2] = k == 2 ? v : a[2];
}

// cut() + setCutAxis() + setCutMargin() on both children, in the order of
// orbit.cpp:235-243.  Returns false where the reference would index with
// cutAxis == -1 (a child without positive extent).
__host__ __device__ inline bool cell_split(const orb_cell &p, orb_cell &l, orb_cell &r)
{
    if (p.cutAxis < 0 || p.cutAxis > 2) return false;
    int32_t nLeft = (int32_t)ceil(p.nLeafCells / 2.0);
    int32_t nRight = p.nLeafCells - nLeft;
    float cut = cell_get_cut(p);
    cell_make(l, left_child_id(p.id), nLeft, p.lower, p.upper);
    set3(l.upper, p.cutAxis, cut);
    l.prevCutAxis = p.cutAxis;
    cell_make(r, right_child_id(p.id), nRight, p.lower, p.upper);
    set3(r.lower, p.cutAxis, cut);
    r.prevCutAxis = p.cutAxis;
    r.cutAxis = cell_cut_axis(r);
    l.cutAxis = cell_cut_axis(l);
    if (r.cutAxis < 0 || l.cutAxis < 0) return false;
    r.cutMarginLeft = get3(r.lower, r.cutAxis);
    r.cutMarginRight = get3(r.upper, r.cutAxis);
    l.cutMarginLeft = get3(l.lower, l.cutAxis);
    l.cutMarginRight = get3(l.upper, l.cutAxis);
    return true;
}

inline orb_cell cell_root(int32_t d)
{
    // orbit.cpp:45-46, 74-76
    const float lo[3] = {-0.5f, -0.5f, -0.5f};
    const float up[3] = {0.5f, 0.5f, 0.5f};
    orb_cell r;
    std::memset(&r, 0, sizeof r);
    cell_make(r, 0, d, lo, up);
    r.cutAxis = 0;
    r.cutMarginLeft = r.lower[0];
    r.cutMarginRight = r.upper[0];
    return r;
}

// Cells per wave of the per-cell kernels (for_cell_groups): one while a
// level has at most 16384 cells, then enough that the level takes about
// 16384 waves, at most 64 per wave; and the blocks (four waves) it needs.
inline int32_t cell_group(int32_t nCells)
{
    int32_t g = 1;
    while (g < 64 && (int64_t)g * 16384 < (int64_t)nCells) g *= 2;
    return g;
}
inline unsigned cell_group_blocks(int32_t nCells, int32_t group)
{
    const int64_t waves = ((int64_t)nCells + group - 1) / group;
    return (unsigned)((waves + 3) / 4);
}

inline int32_t n_levels(int32_t d)
{
    // cell.h:65-67 ceil(log2 d), in integers
    int32_t D = 0;
    while ((int64_t(1) << D) < int64_t(d)) ++D;
    return D;
}

inline int64_t heap_capacity(int32_t d)
{
    if (d < 1) return 0;
    int64_t cap = (int64_t(1) << (n_levels(d) + 1)) - 1;
    int64_t ref = 2 * int64_t(d) - 1;
    return cap > ref ? cap : ref;
}

// ids of the cells processed on level l (1-based), in heap order.
// REFERENCE: orbit.cpp:102-111 slices; FULL: every existing cell of depth l-1
// with nLeafCells >= 2.  Returns false for a slice naming an id never created.
bool level_cells(int32_t d, int32_t levels_mode, int32_t l, const std::vector<uint8_t> &exists,
                 const std::vector<orb_cell> &heap, std::vector<int32_t> &ids, int32_t *bad_id);

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count)
    {
        if (count <= n && p) return ORB_OK;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            set_last_hip_error(e);
            p = nullptr;
            n = 0;
            return ORB_ERR_NOMEM;
        }
        n = count;
        return ORB_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

template <class T>
struct PinBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count)
    {
        if (count <= n && p) return ORB_OK;
        release();
        if (count == 0) count = 1;
        hipError_t e = hipHostMalloc((void **)&p, count * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            set_last_hip_error(e);
            p = nullptr;
            n = 0;
            return ORB_ERR_NOMEM;
        }
        n = count;
        return ORB_OK;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

// Per-cell record the kernels read (derived from orb_cell + the range map).
struct CellRec {
    uint32_t begin;     // absolute index of the first particle
    uint32_t end;       // one past the last
    float cut;          // getCut() of the cell
    int32_t axis;       // cutAxis
    int32_t found;      // foundCut; batched form: 1 + the candidate index of the cut in
                        // the histogram round that found it (0: not found)
    int32_t nLeaf;      // nLeafCells
    float margin_l;     // cutMarginLeft
    float margin_r;     // cutMarginRight
    int32_t id;         // heap id
    int32_t flags;      // batched form: 0 active, 1 inside a finished subtree, 2 small (subtree kernel)
    int32_t passes;     // batched form: bisection passes so far
    // where the cell's build keeps its state (several builds can share a
    // launch): heap/range/flag index = hb + id, the build's heap capacity,
    // its pass counter for this level, its status pair, its last level
    int32_t hb;
    int32_t capr;
    int32_t lp;
    int32_t st;
    int32_t lmax;
};

// One cell of a level step: the build it belongs to and where that build's
// state lives (see CellRec).
struct LevelEntry {
    int32_t id;       // heap id within its build
    int32_t hb;       // heap base of the build
    int32_t capr;     // heap capacity of the build
    int32_t lp;       // pass-counter index: build slot * 64 + level
    int32_t st;       // status pair index: 2 * build slot
    int32_t lmax;     // last level of the build
    int32_t sub_ok;   // the build's schedule lets a subtree finish in LDS
    int32_t pad;
};

// Cut and axis a cell's per-chunk counts were taken at (axis -1: none).
struct CountStamp {
    float cut;
    int32_t axis;
};

// Result of the Hoare split of one cell.
struct SplitRec {
    uint32_t TL;        // keys >= cut
    uint32_t TR;        // keys <= cut
    uint32_t K;         // swapped pairs
    uint32_t i;         // split index (cell-local)
};

// ---------------------------------------------------------------------------
// the context (the PST leaf's LocalData, src/services/pst.h:6-41)
// ---------------------------------------------------------------------------
struct BatchedState;

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0;
    int64_t stride = 0;        // floats between the x, y and z arrays (n rounded up to 64)
    int32_t d = 1;
    orb_meta_params params{};
    int32_t cmp = ORB_CMP_LT;
    int32_t part = ORB_PART_HOARE;
    int64_t cap = 0;           // heap capacity (ids)

    // particles: one allocation of 3 x stride floats (+64 floats of slack), SoA
    DevBuf<float> pos;         // x = pos.p, y = pos.p + stride, z = pos.p + 2 stride
    DevBuf<float> axisT;       // particlesT (orb_make_axis)
    DevBuf<float> tmp;         // same shape: the stable partition's other buffer
    DevBuf<float> saved;       // orb_save_particles snapshot
    bool in_tmp = false;       // the stable partition left the particles in tmp
    // optional particle ids: the input index of the particle at each position
    bool track_ids = false;
    DevBuf<uint32_t> pid, pid_tmp, pid_saved;
    // optional workload weights w(p_i) (documentation/notes.tex:259-271):
    // uint32 per particle, moved with the particles; the bisection counts
    // become weight sums (uint64)
    bool weighted = false;
    DevBuf<uint32_t> wt, wt_tmp, wt_saved;
    DevBuf<uint64_t> chunk_w;        // [max_chunks][2] plain form: weight < cut, <= cut
    DevBuf<uint64_t> cnt_w;          // [cells][2] per-cell sums
    PinBuf<uint64_t> hcnt_w;         // [cells][2]
    DevBuf<uint32_t> ranges;   // [cap][2] cellToRangeMap

    // level tables (sized for the widest level)
    int64_t max_level_cells = 0;
    int64_t max_chunks = 0;
    DevBuf<orb_cell> dcells;
    DevBuf<CellRec> recs;
    DevBuf<uint32_t> cell_nchunks;   // [cells+1]
    DevBuf<uint32_t> cell_first;     // [cells+1] exclusive scan
    DevBuf<uint32_t> chunk_cell;     // [max_chunks]
    DevBuf<uint4> chunk_desc;        // [max_chunks] first key, end, cell, axis (kChunkIdle:
                                     // the cell is not searched this level)
    DevBuf<uint32_t> total_chunks;   // [1]
    DevBuf<uint32_t> cnt;            // [cells][2] lt, le
    DevBuf<CountStamp> stamps;       // [cells] cut/axis of the chunk counts
    DevBuf<uint32_t> chunk_ltle;     // [max_chunks+1] packed lt | le<<16 per chunk
    DevBuf<uint64_t> chunk_cnt;      // [max_chunks+1] packed (nL | nR<<32)
    DevBuf<uint64_t> chunk_pre;      // [max_chunks+1]
    DevBuf<uint64_t> scan_partials;  // scan scratch
    DevBuf<float> packed;            // [n] batched form: window keys packed by the second
                                     // histogram round, per chunk share
    DevBuf<uint64_t> masks;          // [max_chunks][2][64] Hoare stopper masks per chunk
    DevBuf<float> cell_cand;         // [max_level_cells][256] batched form: candidate cuts of
                                     // each searching cell for the second histogram round
                                     // (written by the first round's replay)
    DevBuf<uint32_t> chunk_comp;     // [max_chunks+1][8] packed keys per wave, keys below
    DevBuf<float> sel;               // [n] batched form: each searching cell's keys inside its
                                     // margins after the second round, from the cell's begin
    DevBuf<uint64_t> sel_lists;      // [max_chunks+1] batched form: count of keys gathered into
                                     // a list of sel | keys below << 32, at the list's first chunk
    DevBuf<SplitRec> splits;         // [cells]
    DevBuf<uint4> swap_plan;         // [max_chunks][2] Hoare swap unit of each chunk
    DevBuf<int32_t> status;          // [2]: error flag, failing heap id
    DevBuf<uint32_t> ids_dev;        // [cells]

    // host staging
    PinBuf<orb_cell> hcells;
    PinBuf<uint32_t> hcnt;
    PinBuf<int32_t> hstatus;
    PinBuf<uint32_t> hranges;
    std::vector<int32_t> table_ids;  // ids of the current chunk table
    int32_t table_cells = 0;

    // build results
    std::vector<orb_cell> heap;
    std::vector<uint8_t> exists;
    std::vector<uint64_t> trace_counts;  // [cap][32] left count (weighted: weight) per pass
    bool trace = false;
    bool trace_weighted = false;     // the last traced build summed weights
    bool timing = false;
    bool building = false;     // inside a plain build (its level marks are pending)
    bool use_graph = true;    // batched form: capture the build once, replay it
    bool results_on_device = false;  // batched form: heap/exists/trace not yet mirrored
    std::vector<double> level_ms;    // per-level time of the last timed build
    orb_timing tm{};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t ev_upload = nullptr;  // last H2D out of hcells (pinned staging reuse)
    // timing: events recorded in-stream, read after the build's final sync
    std::vector<hipEvent_t> tpool;
    size_t tnext = 0;
    struct TPair { int a, b, kind; };
    std::vector<TPair> tpairs;
    DevBuf<unsigned long long> swap_pairs;  // timing: K summed over the timed build's swaps
    BatchedState *batched = nullptr;
    // batched-form launch shapes (chosen by orb_autotune; 0 / -1 = default)
    int32_t cur_level = -1;              // level being enqueued, -1 outside a build
    int32_t sub_max_keys = -1;           // cells of at most this many keys go to the subtree
    std::vector<int32_t> tune_chunk_blocks;  // [level] grid cap of the chunk kernels
    std::vector<int32_t> tune_sub_blocks;    // [level] grid of the subtree kernel
    std::vector<double> tune_level_ms;       // [level] device ms of the last autotune's choice
    std::vector<int64_t> level_chunk_hint;   // [level] chunks a batched level expects (< 0: unknown)

    // the buffer that holds the particles now, and the other one
    float *cur() const { return in_tmp ? tmp.p : pos.p; }
    float *other() const { return in_tmp ? pos.p : tmp.p; }
    float *x() const { return cur(); }
    float *y() const { return cur() + stride; }
    float *z() const { return cur() + 2 * stride; }
    uint32_t *ids() const { return track_ids ? (in_tmp ? pid_tmp.p : pid.p) : nullptr; }
    uint32_t *other_ids() const { return track_ids ? (in_tmp ? pid.p : pid_tmp.p) : nullptr; }
    uint32_t *weights() const { return weighted ? (in_tmp ? wt_tmp.p : wt.p) : nullptr; }
    uint32_t *other_weights() const { return weighted ? (in_tmp ? wt.p : wt_tmp.p) : nullptr; }
};

// timing helpers (orb_api.hip): marks are -1 when timing is off
enum { kTimeCount = 0, kTimePartition = 1, kTimeOther = 2, kTimeSwap = 3, kTimeLevel = 100 /* + level */ };
void timing_reset(Ctx &c);
int timing_mark(Ctx &c);
int timing_pair(Ctx &c, int begin_mark, int kind);
int timing_collect(Ctx &c, bool reuse_events = true);

// kernels / launchers (orb_kernels.hip)
int launch_zero_u32(hipStream_t s, uint32_t *p, size_t n);
int launch_zero_u64(hipStream_t s, uint64_t *p, size_t n);
int scan_exclusive_u32(Ctx &c, const uint32_t *in, uint32_t *out, size_t m);
int scan_exclusive_u64(Ctx &c, const uint64_t *in, uint64_t *out, size_t m,
                       const uint32_t *lim = nullptr);
// the same with caller-provided partials (one per 2048 entries), no allocation
int scan_exclusive_u32_into(hipStream_t s, const uint32_t *in, uint32_t *out, size_t m,
                            uint32_t *partials);
unsigned chunk_grid(const Ctx &c);
int upload_level(Ctx &c, const orb_cell *cells, int32_t nCells);
// scanned: cell_first already holds the exclusive scan of the chunk counts
int build_chunk_table(Ctx &c, int32_t nCells, bool with_stamps = true, bool scanned = false);
int launch_count(Ctx &c, int32_t nCells, bool lt_skip_found, const float *axis_override,
                 hipEvent_t after_count = nullptr);
// Weighted count pass into c.cnt_w ([cells][2]: weight < cut, <= cut; with
// total_only both are the cell's weight); skip_found leaves found cells out.
int launch_count_w(Ctx &c, int32_t nCells, bool skip_found, bool total_only);
// The count pass over caller-provided tables (no context, no allocation):
// the chunk table from the exclusive chunk offsets first[] (first[nCells] at
// most max_chunks), the chunk counts, the per-cell lt/le sums into cnt.
int launch_count_explicit(hipStream_t s, const CellRec *recs, int32_t nCells, const uint32_t *first,
                          uint32_t *chunk_cell, uint4 *chunk_desc, uint32_t max_chunks,
                          const float *x, const float *y, const float *z, bool skip_found,
                          uint32_t *chunk_ltle, uint32_t *cnt, CountStamp *stamps,
                          unsigned chunk_blocks);
// Per-chunk histograms of the round that found each cell's cut (batched
// form): the stopper counts come from them instead of a key read.
struct HistCounts {
    const uint16_t *hist = nullptr;   // [chunks][256] bin counts
    const uint16_t *tie = nullptr;    // [chunks][256] keys equal to candidate j
    int le = 0;                       // bins were built with <= (else <)
    const float *pack = nullptr;      // keys packed by the second round, per chunk share
    const uint32_t *comp = nullptr;   // [chunks][8] packed counts per share, keys below
};
int launch_hoare(Ctx &c, int32_t nCells, int32_t *status = nullptr,
                 const HistCounts *hc = nullptr, int rev = 0, bool defer_final = false);
int launch_hoare_final(Ctx &c, int32_t nCells);  // rows i and n-1 of the level's cells
int launch_stopper_counts(Ctx &c, int32_t nCells, bool le_only, const HistCounts *hc = nullptr);
int invalidate_stamps(Ctx &c, int32_t nCells);
int launch_stable_le(Ctx &c, int32_t nCells, int32_t *status = nullptr, bool full_cover = false,
                     const HistCounts *hc = nullptr);
int ensure_stable_buffers(Ctx &c);
int launch_copy_u32(hipStream_t s, const uint32_t *src, uint32_t *dst, size_t n);
int launch_iota(hipStream_t s, uint32_t *p, size_t n);
// raises the status pair to kErrInputWord if x, y or z hold a NaN
int launch_find_nan(hipStream_t s, const float *x, const float *y, const float *z, size_t n,
                    int32_t *status_pair);
int particles_home(Ctx &c);
int schedule_levels(int32_t d, int32_t levels_mode, int64_t cap,
                    std::vector<std::vector<int32_t>> &levels, bool *complete,
                    std::vector<int32_t> *min_leaf = nullptr);
int launch_make_axis(Ctx &c, int32_t nCells);
int launch_cell_sizes(Ctx &c, int32_t nCells, uint32_t *dst);
// The previous level's cells whose last Hoare row swaps are still due
// (entries, count): a batched build runs them in the next level's first
// launch.
struct PendingFinal {
    const LevelEntry *entries = nullptr;
    int32_t nCells = 0;
};
// one level step of the batched form over `entries` (device array)
// (chunk_hist_w, cell_hist_w: the weighted rounds' rows, when c.weighted;
// fin_in: the previous level's due row swaps, run in this level's first
// launch; defer_final: leave this level's to the next)
int enqueue_level(Ctx &c, const LevelEntry *entries, int32_t nCells, orb_cell *heap,
                  uint8_t *exists, uint8_t *sub, uint16_t *chunk_hist, uint32_t *cell_hist,
                  uint64_t *trace, int32_t *level_passes, int32_t *status, bool any_subtree,
                  bool full_cover, int32_t level, uint64_t *chunk_hist_w = nullptr,
                  uint64_t *cell_hist_w = nullptr, const PendingFinal *fin_in = nullptr,
                  bool defer_final = false);
int preallocate_scratch(Ctx &c);
// level tables for nCells cells and (at least) `chunks` chunks
int ensure_level_capacity(Ctx &c, int64_t nCells, int64_t chunks = 0);

// batched build (orb_batched.hip)
int build_batched(Ctx &c, int32_t levels_mode, orb_build_info *info);
void batched_drop_graph(Ctx &c);
int32_t subtree_max_keys(const Ctx &c);
int batched_pull_results(Ctx &c);
void batched_release(Ctx &c);

}  // namespace orb

struct orb_ctx {
    orb::Ctx c;
};
