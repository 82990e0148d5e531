// orb_api.hip -- the C-ABI of liborb_hip.so: context lifecycle, the
// service-shaped entry points and the plain (host-driven) build.
//
// The plain build restates the level loop of src/orbit.cpp:74-275 over the
// service entry points: per level one chunk-table build (ServiceCopyCells),
// one ServiceCount, then up to 32 bisection passes, each one count launch
// whose per-cell counts come back to the host for the termination rule, then
// the split on the host and one partition launch sequence.
#include <atomic>
#include <chrono>
#include <new>

#include "orb_internal.hpp"

namespace orb {

static std::atomic<int> g_last_hip_error{0};
void set_last_hip_error(hipError_t e) { g_last_hip_error.store((int)e); }

bool level_cells(int32_t d, int32_t levels_mode, int32_t l, const std::vector<uint8_t> &exists,
                 const std::vector<orb_cell> &heap, std::vector<int32_t> &ids, int32_t *bad_id)
{
    ids.clear();
    int32_t D = n_levels(d);
    if (l < 1) return true;
    if (levels_mode == ORB_LEVELS_REFERENCE) {
        // orbit.cpp:102-110
        if (l >= D) return true;
        int64_t a = (int64_t(1) << (l - 1)) - 1;
        int64_t last = 2 * int64_t(d) - (int64_t(1) << D);
        int64_t p2 = int64_t(1) << l;
        int64_t b = (last < p2 ? last : p2) - 2;
        for (int64_t id = a; id <= b; ++id) {
            if (!exists[(size_t)id]) {
                if (bad_id) *bad_id = (int32_t)id;
                return false;
            }
            ids.push_back((int32_t)id);
        }
        return true;
    }
    if (l > D) return true;
    int64_t a = (int64_t(1) << (l - 1)) - 1;
    int64_t b = (int64_t(1) << l) - 2;
    for (int64_t id = a; id <= b; ++id)
        if (exists[(size_t)id] && heap[(size_t)id].nLeafCells >= 2) ids.push_back((int32_t)id);
    return true;
}

void timing_reset(Ctx &c)
{
    c.tm = orb_timing{};
    c.level_ms.clear();
    c.tnext = 0;
    c.tpairs.clear();
    if (c.timing && c.swap_pairs.alloc(1) == ORB_OK)
        (void)launch_zero_u64(c.stream, reinterpret_cast<uint64_t *>(c.swap_pairs.p), 1);
}

int timing_mark(Ctx &c)
{
    if (!c.timing) return -1;
    if (c.tnext == c.tpool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        c.tpool.push_back(e);
    }
    int idx = (int)c.tnext++;
    if (hipEventRecord(c.tpool[(size_t)idx], c.stream) != hipSuccess) return -1;
    return idx;
}

int timing_pair(Ctx &c, int begin_mark, int kind)
{
    if (begin_mark < 0) return ORB_OK;
    int end_mark = timing_mark(c);
    if (end_mark < 0) return ORB_OK;
    c.tpairs.push_back(Ctx::TPair{begin_mark, end_mark, kind});
    return ORB_OK;
}

int timing_collect(Ctx &c, bool reuse_events)
{
    for (const auto &p : c.tpairs) {
        float ms = 0.f;
        ORB_HIP(hipEventElapsedTime(&ms, c.tpool[(size_t)p.a], c.tpool[(size_t)p.b]));
        if (p.kind >= kTimeLevel) {
            size_t l = (size_t)(p.kind - kTimeLevel);
            if (c.level_ms.size() <= l) c.level_ms.resize(l + 1, 0.0);
            c.level_ms[l] += ms;
        } else if (p.kind == kTimeCount) {
            c.tm.count_ms += ms;
            c.tm.count_launches += 1;
        } else if (p.kind == kTimePartition) {
            c.tm.partition_ms += ms;
        } else if (p.kind == kTimeSwap) {
            c.tm.swap_ms += ms;
            c.tm.swap_launches += 1;
        } else {
            c.tm.other_ms += ms;
        }
    }
    c.tpairs.clear();
    if (reuse_events) c.tnext = 0;
    return ORB_OK;
}

static int reset_ranges(Ctx &c)
{
    ORB_TRY(launch_zero_u32(c.stream, c.ranges.p, 2 * (size_t)c.cap));
    c.hranges.p[0] = 0;
    c.hranges.p[1] = (uint32_t)c.n;
    ORB_HIP(hipMemcpyAsync(c.ranges.p, c.hranges.p, 2 * sizeof(uint32_t), hipMemcpyHostToDevice,
                           c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    c.table_ids.clear();
    return ORB_OK;
}

static int clear_status(Ctx &c)
{
    return launch_zero_u32(c.stream, reinterpret_cast<uint32_t *>(c.status.p), 2);
}

static int read_status(Ctx &c, int32_t *bad_id)
{
    ORB_HIP(hipMemcpyAsync(c.hstatus.p, c.status.p, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    if (status_error(c.hstatus.p, bad_id)) return ORB_ERR_UB;
    return ORB_OK;
}

// Uploads the cells and (re)builds the chunk table if the cell list changed.
static int prepare(Ctx &c, const orb_cell *cells, int32_t nCells)
{
    if (nCells < 0 || (nCells > 0 && !cells)) return ORB_ERR_ARG;
    for (int32_t i = 0; i < nCells; ++i)
        if (cells[i].id < 0 || cells[i].id >= c.cap) return ORB_ERR_ARG;
    ORB_TRY(upload_level(c, cells, nCells));
    bool same = (int32_t)c.table_ids.size() == nCells;
    for (int32_t i = 0; same && i < nCells; ++i) same = c.table_ids[i] == cells[i].id;
    if (!same) {
        ORB_TRY(build_chunk_table(c, nCells));
        // a service call's cells may overlap (a level's never do): when they
        // hold more chunks than the table has room for, it grows and is built
        // again (every chunk kernel reads up to first[nCells])
        ORB_HIP(hipMemcpyAsync(c.hcnt.p, c.cell_first.p + nCells, sizeof(uint32_t),
                               hipMemcpyDeviceToHost, c.stream));
        ORB_HIP(hipStreamSynchronize(c.stream));
        const int64_t total = c.hcnt.p[0];
        if (total > c.max_chunks) {
            ORB_TRY(ensure_level_capacity(c, nCells, total + 1));
            ORB_TRY(upload_level(c, cells, nCells));
            ORB_TRY(build_chunk_table(c, nCells));
        }
        c.table_ids.resize((size_t)nCells);
        for (int32_t i = 0; i < nCells; ++i) c.table_ids[i] = cells[i].id;
    }
    return ORB_OK;
}

static int timed_count(Ctx &c, int32_t nCells, bool skip_found, const float *axis)
{
    if (c.timing) ORB_HIP(hipEventRecord(c.ev[0], c.stream));
    ORB_TRY(launch_count(c, nCells, skip_found, axis, c.timing ? c.ev[1] : nullptr));
    return ORB_OK;
}

static int finish_timed_count(Ctx &c)
{
    if (!c.timing) return ORB_OK;
    float ms = 0.f;
    ORB_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
    c.tm.count_ms += ms;
    c.tm.count_launches += 1;
    return ORB_OK;
}

// one count-left launch; out[i] gets lt (skip_found) or le counts
static int count_service(Ctx &c, const orb_cell *cells, int32_t nCells, uint32_t *out,
                         bool lt_skip_found, const float *axis)
{
    if (!out && nCells > 0) return ORB_ERR_ARG;
    ORB_TRY(prepare(c, cells, nCells));
    if (nCells == 0) return 0;
    ORB_TRY(timed_count(c, nCells, lt_skip_found, axis));
    ORB_HIP(hipMemcpyAsync(c.hcnt.p, c.cnt.p, 2 * sizeof(uint32_t) * (size_t)nCells,
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    ORB_TRY(finish_timed_count(c));
    for (int32_t i = 0; i < nCells; ++i) {
        if (lt_skip_found) {
            if (!cells[i].foundCut) out[i] = c.hcnt.p[2 * i];
        } else {
            out[i] = c.hcnt.p[2 * i + 1];
        }
    }
    return (int)(nCells * sizeof(uint32_t));
}

static int partition_service(Ctx &c, const orb_cell *cells, int32_t nCells, bool hoare,
                             int32_t *bad_id, bool full_cover = false)
{
    ORB_TRY(prepare(c, cells, nCells));
    if (nCells == 0) return 0;
    for (int32_t i = 0; i < nCells; ++i)
        if (right_child_id(cells[i].id) >= c.cap) return ORB_ERR_ARG;
    ORB_TRY(clear_status(c));
    if (c.timing) ORB_HIP(hipEventRecord(c.ev[0], c.stream));
    if (hoare) ORB_TRY(launch_hoare(c, nCells));
    else ORB_TRY(launch_stable_le(c, nCells, nullptr, full_cover));
    if (c.timing) ORB_HIP(hipEventRecord(c.ev[1], c.stream));
    int rc = read_status(c, bad_id);
    if (c.timing) {
        float ms = 0.f;
        ORB_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        c.tm.partition_ms += ms;
        ORB_TRY(timing_collect(c, !c.building));  // the swap kernel's own pair
    }
    c.table_ids.clear();  // ranges of the children changed; next level rebuilds
    if (rc < 0) return rc;
    return 0;
}

static int count_sizes(Ctx &c, const orb_cell *cells, int32_t nCells, uint32_t *out)
{
    if (nCells > 0 && !out) return ORB_ERR_ARG;
    ORB_TRY(prepare(c, cells, nCells));
    if (nCells == 0) return 0;
    ORB_TRY(launch_cell_sizes(c, nCells, c.cnt.p));
    ORB_HIP(hipMemcpyAsync(c.hcnt.p, c.cnt.p, sizeof(uint32_t) * (size_t)nCells,
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(out, c.hcnt.p, sizeof(uint32_t) * (size_t)nCells);
    return (int)(nCells * sizeof(uint32_t));
}

// Weighted services (workload weights on): per-cell weight sums in 64 bits.
// total: the cell's weight (ServiceCount with weights); otherwise the weight
// left of the cut, < with found cells skipped (their out[i] untouched) or <=
// for every cell.
static int weight_service(Ctx &c, const orb_cell *cells, int32_t nCells, uint64_t *out,
                          bool total, bool lt_skip_found)
{
    if (!c.weighted) return ORB_ERR_STATE;
    if (!out && nCells > 0) return ORB_ERR_ARG;
    ORB_TRY(prepare(c, cells, nCells));
    if (nCells == 0) return 0;
    if (c.timing && !total) ORB_HIP(hipEventRecord(c.ev[0], c.stream));
    ORB_TRY(launch_count_w(c, nCells, lt_skip_found, total));
    if (c.timing && !total) ORB_HIP(hipEventRecord(c.ev[1], c.stream));
    ORB_HIP(hipMemcpyAsync(c.hcnt_w.p, c.cnt_w.p, 2 * sizeof(uint64_t) * (size_t)nCells,
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    if (!total) ORB_TRY(finish_timed_count(c));
    for (int32_t i = 0; i < nCells; ++i) {
        if (total) out[i] = c.hcnt_w.p[2 * i];
        else if (!lt_skip_found) out[i] = c.hcnt_w.p[2 * i + 1];
        else if (!cells[i].foundCut) out[i] = c.hcnt_w.p[2 * i];
    }
    return (int)(nCells * sizeof(uint64_t));
}

// ---------------------------------------------------------------------------
// plain build
// ---------------------------------------------------------------------------
static int start_build(Ctx &c)
{
    ORB_TRY(particles_home(c));
    c.results_on_device = false;
    c.heap.assign((size_t)c.cap, orb_cell{});
    std::memset(c.heap.data(), 0, sizeof(orb_cell) * c.heap.size());
    c.exists.assign((size_t)c.cap, 0);
    c.heap[0] = cell_root(c.d);
    c.exists[0] = 1;
    if (c.trace) c.trace_counts.assign((size_t)c.cap * ORB_MAX_PASSES, ~0ull);
    timing_reset(c);
    return reset_ranges(c);
}

static int build_plain(Ctx &c, int32_t levels_mode, orb_build_info *info)
{
    struct Building {
        Ctx &c;
        explicit Building(Ctx &cc) : c(cc) { c.building = true; }
        ~Building() { c.building = false; }
    } building(c);
    ORB_TRY(start_build(c));
    const int32_t D = n_levels(c.d);
    std::vector<int32_t> ids;
    std::vector<orb_cell> cells;
    std::vector<uint32_t> counts, left;
    std::vector<uint64_t> wcounts, wleft;  // weighted: weight sums
    c.trace_weighted = c.weighted;
    for (int32_t l = 1; l <= D; ++l) {
        int32_t bad = -1;
        if (!level_cells(c.d, levels_mode, l, c.exists, c.heap, ids, &bad)) {
            info->error_cell = bad;
            return ORB_ERR_UB;
        }
        if (levels_mode == ORB_LEVELS_REFERENCE && l >= D) break;
        info->nlevels = l;
        const int32_t nCells = (int32_t)ids.size();
        if (nCells == 0) continue;
        const int tlevel = timing_mark(c);
        cells.resize((size_t)nCells);
        for (int32_t i = 0; i < nCells; ++i) cells[i] = c.heap[(size_t)ids[i]];
        counts.assign((size_t)nCells, 0u);
        left.assign((size_t)nCells, 0u);

        ORB_TRY(prepare(c, cells.data(), nCells));          // ServiceCopyCells
        ORB_TRY(count_sizes(c, cells.data(), nCells, counts.data()));  // ServiceCount
        if (c.weighted) {  // the cells' weights
            wcounts.assign((size_t)nCells, 0ull);
            wleft.assign((size_t)nCells, 0ull);
            ORB_TRY(weight_service(c, cells.data(), nCells, wcounts.data(), true, false));
        }

        // bisection controller (orbit.cpp:146-232)
        bool foundAll = false;
        int32_t j = 0;
        while (!foundAll && j < ORB_MAX_PASSES) {
            j++;
            foundAll = true;
            int64_t bytes = 0;
            const int64_t per_key = c.weighted ? 8 : 4;  // key (+ weight)
            for (int32_t i = 0; i < nCells; ++i)
                if (c.cmp != ORB_CMP_LT || !cells[i].foundCut) bytes += per_key * (int64_t)counts[i];
            const bool lt = c.cmp == ORB_CMP_LT;
            if (c.weighted)
                ORB_TRY(weight_service(c, cells.data(), nCells, wleft.data(), false, lt));
            else
                ORB_TRY(count_service(c, cells.data(), nCells, left.data(), lt, nullptr));
            c.tm.count_bytes += bytes;
            for (int32_t i = 0; i < nCells; ++i) {
                orb_cell &cl = cells[i];
                if (cl.foundCut) continue;
                if (c.trace)
                    c.trace_counts[(size_t)cl.id * ORB_MAX_PASSES + (size_t)(j - 1)] =
                        c.weighted ? wleft[i] : left[i];
                // orbit.cpp:204-205; weighted: the same rule on weight sums
                const int64_t difference =
                    c.weighted ? difference_from_target_w(
                                     wleft[i], difference_target_w(wcounts[i], cl.nLeafCells))
                               : difference_rule(left[i], counts[i], cl.nLeafCells);
                if (difference < 3 && difference > -3) {
                    cl.foundCut = true;
                } else if (difference > 0) {
                    cl.cutMarginRight = cell_get_cut(cl);
                    foundAll = false;
                } else {
                    cl.cutMarginLeft = cell_get_cut(cl);
                    foundAll = false;
                }
            }
        }
        info->level_passes[l - 1] = j;
        info->total_passes += j;

        // split (orbit.cpp:235-250); the heap keeps the final margins
        for (int32_t i = 0; i < nCells; ++i) {
            c.heap[(size_t)ids[i]] = cells[i];
            orb_cell lc, rc;
            std::memset(&lc, 0, sizeof lc);
            std::memset(&rc, 0, sizeof rc);
            if (!cell_split(cells[i], lc, rc)) {
                info->error_cell = cells[i].id;
                return ORB_ERR_UB;
            }
            c.heap[(size_t)lc.id] = lc;
            c.heap[(size_t)rc.id] = rc;
            c.exists[(size_t)lc.id] = 1;
            c.exists[(size_t)rc.id] = 1;
        }

        // a level whose cells hold every particle lets the stable partition
        // flip buffers instead of copying back
        int64_t covered = 0;
        for (int32_t i = 0; i < nCells; ++i) covered += counts[i];
        int32_t bad_id = -1;
        int rc = partition_service(c, cells.data(), nCells, c.part == ORB_PART_HOARE, &bad_id,
                                   covered == c.n);
        if (rc < 0) {
            info->error_cell = bad_id;
            return rc;
        }
        ORB_TRY(timing_pair(c, tlevel, kTimeLevel + l));
    }
    ORB_TRY(particles_home(c));
    ORB_HIP(hipStreamSynchronize(c.stream));
    return timing_collect(c);
}

}  // namespace orb

using namespace orb;

static inline Ctx *C(orb_ctx *ctx) { return ctx ? &ctx->c : nullptr; }

extern "C" {

const char *orb_error_string(int code)
{
    switch (code) {
    case ORB_OK: return "ok";
    case ORB_ERR_ARG: return "invalid argument";
    case ORB_ERR_HIP: return "HIP runtime error";
    case ORB_ERR_NOMEM: return "out of memory";
    case ORB_ERR_UB: return "input on which the reference is undefined";
    case ORB_ERR_STATE: return "call out of sequence";
    case ORB_ERR_NODEV: return "no GPU";
    case ORB_ERR_BUSY: return "not admitted";
    default: return code >= 0 ? "ok" : "unknown error";
    }
}

int orb_last_hip_error(void) { return g_last_hip_error.load(); }

int orb_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int64_t orb_heap_capacity(int32_t d) { return heap_capacity(d); }

int orb_init(orb_ctx **out, int64_t nParticles, int32_t d, bool generate, orb_meta_params params)
{
    if (!out) return ORB_ERR_ARG;
    *out = nullptr;
    if (nParticles < 0 || nParticles > (int64_t(1) << 31) || d < 1 || d > (1 << 24))
        return ORB_ERR_ARG;
    if (orb_device_count() < 1) return ORB_ERR_NODEV;
    orb_ctx *o = new (std::nothrow) orb_ctx;
    if (!o) return ORB_ERR_NOMEM;
    Ctx &c = o->c;
    auto fail = [&](int rc) {
        orb_finalize(o);
        return rc;
    };
    if (hipGetDevice(&c.device) != hipSuccess) return fail(ORB_ERR_HIP);
    hipError_t e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_last_hip_error(e);
        return fail(ORB_ERR_HIP);
    }
    if (hipEventCreate(&c.ev[0]) != hipSuccess || hipEventCreate(&c.ev[1]) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_upload, hipEventDisableTiming) != hipSuccess)
        return fail(ORB_ERR_HIP);
    c.n = nParticles;
    c.stride = ((nParticles + 63) / 64) * 64;
    if (c.stride == 0) c.stride = 64;
    c.d = d;
    c.params = params;
    c.cmp = params.GPU_COUNT ? ORB_CMP_LE : ORB_CMP_LT;
    c.part = params.GPU_PARTITION ? ORB_PART_STABLE_LE : ORB_PART_HOARE;
    c.cap = heap_capacity(d);
    int rc;
    // 64 floats of slack: 16-byte loads of a range ending at n may touch 3
    // floats past the z array
    if ((rc = c.pos.alloc(3 * (size_t)c.stride + 64)) < 0) return fail(rc);
    if ((rc = c.ranges.alloc(2 * (size_t)c.cap)) < 0) return fail(rc);
    if ((rc = c.packed.alloc((size_t)c.stride)) < 0) return fail(rc);
    if ((rc = c.status.alloc(2)) < 0) return fail(rc);
    if ((rc = c.hstatus.alloc(2)) < 0) return fail(rc);
    if ((rc = c.hranges.alloc(2 * (size_t)c.cap)) < 0) return fail(rc);
    if ((rc = ensure_level_capacity(c, (c.cap + 1) / 2)) < 0) return fail(rc);
    if (generate && nParticles > 0) {
        PinBuf<float> h;
        if ((rc = h.alloc(3 * (size_t)c.stride)) < 0) return fail(rc);
        // the reference generator (init.cu:11-25, 47-53): seed 0
        if ((rc = orb_generate(0, nParticles, h.p, h.p + c.stride, h.p + 2 * c.stride)) < 0)
            return fail(rc);
        e = hipMemcpyAsync(c.pos.p, h.p, 3 * sizeof(float) * (size_t)c.stride,
                           hipMemcpyHostToDevice, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        h.release();
        if (e != hipSuccess) {
            set_last_hip_error(e);
            return fail(ORB_ERR_HIP);
        }
    }
    if ((rc = reset_ranges(c)) < 0) return fail(rc);
    if ((rc = clear_status(c)) < 0) return fail(rc);
    *out = o;
    return ORB_OK;
}

int orb_finalize(orb_ctx *ctx)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    batched_release(*c);
    c->pos.release();
    c->axisT.release();
    c->tmp.release();
    c->saved.release();
    c->pid.release();
    c->pid_tmp.release();
    c->pid_saved.release();
    c->wt.release();
    c->wt_tmp.release();
    c->wt_saved.release();
    c->chunk_w.release();
    c->cnt_w.release();
    c->hcnt_w.release();
    c->ranges.release();
    c->swap_pairs.release();
    c->dcells.release();
    c->recs.release();
    c->cell_nchunks.release();
    c->cell_first.release();
    c->chunk_cell.release();
    c->chunk_desc.release();
    c->total_chunks.release();
    c->cnt.release();
    c->stamps.release();
    c->chunk_ltle.release();
    c->chunk_cnt.release();
    c->chunk_pre.release();
    c->scan_partials.release();
    c->packed.release();
    c->masks.release();
    c->cell_cand.release();
    c->splits.release();
    c->chunk_comp.release();
    c->sel.release();
    c->sel_lists.release();
    c->swap_plan.release();
    c->status.release();
    c->ids_dev.release();
    c->hcells.release();
    c->hcnt.release();
    c->hstatus.release();
    c->hranges.release();
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->ev_upload) (void)hipEventDestroy(c->ev_upload);
    for (auto &ev : c->tpool) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete ctx;
    return ORB_OK;
}

int orb_set_semantics(orb_ctx *ctx, int32_t cmp, int32_t partition)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    if (cmp != -1 && cmp != ORB_CMP_LT && cmp != ORB_CMP_LE) return ORB_ERR_ARG;
    if (partition != -1 && partition != ORB_PART_HOARE && partition != ORB_PART_STABLE_LE)
        return ORB_ERR_ARG;
    if (cmp != -1) c->cmp = cmp;
    if (partition != -1) c->part = partition;
    return ORB_OK;
}

int orb_set_particles(orb_ctx *ctx, const float *x, const float *y, const float *z, int64_t n)
{
    Ctx *c = C(ctx);
    if (!c || n != c->n || (n > 0 && (!x || !y || !z))) return ORB_ERR_ARG;
    c->in_tmp = false;
    c->weighted = false;  // new particles: their weights come with orb_set_weights
    if (c->track_ids) ORB_TRY(launch_iota(c->stream, c->pid.p, (size_t)c->stride));
    if (n > 0) {
        ORB_HIP(hipMemcpyAsync(c->x(), x, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        ORB_HIP(hipMemcpyAsync(c->y(), y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        ORB_HIP(hipMemcpyAsync(c->z(), z, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        // a NaN coordinate is refused (ORB_ERR_ARG): checked on the device,
        // through the status pair, which every build clears before use
        ORB_TRY(clear_status(*c));
        ORB_TRY(launch_find_nan(c->stream, c->x(), c->y(), c->z(), (size_t)n, c->status.p));
        ORB_HIP(hipMemcpyAsync(c->hstatus.p, c->status.p, 2 * sizeof(int32_t),
                               hipMemcpyDeviceToHost, c->stream));
        ORB_HIP(hipStreamSynchronize(c->stream));
        const bool has_nan = status_error(c->hstatus.p, nullptr);
        ORB_TRY(clear_status(*c));
        if (has_nan) {
            ORB_TRY(reset_ranges(*c));
            return ORB_ERR_ARG;
        }
    }
    return reset_ranges(*c);
}

int orb_get_particles(orb_ctx *ctx, float *x, float *y, float *z, int64_t n)
{
    Ctx *c = C(ctx);
    if (!c || n != c->n || (n > 0 && (!x || !y || !z))) return ORB_ERR_ARG;
    if (n > 0) {
        ORB_HIP(hipMemcpyAsync(x, c->x(), sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        ORB_HIP(hipMemcpyAsync(y, c->y(), sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        ORB_HIP(hipMemcpyAsync(z, c->z(), sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    }
    ORB_HIP(hipStreamSynchronize(c->stream));
    return ORB_OK;
}

int orb_reset_ranges(orb_ctx *ctx)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return reset_ranges(*c);
}

int orb_save_particles(orb_ctx *ctx)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    ORB_TRY(particles_home(*c));
    ORB_TRY(c->saved.alloc(3 * (size_t)c->stride));
    ORB_HIP(hipMemcpyAsync(c->saved.p, c->pos.p, 3 * sizeof(float) * (size_t)c->stride,
                           hipMemcpyDeviceToDevice, c->stream));
    if (c->track_ids) {
        ORB_TRY(c->pid_saved.alloc((size_t)c->stride));
        ORB_HIP(hipMemcpyAsync(c->pid_saved.p, c->pid.p, sizeof(uint32_t) * (size_t)c->stride,
                               hipMemcpyDeviceToDevice, c->stream));
    } else {
        c->pid_saved.release();
    }
    if (c->weighted) {
        ORB_TRY(c->wt_saved.alloc((size_t)c->stride));
        ORB_HIP(hipMemcpyAsync(c->wt_saved.p, c->wt.p, sizeof(uint32_t) * (size_t)c->stride,
                               hipMemcpyDeviceToDevice, c->stream));
    } else {
        c->wt_saved.release();
    }
    ORB_HIP(hipStreamSynchronize(c->stream));
    return ORB_OK;
}

int orb_restore_particles(orb_ctx *ctx)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    if (!c->saved.p) return ORB_ERR_STATE;
    c->in_tmp = false;
    ORB_HIP(hipMemcpyAsync(c->pos.p, c->saved.p, 3 * sizeof(float) * (size_t)c->stride,
                           hipMemcpyDeviceToDevice, c->stream));
    if (c->track_ids) {
        // ids saved with the particles, else the snapshot is the input order
        if (c->pid_saved.p)
            ORB_HIP(hipMemcpyAsync(c->pid.p, c->pid_saved.p, sizeof(uint32_t) * (size_t)c->stride,
                                   hipMemcpyDeviceToDevice, c->stream));
        else
            ORB_TRY(launch_iota(c->stream, c->pid.p, (size_t)c->stride));
    }
    // the weights saved with the particles come back with them; weights set
    // after the snapshot have no saved order to come back in
    if (c->weighted) {
        if (!c->wt_saved.p) return ORB_ERR_STATE;
        ORB_HIP(hipMemcpyAsync(c->wt.p, c->wt_saved.p, sizeof(uint32_t) * (size_t)c->stride,
                               hipMemcpyDeviceToDevice, c->stream));
    }
    return reset_ranges(*c);
}

int orb_set_weights(orb_ctx *ctx, const uint32_t *w, int64_t n)
{
    Ctx *c = C(ctx);
    if (!c || n != c->n) return ORB_ERR_ARG;
    if (!w) {
        c->weighted = false;
        return ORB_OK;
    }
    ORB_TRY(particles_home(*c));
    ORB_TRY(c->wt.alloc((size_t)c->stride));
    if (n > 0)
        ORB_HIP(hipMemcpyAsync(c->wt.p, w, sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice,
                               c->stream));
    c->weighted = true;
    if (c->part == ORB_PART_STABLE_LE) ORB_TRY(ensure_stable_buffers(*c));
    c->wt_saved.release();  // a snapshot taken before holds other weights
    ORB_HIP(hipStreamSynchronize(c->stream));
    return ORB_OK;
}

int orb_get_weights(orb_ctx *ctx, uint32_t *w, int64_t n)
{
    Ctx *c = C(ctx);
    if (!c || n != c->n || (n > 0 && !w)) return ORB_ERR_ARG;
    if (!c->weighted) return ORB_ERR_STATE;
    if (n > 0)
        ORB_HIP(hipMemcpyAsync(w, c->weights(), sizeof(uint32_t) * (size_t)n,
                               hipMemcpyDeviceToHost, c->stream));
    ORB_HIP(hipStreamSynchronize(c->stream));
    return ORB_OK;
}

int orb_count_weight(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint64_t *out)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return weight_service(*c, cells, nCells, out, true, false);
}

int orb_count_left_weight(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, int32_t cmp,
                          uint64_t *out)
{
    Ctx *c = C(ctx);
    if (!c || (cmp != ORB_CMP_LT && cmp != ORB_CMP_LE)) return ORB_ERR_ARG;
    return weight_service(*c, cells, nCells, out, false, cmp == ORB_CMP_LT);
}

int orb_enable_ids(orb_ctx *ctx, bool on)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    if (on == c->track_ids) return ORB_OK;
    ORB_TRY(particles_home(*c));
    if (on) {
        ORB_TRY(c->pid.alloc((size_t)c->stride));
        ORB_TRY(launch_iota(c->stream, c->pid.p, (size_t)c->stride));
        if (c->part == ORB_PART_STABLE_LE) ORB_TRY(c->pid_tmp.alloc((size_t)c->stride));
    } else {
        c->pid.release();
        c->pid_tmp.release();
        c->pid_saved.release();
    }
    c->track_ids = on;
    ORB_HIP(hipStreamSynchronize(c->stream));
    return ORB_OK;
}

int orb_get_ids(orb_ctx *ctx, uint32_t *ids, int64_t n)
{
    Ctx *c = C(ctx);
    if (!c || n != c->n || (n > 0 && !ids)) return ORB_ERR_ARG;
    if (!c->track_ids) return ORB_ERR_STATE;
    if (n > 0)
        ORB_HIP(hipMemcpyAsync(ids, c->ids(), sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost,
                               c->stream));
    ORB_HIP(hipStreamSynchronize(c->stream));
    return ORB_OK;
}

int orb_copy_particles(orb_ctx *ctx)
{
    if (!C(ctx)) return ORB_ERR_ARG;
    return (int)sizeof(int);
}

int orb_copy_cells(orb_ctx *ctx, const orb_cell *cells, int32_t nCells)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    ORB_TRY(prepare(*c, cells, nCells));
    return 0;
}

int orb_count(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return count_sizes(*c, cells, nCells, out);
}

int orb_make_axis(orb_ctx *ctx, const orb_cell *cells, int32_t nCells)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    ORB_TRY(prepare(*c, cells, nCells));
    if (nCells > 0) ORB_TRY(launch_make_axis(*c, nCells));
    ORB_HIP(hipStreamSynchronize(c->stream));
    return (int)sizeof(int);
}

int orb_count_left(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return count_service(*c, cells, nCells, out, true, nullptr);
}

int orb_count_left_gpu_axis(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    if (!c->axisT.p) return ORB_ERR_STATE;  // orb_make_axis fills particlesT first
    return count_service(*c, cells, nCells, out, false, c->axisT.p);
}

int orb_count_left_gpu(orb_ctx *ctx, const orb_cell *cells, int32_t nCells, uint32_t *out)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return count_service(*c, cells, nCells, out, false, nullptr);
}

int orb_partition(orb_ctx *ctx, const orb_cell *cells, int32_t nCells)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return partition_service(*c, cells, nCells, true, nullptr);
}

int orb_partition_gpu(orb_ctx *ctx, const orb_cell *cells, int32_t nCells)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    return partition_service(*c, cells, nCells, false, nullptr);
}

int orb_run_service(orb_ctx *ctx, int32_t id, const void *vin, int32_t nIn, void *vout,
                    int32_t nOut)
{
    Ctx *c = C(ctx);
    if (!c || nIn < 0 || nOut < 0) return ORB_ERR_ARG;
    const orb_cell *cells = static_cast<const orb_cell *>(vin);
    int32_t nCells = nIn / (int32_t)sizeof(orb_cell);
    uint32_t *out = static_cast<uint32_t *>(vout);
    auto need_out = [&]() { return nOut >= nCells * (int32_t)sizeof(uint32_t) && (nCells == 0 || out); };
    switch (id) {
    case ORB_PST_COPYPARTICLES: return orb_copy_particles(ctx);
    case ORB_PST_COPYCELLS: return orb_copy_cells(ctx, cells, nCells);
    case ORB_PST_COUNT: return need_out() ? orb_count(ctx, cells, nCells, out) : ORB_ERR_ARG;
    case ORB_PST_MAKEAXIS: return orb_make_axis(ctx, cells, nCells);
    case ORB_PST_COUNTLEFT: return need_out() ? orb_count_left(ctx, cells, nCells, out) : ORB_ERR_ARG;
    case ORB_PST_COUNTLEFTAXISGPU:
        return need_out() ? orb_count_left_gpu_axis(ctx, cells, nCells, out) : ORB_ERR_ARG;
    case ORB_PST_COUNTLEFTGPU:
        return need_out() ? orb_count_left_gpu(ctx, cells, nCells, out) : ORB_ERR_ARG;
    case ORB_PST_PARTITION: return orb_partition(ctx, cells, nCells);
    case ORB_PST_PARTITIONGPU: return orb_partition_gpu(ctx, cells, nCells);
    case ORB_PST_SETADD:
    case ORB_PST_INIT:
    case ORB_PST_INITGPU:
    case ORB_PST_FINALIZE:
        return 0;  // lifecycle goes through orb_init / orb_finalize
    default: return ORB_ERR_ARG;
    }
}

int orb_build(orb_ctx *ctx, int32_t form, int32_t levels_mode, orb_build_info *info)
{
    Ctx *c = C(ctx);
    if (!c || (form != ORB_FORM_PLAIN && form != ORB_FORM_BATCHED)) return ORB_ERR_ARG;
    if (levels_mode != ORB_LEVELS_REFERENCE && levels_mode != ORB_LEVELS_FULL) return ORB_ERR_ARG;
    orb_build_info local;
    if (!info) info = &local;
    std::memset(info, 0, sizeof *info);
    info->error_cell = -1;
    auto t0 = std::chrono::steady_clock::now();
    int rc = form == ORB_FORM_PLAIN ? build_plain(*c, levels_mode, info)
                                    : build_batched(*c, levels_mode, info);
    if (rc >= 0) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            set_last_hip_error(e);
            rc = ORB_ERR_HIP;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    info->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return rc;
}

int orb_get_heap(orb_ctx *ctx, orb_cell *heap, int64_t cap)
{
    Ctx *c = C(ctx);
    if (!c || !heap || cap < c->cap) return ORB_ERR_ARG;
    if (c->results_on_device) ORB_TRY(batched_pull_results(*c));
    std::memset(heap, 0, sizeof(orb_cell) * (size_t)cap);
    if (!c->heap.empty()) std::memcpy(heap, c->heap.data(), sizeof(orb_cell) * c->heap.size());
    return ORB_OK;
}

int orb_get_exists(orb_ctx *ctx, uint8_t *exists, int64_t cap)
{
    Ctx *c = C(ctx);
    if (!c || !exists || cap < c->cap) return ORB_ERR_ARG;
    if (c->results_on_device) ORB_TRY(batched_pull_results(*c));
    std::memset(exists, 0, (size_t)cap);
    if (!c->exists.empty()) std::memcpy(exists, c->exists.data(), c->exists.size());
    return ORB_OK;
}

int orb_get_ranges(orb_ctx *ctx, uint32_t *ranges, int64_t cap)
{
    Ctx *c = C(ctx);
    if (!c || !ranges || cap < c->cap) return ORB_ERR_ARG;
    std::memset(ranges, 0, 2 * sizeof(uint32_t) * (size_t)cap);
    ORB_HIP(hipMemcpyAsync(c->hranges.p, c->ranges.p, 2 * sizeof(uint32_t) * (size_t)c->cap,
                           hipMemcpyDeviceToHost, c->stream));
    ORB_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(ranges, c->hranges.p, 2 * sizeof(uint32_t) * (size_t)c->cap);
    return ORB_OK;
}

int orb_enable_trace(orb_ctx *ctx, bool on)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    c->trace = on;
    return ORB_OK;
}

int orb_get_pass_counts(orb_ctx *ctx, uint32_t *counts, int64_t cap)
{
    Ctx *c = C(ctx);
    if (!c || !counts || cap < c->cap) return ORB_ERR_ARG;
    if (c->results_on_device) ORB_TRY(batched_pull_results(*c));
    if (c->trace_counts.size() != (size_t)c->cap * ORB_MAX_PASSES) return ORB_ERR_STATE;
    if (c->trace_weighted) return ORB_ERR_STATE;  // weight sums: orb_get_pass_weights
    for (size_t k = 0; k < c->trace_counts.size(); ++k) {
        const uint64_t v = c->trace_counts[k];
        counts[k] = v == ~0ull ? 0xffffffffu : (uint32_t)v;
    }
    return ORB_OK;
}

int orb_get_pass_weights(orb_ctx *ctx, uint64_t *sums, int64_t cap)
{
    Ctx *c = C(ctx);
    if (!c || !sums || cap < c->cap) return ORB_ERR_ARG;
    if (c->results_on_device) ORB_TRY(batched_pull_results(*c));
    if (c->trace_counts.size() != (size_t)c->cap * ORB_MAX_PASSES) return ORB_ERR_STATE;
    std::memcpy(sums, c->trace_counts.data(), sizeof(uint64_t) * c->trace_counts.size());
    return ORB_OK;
}

int orb_get_level_timing(orb_ctx *ctx, double *ms, int32_t max_levels)
{
    Ctx *c = C(ctx);
    if (!c || !ms || max_levels < 0) return ORB_ERR_ARG;
    // level l (1-based) is ms[l-1]
    int32_t n = 0;
    for (int32_t l = 1; l <= max_levels; ++l) {
        ms[l - 1] = (size_t)l < c->level_ms.size() ? c->level_ms[(size_t)l] : 0.0;
        if (ms[l - 1] > 0.0) n = l;
    }
    return n;
}

// Stream kernels for the bandwidth reference, shaped like the count pass:
// each block streams one tile of kStreamTile float4 (16 KB at four per
// thread), all of a thread's accesses in flight at once, blocks per tile.
constexpr int kStreamPer = 4;
constexpr size_t kStreamTile = (size_t)kBlock * kStreamPer;
__global__ __launch_bounds__(kBlock) void k_stream_read(const float4 *a, size_t n4, float *out)
{
    const size_t base = (size_t)blockIdx.x * kStreamTile + threadIdx.x;
    float4 v[kStreamPer];
#pragma unroll
    for (int k = 0; k < kStreamPer; ++k) {
        const size_t i = base + (size_t)k * kBlock;
        v[k] = i < n4 ? a[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < kStreamPer; ++k) s += (v[k].x + v[k].y) + (v[k].z + v[k].w);
    if (s == 1.0e30f) out[blockIdx.x & 4095u] = s;  // keeps the loads; practically never taken
}

__global__ __launch_bounds__(kBlock) void k_stream_copy(const float4 *a, float4 *b, size_t n4)
{
    const size_t base = (size_t)blockIdx.x * kStreamTile + threadIdx.x;
    float4 v[kStreamPer];
#pragma unroll
    for (int k = 0; k < kStreamPer; ++k) {
        const size_t i = base + (size_t)k * kBlock;
        if (i < n4) v[k] = a[i];
    }
#pragma unroll
    for (int k = 0; k < kStreamPer; ++k) {
        const size_t i = base + (size_t)k * kBlock;
        if (i < n4) b[i] = v[k];
    }
}

int orb_measure_stream(orb_ctx *ctx, int32_t reps, double *read_gbs, double *copy_gbs)
{
    Ctx *c = C(ctx);
    if (!c || reps < 1 || !read_gbs || !copy_gbs) return ORB_ERR_ARG;
    const size_t n4 = 3 * (size_t)c->stride / 4;  // the particle arrays (stride is a multiple of 64)
    DevBuf<float> dst, sink;
    ORB_TRY(dst.alloc(4 * n4));
    ORB_TRY(sink.alloc(4096));
    const unsigned grid = (unsigned)((n4 + kStreamTile - 1) / kStreamTile);
    const float4 *a = reinterpret_cast<const float4 *>(c->pos.p);
    float4 *b = reinterpret_cast<float4 *>(dst.p);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ORB_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return ORB_ERR_HIP;
    }
    int rc = ORB_OK;
    float ms_read = 0.0f, ms_copy = 0.0f;
    auto run = [&](bool copy, float *ms) -> int {
        for (int32_t r = -1; r < reps; ++r) {  // one untimed launch first
            if (r == 0) ORB_HIP(hipEventRecord(e0, c->stream));
            if (copy)
                hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, c->stream, a, b, n4);
            else
                hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(kBlock), 0, c->stream, a, n4,
                                   sink.p);
            ORB_HIP(hipGetLastError());
        }
        ORB_HIP(hipEventRecord(e1, c->stream));
        ORB_HIP(hipEventSynchronize(e1));
        ORB_HIP(hipEventElapsedTime(ms, e0, e1));
        return ORB_OK;
    };
    rc = run(false, &ms_read);
    if (rc == ORB_OK) rc = run(true, &ms_copy);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != ORB_OK) return rc;
    const double bytes = 16.0 * (double)n4 * reps;
    *read_gbs = ms_read > 0.0f ? bytes / (ms_read * 1e6) : 0.0;
    *copy_gbs = ms_copy > 0.0f ? 2.0 * bytes / (ms_copy * 1e6) : 0.0;
    return ORB_OK;
}

int orb_enable_graph(orb_ctx *ctx, bool on)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    c->use_graph = on;
    return ORB_OK;
}

int orb_enable_timing(orb_ctx *ctx, bool on)
{
    Ctx *c = C(ctx);
    if (!c) return ORB_ERR_ARG;
    c->timing = on;
    return ORB_OK;
}

int orb_get_timing(orb_ctx *ctx, orb_timing *t)
{
    Ctx *c = C(ctx);
    if (!c || !t) return ORB_ERR_ARG;
    *t = c->tm;
    if (c->timing && c->swap_pairs.p) {
        unsigned long long k = 0;
        ORB_HIP(hipMemcpyAsync(&k, c->swap_pairs.p, sizeof k, hipMemcpyDeviceToHost, c->stream));
        ORB_HIP(hipStreamSynchronize(c->stream));
        const int64_t row = 12 + (c->track_ids ? 4 : 0) + (c->weighted ? 4 : 0);
        t->swap_pairs = (int64_t)k;
        t->swap_bytes = (int64_t)k * 2 * 2 * row;  // two rows, read and written
    }
    return ORB_OK;
}

}  // extern "C"
