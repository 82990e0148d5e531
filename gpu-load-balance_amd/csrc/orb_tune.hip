// orb_tune.hip -- autotuner for the batched form's launch shapes.
//
// Every shape it chooses between computes the same result, so it can time
// them on the caller's own particles: trial builds from one snapshot, with
// device events around each level.  The subtree threshold moves work
// between levels (a cell taken by the subtree kernel leaves nothing for the
// deeper levels), so it is chosen on the whole build; the per-level grid
// shapes only change how a level's own work is spread, so each level keeps
// its fastest one.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "orb_internal.hpp"

namespace orb {

namespace {

struct Snapshot {
    DevBuf<float> pos;
    DevBuf<uint32_t> pid;
    DevBuf<uint32_t> wt;
};

int take_snapshot(Ctx &c, Snapshot &s)
{
    ORB_TRY(particles_home(c));
    ORB_TRY(s.pos.alloc(3 * (size_t)c.stride + 64));
    ORB_HIP(hipMemcpyAsync(s.pos.p, c.pos.p, sizeof(float) * (3 * (size_t)c.stride + 64),
                           hipMemcpyDeviceToDevice, c.stream));
    if (c.track_ids) {
        ORB_TRY(s.pid.alloc((size_t)c.stride));
        ORB_HIP(hipMemcpyAsync(s.pid.p, c.pid.p, sizeof(uint32_t) * (size_t)c.stride,
                               hipMemcpyDeviceToDevice, c.stream));
    }
    if (c.weighted) {
        ORB_TRY(s.wt.alloc((size_t)c.stride));
        ORB_HIP(hipMemcpyAsync(s.wt.p, c.wt.p, sizeof(uint32_t) * (size_t)c.stride,
                               hipMemcpyDeviceToDevice, c.stream));
    }
    ORB_HIP(hipStreamSynchronize(c.stream));
    return ORB_OK;
}

int put_back(Ctx &c, const Snapshot &s)
{
    c.in_tmp = false;
    ORB_HIP(hipMemcpyAsync(c.pos.p, s.pos.p, sizeof(float) * (3 * (size_t)c.stride + 64),
                           hipMemcpyDeviceToDevice, c.stream));
    if (c.track_ids && s.pid.p)
        ORB_HIP(hipMemcpyAsync(c.pid.p, s.pid.p, sizeof(uint32_t) * (size_t)c.stride,
                               hipMemcpyDeviceToDevice, c.stream));
    if (c.weighted && s.wt.p)
        ORB_HIP(hipMemcpyAsync(c.wt.p, s.wt.p, sizeof(uint32_t) * (size_t)c.stride,
                               hipMemcpyDeviceToDevice, c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    return ORB_OK;
}

// One timed build from the snapshot: per-level device ms into lv (index
// l - 1), returns the sum.
int timed_build(Ctx &c, const Snapshot &s, int32_t levels_mode, int32_t repeats,
                std::vector<double> &lv, double *total)
{
    lv.assign(64, std::numeric_limits<double>::infinity());
    double best = std::numeric_limits<double>::infinity();
    for (int r = 0; r < repeats; ++r) {
        ORB_TRY(put_back(c, s));
        orb_build_info info;
        std::memset(&info, 0, sizeof info);
        info.error_cell = -1;
        ORB_TRY(build_batched(c, levels_mode, &info));
        double sum = 0.0;
        for (size_t l = 1; l < c.level_ms.size() && l <= 64; ++l) {
            lv[l - 1] = std::min(lv[l - 1], c.level_ms[l]);
            sum += c.level_ms[l];
        }
        best = std::min(best, sum);
    }
    *total = best;
    return ORB_OK;
}

}  // namespace

}  // namespace orb

using namespace orb;

static void tuning_from_ctx(const Ctx &c, orb_tuning *t)
{
    std::memset(t, 0, sizeof *t);
    t->subtree_max_keys = c.sub_max_keys;
    int32_t nl = (int32_t)std::max(c.tune_chunk_blocks.size(), c.tune_sub_blocks.size());
    if (nl > 0) nl -= 1;  // index 0 unused
    t->nlevels = std::min(nl, 64);
    for (int32_t l = 1; l <= t->nlevels; ++l) {
        if ((size_t)l < c.tune_chunk_blocks.size()) t->chunk_blocks[l - 1] = c.tune_chunk_blocks[(size_t)l];
        if ((size_t)l < c.tune_sub_blocks.size()) t->subtree_blocks[l - 1] = c.tune_sub_blocks[(size_t)l];
    }
}

extern "C" {

int orb_set_tuning(orb_ctx *ctx, const orb_tuning *t)
{
    if (!ctx) return ORB_ERR_ARG;
    Ctx &c = ctx->c;
    c.tune_level_ms.clear();
    if (!t) {
        c.sub_max_keys = -1;
        c.tune_chunk_blocks.clear();
        c.tune_sub_blocks.clear();
    } else {
        if (t->nlevels < 0 || t->nlevels > 64) return ORB_ERR_ARG;
        for (int32_t l = 0; l < t->nlevels; ++l)
            if (t->chunk_blocks[l] < 0 || t->subtree_blocks[l] < 0) return ORB_ERR_ARG;
        c.sub_max_keys = t->subtree_max_keys;
        c.tune_chunk_blocks.assign((size_t)t->nlevels + 1, 0);
        c.tune_sub_blocks.assign((size_t)t->nlevels + 1, 0);
        for (int32_t l = 1; l <= t->nlevels; ++l) {
            c.tune_chunk_blocks[(size_t)l] = t->chunk_blocks[l - 1];
            c.tune_sub_blocks[(size_t)l] = t->subtree_blocks[l - 1];
        }
    }
    batched_drop_graph(c);
    return ORB_OK;
}

int orb_get_tuning(orb_ctx *ctx, orb_tuning *t)
{
    if (!ctx || !t) return ORB_ERR_ARG;
    tuning_from_ctx(ctx->c, t);
    return ORB_OK;
}

// The context's choices as keyed rows: one per level of its (n, d).
int orb_get_tune_rows(orb_ctx *ctx, orb_tune_row *rows, int32_t cap)
{
    if (!ctx || cap < 0 || (cap > 0 && !rows)) return ORB_ERR_ARG;
    const Ctx &c = ctx->c;
    const int32_t D = n_levels(c.d);
    int32_t n = 0;
    for (int32_t l = 1; l <= D && l <= 64; ++l) {
        int64_t lo = 0, hi = 0;
        const int cells = orb_level_shape(c.n, c.d, l, &lo, &hi);
        if (cells < 0) return cells;
        if (cells == 0) continue;
        if (n == cap) return ORB_ERR_ARG;
        orb_tune_row &r = rows[n++];
        std::memset(&r, 0, sizeof r);
        r.depth = l;
        r.log2_min_keys = lo > 0 ? 63 - __builtin_clzll((unsigned long long)lo) : -1;
        r.log2_max_keys = hi > 0 ? 63 - __builtin_clzll((unsigned long long)hi) : -1;
        r.chunk_blocks = (size_t)l < c.tune_chunk_blocks.size() ? c.tune_chunk_blocks[(size_t)l] : 0;
        r.subtree_blocks = (size_t)l < c.tune_sub_blocks.size() ? c.tune_sub_blocks[(size_t)l] : 0;
        r.subtree_max_keys = c.sub_max_keys;
        r.ms = (size_t)l < c.tune_level_ms.size() ? c.tune_level_ms[(size_t)l] : 0.0;
    }
    return n;
}

// Each level of the context takes the row with its cell-size key, the same
// depth first; a level with no such row keeps the defaults.  The subtree
// threshold comes from the first matched row.
int orb_set_tune_rows(orb_ctx *ctx, const orb_tune_row *rows, int32_t nrows)
{
    if (!ctx || nrows < 0 || (nrows > 0 && !rows)) return ORB_ERR_ARG;
    for (int32_t i = 0; i < nrows; ++i)
        if (rows[i].chunk_blocks < 0 || rows[i].subtree_blocks < 0) return ORB_ERR_ARG;
    Ctx &c = ctx->c;
    const int32_t D = n_levels(c.d);
    std::vector<int32_t> cb((size_t)D + 1, 0), sb((size_t)D + 1, 0);
    int32_t sub_max = -1, matched = 0;
    for (int32_t l = 1; l <= D && l <= 64; ++l) {
        int64_t lo = 0, hi = 0;
        const int cells = orb_level_shape(c.n, c.d, l, &lo, &hi);
        if (cells < 0) return cells;
        if (cells == 0) continue;
        const int32_t klo = lo > 0 ? 63 - __builtin_clzll((unsigned long long)lo) : -1;
        const int32_t khi = hi > 0 ? 63 - __builtin_clzll((unsigned long long)hi) : -1;
        const orb_tune_row *best = nullptr;
        for (int32_t i = 0; i < nrows; ++i) {
            const orb_tune_row &r = rows[i];
            if (r.log2_min_keys != klo || r.log2_max_keys != khi) continue;
            if (!best || (r.depth == l && best->depth != l)) best = &r;
        }
        if (!best) continue;
        cb[(size_t)l] = best->chunk_blocks;
        sb[(size_t)l] = best->subtree_blocks;
        if (matched == 0) sub_max = best->subtree_max_keys;
        matched++;
    }
    c.sub_max_keys = sub_max < 0 ? -1 : sub_max;
    c.tune_chunk_blocks = cb;
    c.tune_sub_blocks = sb;
    c.tune_level_ms.clear();
    batched_drop_graph(c);
    return matched;
}

int orb_autotune(orb_ctx *ctx, int32_t levels_mode, int32_t repeats, orb_tuning *out)
{
    if (!ctx || (levels_mode != ORB_LEVELS_REFERENCE && levels_mode != ORB_LEVELS_FULL))
        return ORB_ERR_ARG;
    Ctx &c = ctx->c;
    if (repeats < 1) repeats = 1;
    if (repeats > 8) repeats = 8;
    const bool timing_was = c.timing;
    const int32_t D = n_levels(c.d);
    Snapshot snap;
    int rc = take_snapshot(c, snap);
    if (rc < 0) return rc;
    c.timing = true;  // eager builds with per-level events

    auto finish = [&](int code) {
        c.timing = timing_was;
        int r2 = put_back(c, snap);
        snap.pos.release();
        snap.pid.release();
        batched_drop_graph(c);
        // as after orb_restore_particles: the input order, range(0) = [0, n)
        int r3 = orb_reset_ranges(ctx);
        return code < 0 ? code : (r2 < 0 ? r2 : r3);
    };

    std::vector<double> lv;
    double t_default = 0.0;
    // defaults first
    c.sub_max_keys = -1;
    c.tune_chunk_blocks.assign((size_t)D + 1, 0);
    c.tune_sub_blocks.assign((size_t)D + 1, 0);
    if ((rc = timed_build(c, snap, levels_mode, repeats, lv, &t_default)) < 0) return finish(rc);
    const std::vector<double> lv_default = lv;

    // 1. subtree threshold, whole build
    const int32_t thresholds[] = {-1, 4400, 2200, 0};
    int32_t best_thr = -1;
    double best_t = t_default;
    // a choice other than the default must win by kMargin: a level whose
    // shapes make no difference keeps the defaults instead of timing noise
    constexpr double kMargin = 0.98;
    for (int32_t thr : thresholds) {
        if (thr == -1) continue;  // measured above
        c.sub_max_keys = thr;
        double t = 0.0;
        if ((rc = timed_build(c, snap, levels_mode, repeats, lv, &t)) < 0) return finish(rc);
        if (t < best_t && t < kMargin * t_default) {
            best_t = t;
            best_thr = thr;
        }
    }
    c.sub_max_keys = best_thr;

    // 2. chunk-kernel grid cap per level
    // per level: the default's time, then the best alternative that beats
    // it by kMargin
    auto pick_per_level = [&](std::vector<int32_t> &shape, const int32_t *opts, int nopt,
                              std::vector<int32_t> &pick) -> int {
        std::vector<double> base(64, std::numeric_limits<double>::infinity()), best(base);
        pick.assign((size_t)D + 1, 0);
        for (int k = 0; k < nopt; ++k) {
            std::fill(shape.begin(), shape.end(), opts[k]);
            double t = 0.0;
            int r = timed_build(c, snap, levels_mode, repeats, lv, &t);
            if (r < 0) return r;
            for (int32_t l = 1; l <= D && l <= 64; ++l) {
                const double v = lv[(size_t)l - 1];
                if (opts[k] == 0) {
                    base[(size_t)l - 1] = v;
                    if (v < best[(size_t)l - 1]) best[(size_t)l - 1] = v;
                } else if (v < best[(size_t)l - 1] && v < kMargin * base[(size_t)l - 1]) {
                    best[(size_t)l - 1] = v;
                    pick[(size_t)l] = opts[k];
                }
            }
        }
        return ORB_OK;
    };
    const int32_t caps[] = {0, 16384, 4096, 1024};  // the default (0) first
    std::vector<int32_t> pick_cap;
    if ((rc = pick_per_level(c.tune_chunk_blocks, caps, 4, pick_cap)) < 0) return finish(rc);
    c.tune_chunk_blocks = pick_cap;

    // 3. subtree grid per level (cells per block)
    const int32_t sblocks[] = {0, 256, 1024, 4096};
    std::vector<int32_t> pick_sb;
    if (subtree_max_keys(c) > 0) {
        if ((rc = pick_per_level(c.tune_sub_blocks, sblocks, 4, pick_sb)) < 0) return finish(rc);
        c.tune_sub_blocks = pick_sb;
    }

    // 4. the combination, kept only if it beats the defaults
    double t_tuned = 0.0;
    if ((rc = timed_build(c, snap, levels_mode, repeats, lv, &t_tuned)) < 0) return finish(rc);
    if (t_tuned > t_default) {
        c.sub_max_keys = -1;
        std::fill(c.tune_chunk_blocks.begin(), c.tune_chunk_blocks.end(), 0);
        std::fill(c.tune_sub_blocks.begin(), c.tune_sub_blocks.end(), 0);
        t_tuned = t_default;
        lv = lv_default;
    }
    c.tune_level_ms.assign((size_t)D + 1, 0.0);
    for (int32_t l = 1; l <= D && l <= 64; ++l)
        c.tune_level_ms[(size_t)l] = std::isfinite(lv[(size_t)l - 1]) ? lv[(size_t)l - 1] : 0.0;
    if (out) {
        tuning_from_ctx(c, out);
        out->default_ms = t_default;
        out->tuned_ms = t_tuned;
    }
    return finish(ORB_OK);
}

}  // extern "C"
