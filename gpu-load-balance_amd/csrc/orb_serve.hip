// orb_serve.hip -- serving ORB builds with continuous batching.
//
// Requests of different particle and leaf counts share one device engine: a
// particle pool (SoA x | y | z in HBM), a cell-heap pool and the level-step
// scratch of the batched form.  A request is admitted (FIFO) when a build
// slot is free and both pools can hold it; admission copies its particles
// into its pool range and seeds its heap.  Every orb_server_step runs one
// level of every active build in the same launches: the cells of all builds
// form one level list, each cell carrying its build's heap base, capacity,
// pass counter and status slot (LevelEntry).  A cell's result depends only on
// its own particles and Cell, so a build's result is the same alone and in
// any mix.  When a build's last level is enqueued, its particles, ranges,
// heap and counters are copied to server-owned pinned buffers and its pool
// ranges are released at once (later uploads are stream-ordered after the
// copies).
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <new>

#include "orb_device.hpp"
#include "orb_serve.h"

namespace orb {

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// first-fit allocator over [0, capacity)
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;  // start -> length
    int64_t capacity = 0;
    void init(int64_t cap)
    {
        capacity = cap;
        free_.clear();
        if (cap > 0) free_[0] = cap;
    }
    int64_t alloc(int64_t len, int64_t align)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            int64_t start = (it->first + align - 1) / align * align;
            int64_t end = it->first + it->second;
            if (start + len <= end) {
                int64_t s0 = it->first;
                free_.erase(it);
                if (start > s0) free_[s0] = start - s0;
                if (start + len < end) free_[start + len] = end - (start + len);
                return start;
            }
        }
        return -1;
    }
    void release(int64_t start, int64_t len)
    {
        if (len <= 0) return;
        auto it = free_.emplace(start, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t free_total() const
    {
        int64_t s = 0;
        for (auto &kv : free_) s += kv.second;
        return s;
    }
};

// Pinned host blocks by power-of-two size class, reused across requests so
// that a submit does not pay for a page-locked allocation.
struct PinPool {
    std::multimap<size_t, void *> free_;
    size_t cached = 0;
    size_t limit = size_t(4) << 30;
    static size_t round(size_t bytes)
    {
        size_t s = 4096;
        while (s < bytes) s <<= 1;
        return s;
    }
    void *get(size_t bytes)
    {
        const size_t s = round(bytes);
        auto it = free_.find(s);
        if (it != free_.end()) {
            void *p = it->second;
            free_.erase(it);
            cached -= s;
            return p;
        }
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, s, hipHostMallocDefault);
        if (e != hipSuccess) {
            set_last_hip_error(e);
            return nullptr;
        }
        return p;
    }
    void put(void *p, size_t bytes)
    {
        if (!p) return;
        const size_t s = round(bytes);
        if (cached + s > limit) {
            (void)hipHostFree(p);
            return;
        }
        free_.emplace(s, p);
        cached += s;
    }
    void clear()
    {
        for (auto &kv : free_) (void)hipHostFree(kv.second);
        free_.clear();
        cached = 0;
    }
};

enum { kPending = 0, kActive = 1, kFinishing = 2, kDone = 3 };

struct Request {
    uint64_t id = 0;
    int32_t slot = -1;
    int64_t n = 0;
    int32_t d = 1;
    int64_t cap = 0;
    int64_t pb = -1, hb = -1;    // pool bases
    std::vector<std::vector<int32_t>> levels;
    bool complete = true;
    int32_t lmax = 0;
    int32_t next_level = 1;
    int state = kPending;
    int error = 0;
    int32_t error_cell = -1;
    float *dev[3] = {nullptr, nullptr, nullptr};  // device request: the caller's arrays
    // pinned blocks: particles (3n floats: in at submit, out at finish; host
    // requests only) and the tree (ranges [cap][2] pool-absolute, heap,
    // exists, 64 level passes + 2 status words)
    float *xyz = nullptr;
    size_t xyz_bytes = 0;
    uint8_t *tree = nullptr;
    size_t tree_bytes = 0;
    uint32_t *ranges = nullptr;
    orb_cell *heap = nullptr;
    uint8_t *exists = nullptr;
    int32_t *counters = nullptr;
    hipEvent_t done = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

}  // namespace

struct Server {
    Ctx c;                        // engine: particle pool = c.pos, ranges pool = c.ranges
    orb_server_config cfg{};
    DevBuf<orb_cell> heap;
    DevBuf<uint8_t> exists, sub;
    DevBuf<uint16_t> chunk_hist;
    DevBuf<uint32_t> cell_hist;
    DevBuf<int32_t> level_passes; // [slots][64]
    DevBuf<int32_t> status;       // [slots][2]
    DevBuf<LevelEntry> entries;
    PinBuf<LevelEntry> hentries[2];
    hipEvent_t ev_entries[2] = {nullptr, nullptr};
    int flip = 0;
    RangeAlloc palloc, halloc;
    std::vector<int32_t> free_slots;
    std::deque<Request *> pending;
    std::vector<Request *> active, finishing;
    std::deque<uint64_t> done_ids;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    PinPool pins;
    uint64_t next_id = 1;
    orb_server_stats stats{};
};

__global__ void k_s_init(orb_cell *heap, uint8_t *exists, uint8_t *sub, uint32_t *ranges,
                         int64_t hb, int64_t cap, orb_cell root, uint32_t pb, uint32_t n,
                         int32_t *passes, int32_t *status)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < cap; k += stride) {
        store_root_or_zero(&heap[hb + k], root, k == 0);
        exists[hb + k] = k == 0 ? 1 : 0;
        sub[hb + k] = 0;
        ranges[2 * (hb + k)] = k == 0 ? pb : 0u;
        ranges[2 * (hb + k) + 1] = k == 0 ? pb + n : 0u;
    }
    if (i < 64) passes[i] = 0;
    if (i < 2) status[i] = 0;
}

static int server_alloc(Server &S)
{
    Ctx &c = S.c;
    const int64_t P = S.cfg.particle_capacity, H = S.cfg.cell_capacity;
    const int32_t slots = S.cfg.max_active;
    c.n = P;
    c.stride = ((P + 63) / 64) * 64;
    c.cap = H;
    c.d = 1;
    ORB_TRY(c.pos.alloc(3 * (size_t)c.stride + 64));
    ORB_TRY(c.ranges.alloc(2 * (size_t)H));
    ORB_TRY(c.packed.alloc((size_t)c.stride));
    ORB_TRY(c.sel.alloc((size_t)c.stride));
    ORB_TRY(c.status.alloc(2));
    ORB_TRY(c.hstatus.alloc(2));
    ORB_TRY(c.hranges.alloc(2));
    ORB_TRY(ensure_level_capacity(c, H / 2 + 1));
    ORB_TRY(preallocate_scratch(c));
    ORB_TRY(S.heap.alloc((size_t)H));
    ORB_TRY(S.exists.alloc((size_t)H));
    ORB_TRY(S.sub.alloc((size_t)H));
    ORB_TRY(S.chunk_hist.alloc(kHistRegions * (size_t)(c.max_chunks + 1) * kBins));
    ORB_TRY(S.cell_hist.alloc((size_t)c.max_level_cells * 256));
    ORB_TRY(S.level_passes.alloc((size_t)slots * 64));
    ORB_TRY(S.status.alloc((size_t)slots * 2));
    ORB_TRY(S.entries.alloc((size_t)c.max_level_cells));
    ORB_TRY(S.hentries[0].alloc((size_t)c.max_level_cells));
    ORB_TRY(S.hentries[1].alloc((size_t)c.max_level_cells));
    ORB_TRY(launch_zero_u32(c.stream, S.cell_hist.p, (size_t)c.max_level_cells * 256));
    ORB_HIP(hipStreamSynchronize(c.stream));
    S.palloc.init(c.stride);
    S.halloc.init(H);
    for (int32_t s = slots - 1; s >= 0; --s) S.free_slots.push_back(s);
    return ORB_OK;
}

static int admit(Server &S)
{
    Ctx &c = S.c;
    while (!S.pending.empty() && !S.free_slots.empty()) {
        Request *r = S.pending.front();
        int64_t pb = S.palloc.alloc(r->n > 0 ? r->n : 1, 64);
        if (pb < 0) break;
        int64_t hb = S.halloc.alloc(r->cap, 1);
        if (hb < 0) {
            S.palloc.release(pb, r->n > 0 ? r->n : 1);
            break;
        }
        S.pending.pop_front();
        r->pb = pb;
        r->hb = hb;
        r->slot = S.free_slots.back();
        S.free_slots.pop_back();
        if (r->n > 0) {
            const size_t nb = sizeof(float) * (size_t)r->n;
            float *dst[3] = {c.x() + pb, c.y() + pb, c.z() + pb};
            for (int a = 0; a < 3; ++a) {
                if (r->dev[0])
                    ORB_HIP(hipMemcpyAsync(dst[a], r->dev[a], nb, hipMemcpyDeviceToDevice, c.stream));
                else
                    ORB_HIP(hipMemcpyAsync(dst[a], r->xyz + a * r->n, nb, hipMemcpyHostToDevice,
                                           c.stream));
            }
        }
        const orb_cell root = cell_root(r->d);
        unsigned g = (unsigned)((r->cap + kBlock - 1) / kBlock);
        if (g > 4096) g = 4096;
        hipLaunchKernelGGL(k_s_init, dim3(g), dim3(kBlock), 0, c.stream, S.heap.p, S.exists.p,
                           S.sub.p, c.ranges.p, (int64_t)hb, (int64_t)r->cap, root, (uint32_t)pb,
                           (uint32_t)r->n, S.level_passes.p + 64 * r->slot,
                           S.status.p + 2 * r->slot);
        ORB_HIP(hipGetLastError());
        // a NaN coordinate fails the request with ORB_ERR_ARG (status word
        // kErrInputWord, above every undefined-cell word)
        ORB_TRY(launch_find_nan(c.stream, c.x() + pb, c.y() + pb, c.z() + pb, (size_t)r->n,
                                S.status.p + 2 * r->slot));
        r->state = kActive;
        r->t_admit = now_ms();
        S.active.push_back(r);
        S.stats.admitted++;
    }
    return ORB_OK;
}

static int finish(Server &S, Request *r)
{
    Ctx &c = S.c;
    const size_t n = (size_t)r->n, cap = (size_t)r->cap;
    if (n > 0) {
        const float *src[3] = {c.x() + r->pb, c.y() + r->pb, c.z() + r->pb};
        for (int a = 0; a < 3; ++a) {
            if (r->dev[0])
                ORB_HIP(hipMemcpyAsync(r->dev[a], src[a], sizeof(float) * n,
                                       hipMemcpyDeviceToDevice, c.stream));
            else
                ORB_HIP(hipMemcpyAsync(r->xyz + a * n, src[a], sizeof(float) * n,
                                       hipMemcpyDeviceToHost, c.stream));
        }
    }
    ORB_HIP(hipMemcpyAsync(r->ranges, c.ranges.p + 2 * r->hb, 2 * sizeof(uint32_t) * cap,
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipMemcpyAsync(r->heap, S.heap.p + r->hb, sizeof(orb_cell) * cap,
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipMemcpyAsync(r->exists, S.exists.p + r->hb, cap, hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipMemcpyAsync(r->counters, S.level_passes.p + 64 * r->slot, 64 * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipMemcpyAsync(r->counters + 64, S.status.p + 2 * r->slot, 2 * sizeof(int32_t),
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipEventRecord(r->done, c.stream));
    // later work on these pool ranges is stream-ordered after the copies
    S.palloc.release(r->pb, r->n > 0 ? r->n : 1);
    S.halloc.release(r->hb, r->cap);
    S.free_slots.push_back(r->slot);
    r->state = kFinishing;
    S.finishing.push_back(r);
    return ORB_OK;
}

static int poll_finishing(Server &S)
{
    for (size_t i = 0; i < S.finishing.size();) {
        Request *r = S.finishing[i];
        hipError_t e = hipEventQuery(r->done);
        if (e == hipSuccess) {
            r->state = kDone;
            r->t_done = now_ms();
            S.done_ids.push_back(r->id);
            int32_t bad = -1;
            if (status_error(r->counters + 64, &bad))
                r->error = bad < 0 ? ORB_ERR_ARG : ORB_ERR_UB;
            r->error_cell = bad;
            S.stats.completed++;
            S.finishing[i] = S.finishing.back();
            S.finishing.pop_back();
        } else if (e == hipErrorNotReady) {
            ++i;
        } else {
            set_last_hip_error(e);
            return ORB_ERR_HIP;
        }
    }
    return ORB_OK;
}

static void release_request(Server &S, Request *r)
{
    S.pins.put(r->xyz, r->xyz_bytes);
    S.pins.put(r->tree, r->tree_bytes);
    r->xyz = nullptr;
    r->tree = nullptr;
    if (r->done) (void)hipEventDestroy(r->done);
    r->done = nullptr;
}

static int submit(Server &S, const float *hx, const float *hy, const float *hz, float *dx,
                  float *dy, float *dz, int64_t n, int32_t d, uint64_t *id)
{
    const double t0 = now_ms();
    const int64_t cap = heap_capacity(d);
    if (n > S.palloc.capacity || cap > S.halloc.capacity || n > (int64_t(1) << 31))
        return ORB_ERR_BUSY;  // could never be admitted
    auto r = std::unique_ptr<Request>(new (std::nothrow) Request);
    if (!r) return ORB_ERR_NOMEM;
    r->n = n;
    r->d = d;
    r->cap = cap;
    ORB_TRY(schedule_levels(d, S.cfg.levels_mode, r->cap, r->levels, &r->complete));
    const int32_t D = n_levels(d);
    r->lmax = S.cfg.levels_mode == ORB_LEVELS_REFERENCE ? D - 1 : D;
    // tree block: ranges | heap | exists | counters, each 256-byte aligned
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t o_heap = up(2 * sizeof(uint32_t) * (size_t)cap);
    const size_t o_ex = o_heap + up(sizeof(orb_cell) * (size_t)cap);
    const size_t o_cnt = o_ex + up((size_t)cap);
    r->tree_bytes = o_cnt + 66 * sizeof(int32_t);
    r->tree = static_cast<uint8_t *>(S.pins.get(r->tree_bytes));
    if (!r->tree) return ORB_ERR_NOMEM;
    r->ranges = reinterpret_cast<uint32_t *>(r->tree);
    r->heap = reinterpret_cast<orb_cell *>(r->tree + o_heap);
    r->exists = r->tree + o_ex;
    r->counters = reinterpret_cast<int32_t *>(r->tree + o_cnt);
    if (dx) {
        r->dev[0] = dx;
        r->dev[1] = dy;
        r->dev[2] = dz;
    } else if (n > 0) {
        r->xyz_bytes = 3 * sizeof(float) * (size_t)n;
        r->xyz = static_cast<float *>(S.pins.get(r->xyz_bytes));
        if (!r->xyz) {
            release_request(S, r.get());
            return ORB_ERR_NOMEM;
        }
        std::memcpy(r->xyz, hx, sizeof(float) * (size_t)n);
        std::memcpy(r->xyz + n, hy, sizeof(float) * (size_t)n);
        std::memcpy(r->xyz + 2 * n, hz, sizeof(float) * (size_t)n);
    }
    if (hipEventCreateWithFlags(&r->done, hipEventDisableTiming) != hipSuccess) {
        release_request(S, r.get());
        return ORB_ERR_HIP;
    }
    r->id = S.next_id++;
    r->t_submit = t0;
    *id = r->id;
    S.pending.push_back(r.get());
    S.all[r->id] = std::move(r);
    S.stats.submitted++;
    return ORB_OK;
}

static int step(Server &S)
{
    Ctx &c = S.c;
    ORB_TRY(admit(S));
    if (!S.active.empty()) {
        // this step's level of every active build, in one list
        LevelEntry *h = S.hentries[S.flip].p;
        ORB_HIP(hipEventSynchronize(S.ev_entries[S.flip]));  // staging buffer free again
        int64_t m = 0;
        bool any_sub = false;
        std::vector<Request *> stepped;
        for (Request *r : S.active) {
            while (r->next_level <= r->lmax && r->levels[(size_t)r->next_level].empty())
                r->next_level++;
            if (r->next_level > r->lmax) continue;
            const int32_t l = r->next_level;
            const auto &ids = r->levels[(size_t)l];
            if (m + (int64_t)ids.size() > c.max_level_cells) break;  // next step takes the rest
            for (int32_t id : ids)
                h[m++] = LevelEntry{id, (int32_t)r->hb, (int32_t)r->cap, 64 * r->slot + l,
                                    2 * r->slot, r->lmax, r->complete ? 1 : 0, 0};
            any_sub = any_sub || r->complete;
            stepped.push_back(r);
        }
        if (m > 0) {
            ORB_HIP(hipMemcpyAsync(S.entries.p, h, sizeof(LevelEntry) * (size_t)m,
                                   hipMemcpyHostToDevice, c.stream));
            ORB_HIP(hipEventRecord(S.ev_entries[S.flip], c.stream));
            S.flip ^= 1;
            ORB_TRY(enqueue_level(c, S.entries.p, (int32_t)m, S.heap.p, S.exists.p, S.sub.p,
                                  S.chunk_hist.p, S.cell_hist.p, nullptr, S.level_passes.p,
                                  S.status.p, any_sub, false, -1));
            S.stats.steps++;
            S.stats.cells_stepped += m;
            if ((int64_t)stepped.size() > S.stats.max_batch) S.stats.max_batch = (int64_t)stepped.size();
        }
        for (Request *r : stepped) r->next_level++;
        // builds with no level left: results out, pool ranges back
        for (size_t i = 0; i < S.active.size();) {
            Request *r = S.active[i];
            while (r->next_level <= r->lmax && r->levels[(size_t)r->next_level].empty())
                r->next_level++;
            if (r->next_level > r->lmax) {
                ORB_TRY(finish(S, r));
                S.active[i] = S.active.back();
                S.active.pop_back();
            } else {
                ++i;
            }
        }
    }
    ORB_TRY(poll_finishing(S));
    return (int)(S.pending.size() + S.active.size() + S.finishing.size());
}

}  // namespace orb

using namespace orb;

struct orb_server {
    Server s;
};

extern "C" {

int orb_server_create(orb_server **out, const orb_server_config *cfg)
{
    if (!out || !cfg) return ORB_ERR_ARG;
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->max_active > 4096 || cfg->levels_mode < 0 ||
        cfg->levels_mode > 1 || cfg->particle_capacity < 0 || cfg->cell_capacity < 0)
        return ORB_ERR_ARG;
    if (orb_device_count() < 1) return ORB_ERR_NODEV;
    orb_server *o = new (std::nothrow) orb_server;
    if (!o) return ORB_ERR_NOMEM;
    Server &S = o->s;
    S.cfg = *cfg;
    Ctx &c = S.c;
    auto fail = [&](int rc) {
        orb_server_destroy(o);
        return rc;
    };
    if (hipGetDevice(&c.device) != hipSuccess) return fail(ORB_ERR_HIP);
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess)
        return fail(ORB_ERR_HIP);
    for (auto &e : S.ev_entries)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(ORB_ERR_HIP);
    if (hipEventCreateWithFlags(&c.ev_upload, hipEventDisableTiming) != hipSuccess)
        return fail(ORB_ERR_HIP);
    for (int k = 0; k < 2; ++k)
        if (hipEventRecord(S.ev_entries[k], c.stream) != hipSuccess) return fail(ORB_ERR_HIP);
    c.cmp = cfg->cmp == ORB_CMP_LE ? ORB_CMP_LE : ORB_CMP_LT;
    c.part = cfg->partition == ORB_PART_STABLE_LE ? ORB_PART_STABLE_LE : ORB_PART_HOARE;
    c.use_graph = false;
    if (S.cfg.particle_capacity == 0 || S.cfg.cell_capacity == 0) {
        // admission by free HBM: the pools take a share of what is free now
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(ORB_ERR_HIP);
        // per particle: x, y, z, two position words, chunk scratch ~ 24 B
        if (S.cfg.particle_capacity == 0)
            S.cfg.particle_capacity = (int64_t)((double)fr * 0.4 / 24.0);
        if (S.cfg.cell_capacity == 0) S.cfg.cell_capacity = int64_t(1) << 22;
        if (S.cfg.particle_capacity > (int64_t(1) << 31)) S.cfg.particle_capacity = int64_t(1) << 31;
    }
    int rc = server_alloc(S);
    if (rc < 0) return fail(rc);
    *out = o;
    return ORB_OK;
}

int orb_server_destroy(orb_server *o)
{
    if (!o) return ORB_ERR_ARG;
    Server &S = o->s;
    Ctx &c = S.c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (auto &kv : S.all) release_request(S, kv.second.get());
    S.all.clear();
    S.pins.clear();
    for (auto &e : S.ev_entries)
        if (e) (void)hipEventDestroy(e);
    S.heap.release();
    S.exists.release();
    S.sub.release();
    S.chunk_hist.release();
    S.cell_hist.release();
    S.level_passes.release();
    S.status.release();
    S.entries.release();
    S.hentries[0].release();
    S.hentries[1].release();
    // engine buffers
    c.pos.release();
    c.ranges.release();
    c.packed.release();
    c.masks.release();
    c.cell_cand.release();
    c.status.release();
    c.hstatus.release();
    c.hranges.release();
    c.dcells.release();
    c.recs.release();
    c.cell_nchunks.release();
    c.cell_first.release();
    c.chunk_cell.release();
    c.chunk_desc.release();
    c.cnt.release();
    c.stamps.release();
    c.chunk_ltle.release();
    c.chunk_cnt.release();
    c.chunk_pre.release();
    c.scan_partials.release();
    c.splits.release();
    c.chunk_comp.release();
    c.sel.release();
    c.sel_lists.release();
    c.swap_plan.release();
    c.ids_dev.release();
    c.hcells.release();
    c.hcnt.release();
    c.tmp.release();
    if (c.ev_upload) (void)hipEventDestroy(c.ev_upload);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete o;
    return ORB_OK;
}

int orb_server_submit(orb_server *o, const float *x, const float *y, const float *z, int64_t n,
                      int32_t d, uint64_t *id)
{
    if (!o || !id || n < 0 || d < 1 || (n > 0 && (!x || !y || !z))) return ORB_ERR_ARG;
    return submit(o->s, x, y, z, nullptr, nullptr, nullptr, n, d, id);
}

int orb_server_submit_device(orb_server *o, float *x, float *y, float *z, int64_t n, int32_t d,
                             uint64_t *id)
{
    if (!o || !id || n < 0 || d < 1 || !x || !y || !z) return ORB_ERR_ARG;
    return submit(o->s, nullptr, nullptr, nullptr, x, y, z, n, d, id);
}

int orb_server_step(orb_server *o)
{
    if (!o) return ORB_ERR_ARG;
    return step(o->s);
}

int orb_server_poll(orb_server *o, uint64_t id)
{
    if (!o) return ORB_ERR_ARG;
    Server &S = o->s;
    ORB_TRY(poll_finishing(S));
    auto it = S.all.find(id);
    if (it == S.all.end()) return ORB_ERR_ARG;
    return it->second->state == kDone ? 1 : 0;
}

int orb_server_next_done(orb_server *o, uint64_t *id)
{
    if (!o || !id) return ORB_ERR_ARG;
    Server &S = o->s;
    ORB_TRY(poll_finishing(S));
    while (!S.done_ids.empty()) {
        const uint64_t v = S.done_ids.front();
        S.done_ids.pop_front();
        if (S.all.count(v)) {  // not fetched already
            *id = v;
            return 1;
        }
    }
    return 0;
}

int orb_server_fetch(orb_server *o, uint64_t id, float *x, float *y, float *z, uint32_t *ranges,
                     orb_cell *heap, uint8_t *exists, int64_t cap, orb_request_info *info)
{
    if (!o) return ORB_ERR_ARG;
    Server &S = o->s;
    auto it = S.all.find(id);
    if (it == S.all.end()) return ORB_ERR_ARG;
    Request *r = it->second.get();
    if (r->state != kDone) return ORB_ERR_STATE;
    if ((heap || ranges || exists) && cap < r->cap) return ORB_ERR_ARG;
    const size_t n = (size_t)r->n;
    float *dst[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (!dst[a] || n == 0) continue;
        if (r->dev[0]) ORB_HIP(hipMemcpy(dst[a], r->dev[a], sizeof(float) * n, hipMemcpyDeviceToHost));
        else std::memcpy(dst[a], r->xyz + a * n, sizeof(float) * n);
    }
    if (heap) {
        std::memset(heap, 0, sizeof(orb_cell) * (size_t)cap);
        std::memcpy(heap, r->heap, sizeof(orb_cell) * (size_t)r->cap);
    }
    if (exists) {
        std::memset(exists, 0, (size_t)cap);
        std::memcpy(exists, r->exists, (size_t)r->cap);
    }
    if (ranges) {
        // pool-absolute ranges back to the build's own [0, n)
        std::memset(ranges, 0, 2 * sizeof(uint32_t) * (size_t)cap);
        for (int64_t k = 0; k < r->cap; ++k) {
            if (!r->exists[k]) continue;
            ranges[2 * k] = r->ranges[2 * k] - (uint32_t)r->pb;
            ranges[2 * k + 1] = r->ranges[2 * k + 1] - (uint32_t)r->pb;
        }
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->n = r->n;
        info->d = r->d;
        info->error = r->error;
        info->error_cell = r->error ? r->error_cell : -1;
        for (int32_t l = 1; l <= r->lmax && l < 64; ++l) {
            info->nlevels = l;
            if (r->levels[(size_t)l].empty()) continue;  // as orb_build reports it
            info->level_passes[l - 1] = r->counters[l];
            info->total_passes += r->counters[l];
        }
        info->submit_ms = r->t_submit;
        info->admit_ms = r->t_admit;
        info->done_ms = r->t_done;
    }
    int err = r->error;
    release_request(S, r);
    S.all.erase(it);
    return err;
}

int orb_server_get_stats(orb_server *o, orb_server_stats *st)
{
    if (!o || !st) return ORB_ERR_ARG;
    Server &S = o->s;
    *st = S.stats;
    st->pending = (int64_t)S.pending.size();
    st->active = (int64_t)S.active.size();
    st->finishing = (int64_t)S.finishing.size();
    st->particle_capacity = S.cfg.particle_capacity;
    st->particles_free = S.palloc.free_total();
    st->cell_capacity = S.cfg.cell_capacity;
    st->cells_free = S.halloc.free_total();
    return ORB_OK;
}

}  // extern "C"
