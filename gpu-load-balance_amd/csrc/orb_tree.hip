// orb_tree.hip -- the built tree as a consumer needs it: leaves with their
// particle ranges, per-particle leaf membership, tight leaf bounding boxes.
#include <algorithm>
#include <vector>

#include "orb_device.hpp"

namespace orb {

struct LeafRec {
    uint32_t begin, end;
    int32_t id;
    uint32_t pad;
};

// Leaves of the last build from the host heap and the device range map.
static int collect_leaves(Ctx &c, std::vector<LeafRec> &leaves)
{
    ORB_TRY(batched_pull_results(c));
    if (c.exists.size() != (size_t)c.cap) return ORB_ERR_STATE;  // no build yet
    std::vector<uint32_t> rg(2 * (size_t)c.cap);
    ORB_HIP(hipMemcpyAsync(rg.data(), c.ranges.p, sizeof(uint32_t) * rg.size(),
                           hipMemcpyDeviceToHost, c.stream));
    ORB_HIP(hipStreamSynchronize(c.stream));
    leaves.clear();
    for (int64_t id = 0; id < c.cap; ++id) {
        if (!c.exists[(size_t)id]) continue;
        const int64_t l = left_child_id((int32_t)id), r = right_child_id((int32_t)id);
        const bool split = (l < c.cap && c.exists[(size_t)l]) || (r < c.cap && c.exists[(size_t)r]);
        if (split) continue;
        leaves.push_back(LeafRec{rg[2 * (size_t)id], rg[2 * (size_t)id + 1], (int32_t)id, 0u});
    }
    return ORB_OK;
}

__global__ void k_fill_i32(int32_t *p, size_t n, int32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = v;
}

// one block per leaf (grid-stride over leaves)
__global__ void k_leaf_fill(const LeafRec *leaves, int64_t nLeaves, int32_t *leaf_of, uint32_t n)
{
    for (int64_t j = blockIdx.x; j < nLeaves; j += gridDim.x) {
        const LeafRec lr = leaves[j];
        const uint32_t e = lr.end < n ? lr.end : n;
        for (uint32_t p = lr.begin + threadIdx.x; p < e; p += blockDim.x) leaf_of[p] = lr.id;
    }
}

__global__ void k_by_input_order(const int32_t *leaf_of, const uint32_t *pid, uint32_t n,
                                 int32_t *out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t p = i; p < n; p += stride) {
        const uint32_t o = pid[p];
        if (o < n) out[o] = leaf_of[p];
    }
}

// Segmented min/max: one block per leaf, a strided pass per thread, then
// wave shuffles and four LDS partials.
__global__ __launch_bounds__(kBlock) void k_leaf_bounds(const LeafRec *leaves, int64_t nLeaves,
                                                        const float *x, const float *y,
                                                        const float *z, uint32_t n, float *lower,
                                                        float *upper)
{
    __shared__ float smin[4][3], smax[4][3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t j = blockIdx.x; j < nLeaves; j += gridDim.x) {
        const LeafRec lr = leaves[j];
        const uint32_t e = lr.end < n ? lr.end : n;
        float mn[3] = {INFINITY, INFINITY, INFINITY};
        float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t p = lr.begin + threadIdx.x; p < e; p += kBlock) {
            const float v[3] = {x[p], y[p], z[p]};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                mn[a] = fminf(mn[a], v[a]);
                mx[a] = fmaxf(mx[a], v[a]);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
                mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                smin[w][a] = mn[a];
                smax[w][a] = mx[a];
            }
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int a = threadIdx.x;
            float lo = smin[0][a], hi = smax[0][a];
            for (int q = 1; q < 4; ++q) {
                lo = fminf(lo, smin[q][a]);
                hi = fmaxf(hi, smax[q][a]);
            }
            lower[3 * j + a] = lo;
            upper[3 * j + a] = hi;
        }
        __syncthreads();  // smin/smax are rewritten for the next leaf
    }
}

static unsigned grid_for(int64_t items)
{
    int64_t g = items < 1 ? 1 : items;
    return (unsigned)(g > 8192 ? 8192 : g);
}

}  // namespace orb

using namespace orb;

extern "C" {

int orb_get_leaves(orb_ctx *ctx, int32_t *ids, uint32_t *ranges, int64_t cap, int64_t *count)
{
    if (!ctx || !count) return ORB_ERR_ARG;
    Ctx &c = ctx->c;
    std::vector<LeafRec> leaves;
    ORB_TRY(collect_leaves(c, leaves));
    *count = (int64_t)leaves.size();
    if (!ids && !ranges) return ORB_OK;
    if (cap < (int64_t)leaves.size()) return ORB_ERR_ARG;
    for (size_t j = 0; j < leaves.size(); ++j) {
        if (ids) ids[j] = leaves[j].id;
        if (ranges) {
            ranges[2 * j] = leaves[j].begin;
            ranges[2 * j + 1] = leaves[j].end;
        }
    }
    return ORB_OK;
}

int orb_get_membership(orb_ctx *ctx, int32_t *leaf_of, int64_t n, bool by_input_order)
{
    if (!ctx || (n > 0 && !leaf_of)) return ORB_ERR_ARG;
    Ctx &c = ctx->c;
    if (n != c.n) return ORB_ERR_ARG;
    if (by_input_order && !c.track_ids) return ORB_ERR_STATE;
    std::vector<LeafRec> leaves;
    ORB_TRY(collect_leaves(c, leaves));
    if (n == 0) return ORB_OK;
    DevBuf<LeafRec> dl;
    DevBuf<int32_t> mem, mem2;
    auto cleanup = [&](int rc) {
        dl.release();
        mem.release();
        mem2.release();
        return rc;
    };
    int rc;
    if ((rc = dl.alloc(leaves.size())) < 0 || (rc = mem.alloc((size_t)n)) < 0) return cleanup(rc);
    if (!leaves.empty() &&
        hipMemcpyAsync(dl.p, leaves.data(), sizeof(LeafRec) * leaves.size(), hipMemcpyHostToDevice,
                       c.stream) != hipSuccess)
        return cleanup(ORB_ERR_HIP);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n / kBlock + 1)), dim3(kBlock), 0, c.stream,
                       mem.p, (size_t)n, -1);
    hipLaunchKernelGGL(k_leaf_fill, dim3(grid_for((int64_t)leaves.size())), dim3(kBlock), 0,
                       c.stream, dl.p, (int64_t)leaves.size(), mem.p, (uint32_t)n);
    int32_t *src = mem.p;
    if (by_input_order) {
        if ((rc = mem2.alloc((size_t)n)) < 0) return cleanup(rc);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n / kBlock + 1)), dim3(kBlock), 0, c.stream,
                           mem2.p, (size_t)n, -1);
        hipLaunchKernelGGL(k_by_input_order, dim3(grid_for(n / kBlock + 1)), dim3(kBlock), 0,
                           c.stream, mem.p, c.ids(), (uint32_t)n, mem2.p);
        src = mem2.p;
    }
    if (hipGetLastError() != hipSuccess) return cleanup(ORB_ERR_HIP);
    if (hipMemcpyAsync(leaf_of, src, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                       c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return cleanup(ORB_ERR_HIP);
    return cleanup(ORB_OK);
}

int orb_get_leaf_bounds(orb_ctx *ctx, float *lower, float *upper, int64_t cap)
{
    if (!ctx || !lower || !upper) return ORB_ERR_ARG;
    Ctx &c = ctx->c;
    std::vector<LeafRec> leaves;
    ORB_TRY(collect_leaves(c, leaves));
    const int64_t m = (int64_t)leaves.size();
    if (cap < m) return ORB_ERR_ARG;
    if (m == 0) return ORB_OK;
    DevBuf<LeafRec> dl;
    DevBuf<float> lo, hi;
    auto cleanup = [&](int rc) {
        dl.release();
        lo.release();
        hi.release();
        return rc;
    };
    int rc;
    if ((rc = dl.alloc((size_t)m)) < 0 || (rc = lo.alloc(3 * (size_t)m)) < 0 ||
        (rc = hi.alloc(3 * (size_t)m)) < 0)
        return cleanup(rc);
    if (hipMemcpyAsync(dl.p, leaves.data(), sizeof(LeafRec) * (size_t)m, hipMemcpyHostToDevice,
                       c.stream) != hipSuccess)
        return cleanup(ORB_ERR_HIP);
    hipLaunchKernelGGL(k_leaf_bounds, dim3(grid_for(m)), dim3(kBlock), 0, c.stream, dl.p, m,
                       c.x(), c.y(), c.z(), (uint32_t)c.n, lo.p, hi.p);
    if (hipGetLastError() != hipSuccess) return cleanup(ORB_ERR_HIP);
    if (hipMemcpyAsync(lower, lo.p, sizeof(float) * 3 * (size_t)m, hipMemcpyDeviceToHost,
                       c.stream) != hipSuccess ||
        hipMemcpyAsync(upper, hi.p, sizeof(float) * 3 * (size_t)m, hipMemcpyDeviceToHost,
                       c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return cleanup(ORB_ERR_HIP);
    return cleanup(ORB_OK);
}

}  // extern "C"
