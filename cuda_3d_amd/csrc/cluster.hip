// cluster.hip -- Euclidean clustering as connected components of the radius graph.
//
// Reference: cuda_euclidean_clustering (src/geometry/cuda_clustering.cu:31-110) downloads the
// radius-neighbour lists and runs a breadth-first search on the host, seeding from the lowest
// unvisited index; a cluster is kept when it has at least min_cluster_size points and clusters
// are numbered in the order of their lowest index.
//
// Here the components come from a lock-free union-find on the device:
//   1. init     parent[i] = i, size[i] = 0
//   2. hook     every point, walked in cell order for locality, links its root and the root of
//               each lower-indexed neighbour within the tolerance (the radius search's own
//               predicate: squared FP64 distance < r^2); a root is only ever linked to a smaller
//               one by compare-and-swap, so every tree's root is its component's lowest index
//   3. compress parent[i] = root (a read-only walk, then a store to the point's own entry),
//               size[root] += 1 (one atomic per wave and root run)
//   4. rank     kept roots (size >= min) are numbered by an exclusive scan in index order
//   5. label    label[i] = number of its root's cluster, or -1
// Parent pointers only ever decrease and always name an ancestor, so the path-halving stores in
// find() may race with the compare-and-swaps without losing a link.
//
// One difference: the reference's search stops growing a cluster once its queue holds
// max_cluster_size points, leaving the rest of that component to later seeds in an order that
// depends on the neighbour lists.  Components up to max_cluster_size points come out identical;
// larger ones are kept whole and counted in n_oversized.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

namespace {

constexpr int kCcBlock = 256;

// Workgroup scope lets the load hit L1.  A stale parent is still an ancestor and pointers only
// decrease, so walks on stale values end, and every link is validated by an agent-scope
// compare-and-swap; each failed one strictly lowers the larger of the two roots.
__device__ __forceinline__ int32_t cc_load(int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of x with path halving.
__device__ __forceinline__ int32_t cc_find(int32_t *parent, int32_t x)
{
    int32_t y = cc_load(parent + x);
    while (y != x) {
        const int32_t z = cc_load(parent + y);
        if (z == y) return y;
        __hip_atomic_store(parent + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = z;
        y = cc_load(parent + x);
    }
    return x;
}

// Root of x without writes: the compress pass stores each point's root into its own entry
// only, so a halving store from another thread can never replace a root with an ancestor.
__device__ __forceinline__ int32_t cc_root(int32_t *parent, int32_t x)
{
    int32_t y = cc_load(parent + x);
    while (y != x) {
        x = y;
        y = cc_load(parent + x);
    }
    return x;
}

// Links the trees of u and v (any nodes of them; the point's own root is passed in).
__device__ __forceinline__ void cc_union(int32_t *parent, int32_t u, int32_t v)
{
    while (u != v) {
        if (u > v) {
            const int32_t t = u;
            u = v;
            v = t;
        }
        int32_t expect = v;  // link the larger root under the smaller one
        if (__hip_atomic_compare_exchange_strong(parent + v, &expect, u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
        u = cc_find(parent, u);
        v = cc_find(parent, expect);
    }
}

__global__ __launch_bounds__(kCcBlock) void k_cc_init(const CloudHdr *__restrict__ h, int32_t *__restrict__ parent,
                                                     uint32_t *__restrict__ size, uint32_t *__restrict__ oversized)
{
    const int n = h->n;
    if (blockIdx.x == 0 && threadIdx.x == 0) *oversized = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        parent[i] = i;
        size[i] = 0;
    }
}

__global__ __launch_bounds__(kCcBlock) void k_cc_hook(const CloudHdr *__restrict__ h, GridDesc g,
                                                     const uint32_t *__restrict__ start,
                                                     const float4 *__restrict__ sxyzi,
                                                     const int32_t *__restrict__ order, int32_t *parent)
{
    const int n = h->n;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double r2 = (double)g.radius * (double)g.radius;
    const float4 p = sxyzi[s];
    const int32_t i = order[s];
    int32_t lo[3], hi[3];
    if (!grid_range(g, p.x, p.y, p.z, lo, hi)) return;
    int32_t ri = cc_find(parent, i);  // this point's root, refreshed after each link
    for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
        for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
            const int64_t row = ((int64_t)ix * g.dims[1] + iy) * g.dims[2];
            const uint32_t e = start[row + hi[2] + 1];
            for (uint32_t t = start[row + lo[2]]; t < e; ++t) {
                const float4 q = sxyzi[t];
                const int32_t j = order[t];
                if (j < i && dist2(q.x, q.y, q.z, p.x, p.y, p.z) < r2) {
                    const int32_t rj = cc_find(parent, j);
                    if (rj != ri) {
                        ri = cc_find(parent, ri);  // short walk: ri was a root not long ago
                        if (rj != ri) {
                            cc_union(parent, ri, rj);  // a compare-and-swap only for distinct roots
                            ri = cc_find(parent, ri);
                        }
                    }
                }
            }
        }
}

__global__ __launch_bounds__(kCcBlock) void k_cc_compress(const CloudHdr *__restrict__ h, int32_t *parent,
                                                         uint32_t *__restrict__ size)
{
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < n;
    const int32_t r = on ? cc_root(parent, i) : -1;
    if (on) __hip_atomic_store(parent + i, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // lanes sharing a root add once: the lowest such lane adds the run's count
    uint64_t pending = __ballot(on);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int32_t lr = __shfl(r, leader);
        const uint64_t same = __ballot(on && r == lr);
        if (lane_id() == leader) atomicAdd(size + lr, (uint32_t)__popcll(same));
        pending &= ~same;
    }
}

// kept roots: flag + per-block count for the scan; oversized kept roots counted
__global__ __launch_bounds__(kCcBlock) void k_cc_flag(const CloudHdr *__restrict__ h,
                                                     const int32_t *__restrict__ parent,
                                                     const uint32_t *__restrict__ size, uint32_t min_size,
                                                     uint32_t max_size, uint8_t *__restrict__ kept,
                                                     uint32_t *__restrict__ bcnt, uint32_t *__restrict__ oversized)
{
    __shared__ uint32_t wc[kCcBlock / kWave];
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool k = false;
    if (i < n) {
        const uint32_t sz = parent[i] == i ? size[i] : 0u;
        k = sz > 0 && sz >= min_size;
        kept[i] = k ? 1 : 0;
        if (k && sz > max_size) atomicAdd(oversized, 1u);
    }
    const uint64_t m = __ballot(k);
    if (lane_id() == 0) wc[threadIdx.x / kWave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kCcBlock / kWave; ++w) t += wc[w];
        bcnt[blockIdx.x] = t;
    }
}

// cluster number of each kept root, written over its size entry
__global__ __launch_bounds__(kCcBlock) void k_cc_rank(const CloudHdr *__restrict__ h,
                                                     const uint8_t *__restrict__ kept,
                                                     const uint32_t *__restrict__ boff, uint32_t *__restrict__ size)
{
    __shared__ uint32_t wo[kCcBlock / kWave];
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool k = i < n && kept[i];
    const uint64_t m = __ballot(k);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) wo[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int v = 0; v < kCcBlock / kWave; ++v) {
            const uint32_t c = wo[v];
            wo[v] = t;
            t += c;
        }
    }
    __syncthreads();
    if (k) size[i] = boff[blockIdx.x] + wo[w] + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
}

__global__ __launch_bounds__(kCcBlock) void k_cc_label(const CloudHdr *__restrict__ h, const uint8_t *__restrict__ kept,
                                                      const uint32_t *__restrict__ rank, int32_t *__restrict__ label)
{
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t r = label[i];
    label[i] = kept[r] ? (int32_t)rank[r] : -1;
}

}  // namespace

int32_t launch_euclidean_clustering(gca_ctx ctx, gca_cloud c, float tolerance, uint32_t min_size,
                                    uint32_t max_size, int32_t *label, uint32_t *counts_dev)
{
    Grid G = c->grid;  // the grid may be one built for a nearby radius: search at the tolerance
    G.desc.radius = tolerance;
    const int64_t nb = std::max<int64_t>(1, (c->cap + kCcBlock - 1) / kCcBlock);
    uint32_t *size = (uint32_t *)ctx_scratch(ctx, 25, sizeof(uint32_t) * (size_t)c->cap, false);
    uint8_t *kept = (uint8_t *)ctx_scratch(ctx, 2, (size_t)c->cap, false);
    uint32_t *bcnt = (uint32_t *)ctx_scratch(ctx, 3, sizeof(uint32_t) * (nb + 1), false);
    uint32_t *boff = (uint32_t *)ctx_scratch(ctx, 4, sizeof(uint32_t) * (nb + 1), false);
    if (!size || !kept || !bcnt || !boff) return GCA_ERR_OUT_OF_MEMORY;
    const dim3 grid((unsigned)nb), block(kCcBlock);
    hipLaunchKernelGGL(k_cc_init, grid, block, 0, ctx->stream, c->hdr, label, size, counts_dev + 1);
    hipLaunchKernelGGL(k_cc_hook, grid, block, 0, ctx->stream, c->hdr, G.desc, G.start, G.sxyzi, G.order, label);
    hipLaunchKernelGGL(k_cc_compress, grid, block, 0, ctx->stream, c->hdr, label, size);
    hipLaunchKernelGGL(k_cc_flag, grid, block, 0, ctx->stream, c->hdr, label, size, min_size, max_size, kept, bcnt,
                       counts_dev + 1);
    GCA_HIP(hipGetLastError());
    GCA_TRY(exclusive_scan_u32(ctx, bcnt, boff, nb, counts_dev));
    hipLaunchKernelGGL(k_cc_rank, grid, block, 0, ctx->stream, c->hdr, kept, boff, size);
    hipLaunchKernelGGL(k_cc_label, grid, block, 0, ctx->stream, c->hdr, kept, size, label);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

}  // namespace gca
