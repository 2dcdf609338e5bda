// cluster.hip -- Euclidean clustering and residual-based segmentation as connected components
// of the radius graph.
//
// Reference: cuda_euclidean_clustering (src/geometry/cuda_clustering.cu:31-110) and
// cuda_residual_based_segmentation (:251-341) download the radius-neighbour lists and run a
// breadth-first search on the host, seeding from the lowest unvisited index.  The residual form
// follows an edge only when the two points' absolute ICP residuals differ by less than 0.003
// (:291-293).  A search stops expanding once its queue holds max_cluster_size points (:75-96,
// :308-328), leaving the rest of that component to later seeds; a queue is kept as a cluster when
// it has at least min_cluster_size points, and clusters are numbered in seed order.
//
// Here the components come from a lock-free union-find on the device:
//   1. init     parent[i] = i, size[i] = 0
//   2. hook     every point, walked in cell order for locality, links its root and the root of
//               each neighbour within the tolerance that comes before it in cell order (the
//               radius search's own predicate: squared FP64 distance < r^2, and for the residual
//               form the residual gate); a root is only ever linked to a smaller one by
//               compare-and-swap, so every tree's root is its component's lowest index, which is
//               the reference's seed
//   3. compress parent[i] = root (a read-only walk, then a store to the point's own entry),
//               size[root] += 1 (one atomic per wave and root run)
//   4. cut      only for components above max_cluster_size (listed by k_cc_over): the
//               reference's truncated searches, replayed exactly on the device (k_cc_truncate),
//               re-root the component's points on the seeds of the pieces it is cut into
//   5. rank     kept roots (size >= min) are numbered by an exclusive scan in index order
//   6. label    label[i] = number of its root's cluster, or -1
// Parent pointers only ever decrease and always name an ancestor, so the path-halving stores in
// find() may race with the compare-and-swaps without losing a link.
//
// Neighbour lists are visited in the radius search's canonical order, (squared distance,
// index), where the reference visits them in its grid-cell order; only the pieces of a cut
// component depend on that order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <vector>

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

// Edge gate of the residual-based segmentation (cuda_clustering.cu:291-293): |r_a - r_b| in FP32,
// compared with the double 0.003.  res == nullptr: plain Euclidean clustering.
__device__ __forceinline__ bool residual_gate(const float *__restrict__ res, int32_t a, int32_t b)
{
    return res == nullptr || (double)fabsf(res[a] - res[b]) < 0.003;
}

// Hook: kCcLanes lanes per point (points in cell order), each lane a share of the point's
// candidates from the 27 cells around it, kCcUnroll loads in flight per lane.  Each edge is taken
// from its later endpoint in cell order only (candidates at sorted positions before the point's
// own), which halves the candidates and changes nothing: the components, and their roots (each
// root is its component's lowest original index), do not depend on which end links an edge.
constexpr int kCcLanes = 8;
constexpr int kCcUnroll = 4;
constexpr int kCcRuns = 9;  // z-runs of cells per point when the cell is at least the radius

__device__ __forceinline__ void cc_link(int32_t *parent, int32_t &ri, int32_t j)
{
    const int32_t rj = cc_find(parent, j);
    if (rj != ri) {
        ri = cc_find(parent, ri);  // short walk: ri was a root not long ago
        if (rj != ri) {
            cc_union(parent, ri, rj);  // a compare-and-swap only for distinct roots
            ri = cc_find(parent, ri);
        }
    }
}

__global__ __launch_bounds__(kCcBlock) void k_cc_hook(const CloudHdr *__restrict__ h, GridDesc g,
                                                     const uint32_t *__restrict__ start,
                                                     const float4 *__restrict__ sxyzi,
                                                     const int32_t *__restrict__ order,
                                                     const float *__restrict__ res, int32_t *parent)
{
    const int n = h->n;
    const int gl = (int)threadIdx.x & (kCcLanes - 1);
    const int s = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kCcLanes);  // the group's point
    if (s >= n) return;  // the whole group
    const double r2 = (double)g.radius * (double)g.radius;
    const float4 p = sxyzi[s];
    const int32_t i = order[s];
    int32_t lo[3], hi[3];
    if (!grid_range(g, p.x, p.y, p.z, lo, hi)) return;
    const int nry = hi[1] - lo[1] + 1;
    const int nruns = (hi[0] - lo[0] + 1) * nry;
    int32_t ri = cc_find(parent, i);  // this point's root, refreshed after each link
    auto consider = [&](const float4 &q, int32_t j) {
        if (dist2(q.x, q.y, q.z, p.x, p.y, p.z) < r2 && residual_gate(res, i, j)) cc_link(parent, ri, j);
    };
    if (nruns <= kCcRuns) {
        // run bounds up front, each run clamped to the positions before s; one flattened walk
        uint32_t rs[kCcRuns], cum[kCcRuns + 1];
        cum[0] = 0;
#pragma unroll
        for (int r = 0; r < kCcRuns; ++r) {
            uint32_t a = 0, b = 0;
            if (r < nruns) {
                const int ix = lo[0] + r / nry, iy = lo[1] + r % nry;
                const int64_t row = ((int64_t)ix * g.dims[1] + iy) * g.dims[2];
                a = start[row + lo[2]];
                b = min(start[row + hi[2] + 1], (uint32_t)s);
            }
            rs[r] = a;
            cum[r + 1] = cum[r] + (b > a ? b - a : 0u);
        }
        const uint32_t total = cum[kCcRuns];
        for (uint32_t f0 = gl; f0 < total; f0 += kCcLanes * kCcUnroll) {
            float4 q[kCcUnroll];
            int32_t jj[kCcUnroll];
#pragma unroll
            for (int u = 0; u < kCcUnroll; ++u) {
                const uint32_t f = f0 + u * kCcLanes;
                jj[u] = -1;
                if (f < total) {
                    uint32_t t = rs[0] + f;
#pragma unroll
                    for (int r = 1; r < kCcRuns; ++r)
                        if (f >= cum[r]) t = rs[r] + (f - cum[r]);
                    q[u] = sxyzi[t];
                    jj[u] = order[t];
                }
            }
#pragma unroll
            for (int u = 0; u < kCcUnroll; ++u)
                if (jj[u] >= 0) consider(q[u], jj[u]);
        }
    } else {
        // cell smaller than the radius (not produced by grid_build): plain run walk
        for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
            for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
                const int64_t row = ((int64_t)ix * g.dims[1] + iy) * g.dims[2];
                const uint32_t e = min(start[row + hi[2] + 1], (uint32_t)s);
                for (uint32_t t = start[row + lo[2]] + gl; t < e; t += kCcLanes) consider(sxyzi[t], order[t]);
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

// Components above max_size: one job each (root, size) for the truncated searches; *n_jobs counts
// them (the jobs' order does not matter: each cut is independent of the others).
struct CutJob {
    int32_t root, size, base;  // base: the job's region of the queue buffer (set on the host)
    int32_t pad;
};

__global__ __launch_bounds__(kCcBlock) void k_cc_over(const CloudHdr *__restrict__ h, const int32_t *__restrict__ parent,
                                                     const uint32_t *__restrict__ size, uint32_t max_size,
                                                     CutJob *__restrict__ jobs, uint32_t *__restrict__ n_jobs)
{
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || parent[i] != i || size[i] <= max_size) return;
    const uint32_t k = atomicAdd(n_jobs, 1u);
    jobs[k] = CutJob{i, (int32_t)size[i], 0, 0};
}

// Claim words of the truncated searches: kUnvisited, kVisited, or the key (queue position + 1,
// list position) of the lowest proposal in the current round.
constexpr unsigned long long kUnvisited = ~0ull;
constexpr unsigned long long kVisited = 0ull;

__global__ __launch_bounds__(kCcBlock) void k_cc_claims_clear(const CloudHdr *__restrict__ h,
                                                             unsigned long long *__restrict__ claim)
{
    const int n = h->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) claim[i] = kUnvisited;
}

__device__ __forceinline__ unsigned long long claim_load(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void claim_store(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The reference's searches over one component above max_size (one workgroup per component):
// seeds in ascending index among the component's unvisited points; from each, a breadth-first
// queue that stops expanding once it holds max_size points.  The queue is grown a frontier chunk
// of up to kCutBlock elements at a time, reproducing the sequential order exactly:
//   1. each element proposes (its queue position, list position) for every unvisited, gated
//      neighbour with an atomic minimum on the neighbour's claim word;
//   2. an element's new points are the neighbours whose claim it holds, in list order (a point
//      goes to the first element, in queue order, that lists it, as in the sequential search);
//   3. element t is expanded only if the queue held fewer than max_size points before it (an
//      exclusive scan of the counts), so the chunk's expanded elements are a prefix;
//   4. claims of elements cut off (and every other non-winning proposal) are reset.
// A piece's points are then re-rooted on its seed, and size[seed] is the piece's size.  Claims
// never leave the component (its points' neighbours are its points), so workgroups are
// independent.
constexpr int kCutBlock = 1024;

__global__ __launch_bounds__(kCutBlock) void k_cc_truncate(const CloudHdr *__restrict__ h, const CutJob *__restrict__ jobs,
                                                          int32_t *__restrict__ parent, uint32_t *__restrict__ size,
                                                          const int32_t *__restrict__ off,
                                                          const int32_t *__restrict__ nbr,
                                                          const float *__restrict__ res,
                                                          unsigned long long *__restrict__ claim,
                                                          int32_t *__restrict__ queue, uint32_t max_size)
{
    __shared__ uint32_t lds[kCutBlock / kWave + 1];
    __shared__ int32_t s_seed;
    __shared__ uint32_t s_add, s_proc;
    const int n = h->n;
    const CutJob J = jobs[blockIdx.x];
    int32_t *q = queue + J.base;
    const int32_t root = J.root;
    const int t = threadIdx.x;
    int32_t cursor = root;
    for (;;) {
        // next seed: the lowest unvisited point of the component at or after the cursor
        if (t == 0) s_seed = INT_MAX;
        __syncthreads();
        for (; cursor < n; cursor += kCutBlock) {
            const int32_t i = cursor + t;
            if (i < n && parent[i] == root && claim_load(claim + i) != kVisited) atomicMin(&s_seed, i);
            __syncthreads();
            if (s_seed != INT_MAX) break;
        }
        const int32_t seed = s_seed;
        __syncthreads();
        if (seed == INT_MAX) break;
        cursor = seed + 1;
        if (t == 0) {
            q[0] = seed;
            claim_store(claim + seed, kVisited);
        }
        uint32_t qlen = 1, head = 0;
        __syncthreads();
        while (head < qlen && qlen < max_size) {
            const uint32_t level_end = qlen;
            while (head < level_end && qlen < max_size) {
                const uint32_t cnt = min((uint32_t)kCutBlock, level_end - head);
                const bool on = (uint32_t)t < cnt;
                const int32_t e = on ? q[head + t] : -1;
                const int32_t l0 = on ? off[e] : 0, l1 = on ? off[e + 1] : 0;
                const unsigned long long key_hi = (unsigned long long)(head + t + 1) << 32;
                for (int32_t l = l0; l < l1; ++l) {  // 1. proposals
                    const int32_t v = nbr[l];
                    if (residual_gate(res, e, v) && claim_load(claim + v) != kVisited)
                        atomicMin(claim + v, key_hi | (unsigned long long)(l - l0));
                }
                __syncthreads();
                uint32_t won = 0;  // 2. points this element takes
                for (int32_t l = l0; l < l1; ++l) {
                    const int32_t v = nbr[l];
                    won += residual_gate(res, e, v) && claim_load(claim + v) == (key_hi | (unsigned long long)(l - l0));
                }
                uint32_t total;
                const uint32_t before = block_excl_scan<kCutBlock>(won, lds, &total);
                (void)total;
                const bool expand = on && qlen + before < max_size;  // 3. a prefix of the chunk
                if (expand) {
                    uint32_t pos = qlen + before;
                    for (int32_t l = l0; l < l1; ++l) {
                        const int32_t v = nbr[l];
                        if (residual_gate(res, e, v) && claim_load(claim + v) == (key_hi | (unsigned long long)(l - l0))) {
                            q[pos++] = v;
                            claim_store(claim + v, kVisited);
                        }
                    }
                }
                // the last expanded element (the chunk's last, or the last before the queue is
                // full) publishes how many were expanded and how many points they added
                if (expand && (t + 1 == (int)cnt || qlen + before + won >= max_size)) {
                    s_proc = (uint32_t)t + 1;
                    s_add = before + won;
                }
                __syncthreads();
                for (int32_t l = l0; l < l1; ++l) {  // 4. reset proposals that did not win
                    const int32_t v = nbr[l];
                    if (residual_gate(res, e, v)) {
                        const unsigned long long c = claim_load(claim + v);
                        if (c != kVisited && c != kUnvisited) claim_store(claim + v, kUnvisited);
                    }
                }
                head += s_proc;
                qlen += s_add;
                __syncthreads();
            }
        }
        // the piece: every queued point (expanded or not) is re-rooted on the seed
        for (uint32_t k = t; k < qlen; k += kCutBlock) parent[q[k]] = seed;
        if (t == 0) size[seed] = qlen;
        __syncthreads();
    }
}

// kept roots: flag + per-block count for the scan
__global__ __launch_bounds__(kCcBlock) void k_cc_flag(const CloudHdr *__restrict__ h,
                                                     const int32_t *__restrict__ parent,
                                                     const uint32_t *__restrict__ size, uint32_t min_size,
                                                     uint8_t *__restrict__ kept, uint32_t *__restrict__ bcnt)
{
    __shared__ uint32_t wc[kCcBlock / kWave];
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool k = false;
    if (i < n) {
        const uint32_t sz = parent[i] == i ? size[i] : 0u;
        k = sz > 0 && sz >= min_size;
        kept[i] = k ? 1 : 0;
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
                                    uint32_t max_size, int32_t *label, uint32_t *counts_dev, const float *res)
{
    Grid G = c->grid;  // the grid may be one built for a nearby radius: search at the tolerance
    G.desc.radius = tolerance;
    const int64_t nb = std::max<int64_t>(1, (c->cap + kCcBlock - 1) / kCcBlock);
    uint32_t *size = (uint32_t *)ctx_scratch(ctx, 25, sizeof(uint32_t) * (size_t)c->cap, false);
    uint8_t *kept = (uint8_t *)ctx_scratch(ctx, 2, (size_t)c->cap, false);
    uint32_t *bcnt = (uint32_t *)ctx_scratch(ctx, 3, sizeof(uint32_t) * (nb + 1), false);
    uint32_t *boff = (uint32_t *)ctx_scratch(ctx, 4, sizeof(uint32_t) * (nb + 1), false);
    CutJob *jobs = (CutJob *)ctx_scratch(ctx, 28, sizeof(CutJob) * (size_t)std::max<int64_t>(c->cap, 1), false);
    if (!size || !kept || !bcnt || !boff || !jobs) return GCA_ERR_OUT_OF_MEMORY;
    const dim3 grid((unsigned)nb), block(kCcBlock);
    hipLaunchKernelGGL(k_cc_init, grid, block, 0, ctx->stream, c->hdr, label, size, counts_dev + 1);
    const dim3 grid_hook((unsigned)std::max<int64_t>(1, (c->cap * kCcLanes + kCcBlock - 1) / kCcBlock));
    hipLaunchKernelGGL(k_cc_hook, grid_hook, block, 0, ctx->stream, c->hdr, G.desc, G.start, G.sxyzi, G.order, res,
                       label);
    hipLaunchKernelGGL(k_cc_compress, grid, block, 0, ctx->stream, c->hdr, label, size);
    hipLaunchKernelGGL(k_cc_over, grid, block, 0, ctx->stream, c->hdr, label, size, max_size, jobs, counts_dev + 1);
    GCA_HIP(hipGetLastError());
    uint32_t n_jobs = 0;
    GCA_HIP(hipMemcpyAsync(&n_jobs, counts_dev + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    if (n_jobs > 0) {
        // Components above max_size are cut as the reference's searches cut them.  Their
        // neighbour lists (radius search of the cloud against itself, canonical order) and a
        // queue region per component, laid out on the host.
        std::vector<CutJob> hj(n_jobs);
        GCA_HIP(hipMemcpyAsync(hj.data(), jobs, sizeof(CutJob) * n_jobs, hipMemcpyDeviceToHost, ctx->stream));
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        int64_t base = 0;
        for (auto &j : hj) {
            j.base = (int32_t)base;
            base += j.size;
        }
        GCA_HIP(hipMemcpyAsync(jobs, hj.data(), sizeof(CutJob) * n_jobs, hipMemcpyHostToDevice, ctx->stream));
        const int64_t n = c->host_hdr.n;
        int32_t *off = (int32_t *)ctx_scratch(ctx, 29, sizeof(int32_t) * (size_t)(n + 1), false);
        int32_t *queue = (int32_t *)ctx_scratch(ctx, 30, sizeof(int32_t) * (size_t)std::max<int64_t>(base, 1), false);
        unsigned long long *claim =
            (unsigned long long *)ctx_scratch(ctx, 31, sizeof(unsigned long long) * (size_t)std::max<int64_t>(n, 1), false);
        if (!off || !queue || !claim) return GCA_ERR_OUT_OF_MEMORY;
        int64_t total = 0;
        GCA_TRY(launch_radius(ctx, c, c, tolerance, 0, off, nullptr, 0, &total));
        int32_t *nbr = (int32_t *)ctx_scratch(ctx, 32, sizeof(int32_t) * (size_t)std::max<int64_t>(total, 1), false);
        if (!nbr) return GCA_ERR_OUT_OF_MEMORY;
        GCA_TRY(launch_radius(ctx, c, c, tolerance, 0, off, nbr, total, &total));
        hipLaunchKernelGGL(k_cc_claims_clear, grid, block, 0, ctx->stream, c->hdr, claim);
        hipLaunchKernelGGL(k_cc_truncate, dim3(n_jobs), dim3(kCutBlock), 0, ctx->stream, c->hdr, jobs, label, size, off, nbr,
                           res, claim, queue, max_size);
        GCA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_cc_flag, grid, block, 0, ctx->stream, c->hdr, label, size, min_size, kept, bcnt);
    GCA_HIP(hipGetLastError());
    GCA_TRY(exclusive_scan_u32(ctx, bcnt, boff, nb, counts_dev));
    hipLaunchKernelGGL(k_cc_rank, grid, block, 0, ctx->stream, c->hdr, kept, boff, size);
    hipLaunchKernelGGL(k_cc_label, grid, block, 0, ctx->stream, c->hdr, kept, size, label);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

}  // namespace gca
