// grid.hip -- uniform-grid neighbour search (K4/K5/K6) and radius outlier removal.
//
// Replaces the grid build + nn_search_functor + nn_search_radius_step1/2 +
// check_if_enough_radius_nn_functor of src/geometry/cuda_nn_search.cu:29-800.
//
// Grid build is a counting sort: count points per dense cell (atomics), exclusive scan of
// the counts, scatter (atomicSub hands out slots and leaves the counts at zero again).  The
// points are copied into cell order (sxyzi) so that neighbour scans read contiguous float4s:
// cells (ix, iy, lo_z..hi_z) are adjacent in the z-fastest linear order, so a query reads at
// most 9 contiguous runs.  The grid is built once per cloud and radius and cached, where the
// reference rebuilds it (including a sort) on every call.
//
// Results are canonical: neighbour set = { j : d2_fp64 < r^2 }, nearest = argmin (d2, index),
// K-cap = the K smallest (d2, index); so they do not depend on the order inside a cell.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

__global__ __launch_bounds__(256) void k_grid_count(const float4 *__restrict__ xyzi,
                                                    const CloudHdr *__restrict__ hdr, GridDesc g,
                                                    uint32_t *__restrict__ count,
                                                    uint32_t *__restrict__ cell_of)
{
    const int n = hdr->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = xyzi[i];
        int32_t cx = grid_cell(p.x, g.origin[0], g.cell, g.dims[0]);
        int32_t cy = grid_cell(p.y, g.origin[1], g.cell, g.dims[1]);
        int32_t cz = grid_cell(p.z, g.origin[2], g.cell, g.dims[2]);
        uint32_t lin = (uint32_t)(((int64_t)cx * g.dims[1] + cy) * g.dims[2] + cz);
        cell_of[i] = lin;
        atomicAdd(&count[lin], 1u);
    }
}

struct ScatterArgs {
    const float4 *xyzi;
    const CloudHdr *hdr;
    uint32_t *count;
    const uint32_t *start, *cell_of;
    int32_t *order;
    float4 *sxyzi;
};

template <int N>
__global__ __launch_bounds__(256) void k_grid_scatter(Jobs<ScatterArgs, N> J)
{
    const ScatterArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    const float4 *__restrict__ xyzi = A.xyzi;
    uint32_t *__restrict__ count = A.count;
    const uint32_t *__restrict__ start = A.start;
    const uint32_t *__restrict__ cell_of = A.cell_of;
    int32_t *__restrict__ order = A.order;
    float4 *__restrict__ sxyzi = A.sxyzi;
    const int n = A.hdr->n;
    const int lane = lane_id();
    for (int b = blockIdx.x * blockDim.x; b < n; b += gridDim.x * blockDim.x) {  // block-uniform
        const int i = b + (int)threadIdx.x;
        const bool on = i < n;
        const uint32_t lin = on ? cell_of[i] : 0u;
        // a run of equal cells in the wavefront takes its slots with one atomic (the order of
        // points within a cell is free: every query breaks ties by original index)
        const int head = run_head(on ? (int64_t)lin : -1 - (int64_t)lane);
        const uint32_t len = run_reduce(on ? 1u : 0u, head, [](uint32_t x, uint32_t y) { return x + y; });
        uint32_t top = 0u;
        if (on && head == lane) top = atomicSub(&count[lin], len);
        top = __shfl(top, head, kWave);
        if (on) {
            const uint32_t pos = start[lin] + top - 1u - (uint32_t)(lane - head);
            order[pos] = i;
            sxyzi[pos] = xyzi[i];
        }
    }
}

int32_t grid_prepare(gca_ctx ctx, gca_cloud c, const GridDesc &d)
{
    Grid &G = c->grid;
    G.valid = false;
    if (G.cap_cells < d.ncells) {
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        hipFree(G.count);
        hipFree(G.start);
        G.count = G.start = nullptr;
        G.cap_cells = 0;
        int64_t want = std::min<int64_t>(kGridCellCap, d.ncells + d.ncells / 4 + 1024);
        GCA_HIP(hipMalloc(&G.count, sizeof(uint32_t) * want));
        GCA_HIP(hipMalloc(&G.start, sizeof(uint32_t) * (want + 1)));
        GCA_HIP(hipMemsetAsync(G.count, 0, sizeof(uint32_t) * want, ctx->stream));
        G.cap_cells = want;
    }
    if (G.cap_pts < c->cap) {
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        hipFree(G.order);
        hipFree(G.sxyzi);
        hipFree(G.cell_of);
        hipFree(G.snrm);
        hipFree(G.sgrad);
        G.order = nullptr;
        G.sxyzi = nullptr;
        G.cell_of = nullptr;
        G.snrm = G.sgrad = nullptr;
        G.cap_pts = 0;
        GCA_HIP(hipMalloc(&G.order, sizeof(int32_t) * c->cap));
        GCA_HIP(hipMalloc(&G.sxyzi, sizeof(float4) * c->cap));
        GCA_HIP(hipMalloc(&G.cell_of, sizeof(uint32_t) * c->cap));
        G.cap_pts = c->cap;
    }
    return GCA_OK;
}

int32_t launch_grid_build(gca_ctx ctx, gca_cloud c, const GridDesc &d, bool counted)
{
    Grid &G = c->grid;
    GCA_TRY(grid_prepare(ctx, c, d));
    const unsigned nb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (c->cap + 255) / 256));
    StageTimer timer(ctx, GCA_STAGE_GRID);
    if (!counted) {
        hipLaunchKernelGGL(k_grid_count, dim3(nb), dim3(256), 0, ctx->stream, c->xyzi, c->hdr, d, G.count,
                           G.cell_of);
        GCA_HIP(hipGetLastError());
    }
    GCA_TRY(exclusive_scan_u32(ctx, G.count, G.start, d.ncells, nullptr));
    const Jobs<ScatterArgs, 1> sc{{{c->xyzi, c->hdr, G.count, G.start, G.cell_of, G.order, G.sxyzi}}};
    hipLaunchKernelGGL(k_grid_scatter<1>, dim3(nb), dim3(256), 0, ctx->stream, sc);
    GCA_HIP(hipGetLastError());
    G.desc = d;
    G.valid = true;
    G.attrs_valid = false;
    return GCA_OK;
}

int32_t launch_grid_levels(gca_ctx launch, const gca_ctx *ctxs, const gca_cloud *cs, const GridDesc *ds, int nj)
{
    if (nj < 1 || nj > kMaxJobs) return GCA_ERR_INVALID_VALUE;
    bool wide = false;
    for (int j = 0; j < nj; ++j) {
        GCA_TRY(grid_prepare(ctxs[j], cs[j], ds[j]));
        wide = wide || scan_wide(ds[j].ncells);
    }
    ScanJob scans[kMaxJobs];
    Jobs<ScatterArgs, kMaxJobs> sc{};
    unsigned nb = 1;
    for (int j = 0; j < nj; ++j) {
        const Grid &G = cs[j]->grid;
        GCA_TRY(scan_job(ctxs[j], G.count, G.start, ds[j].ncells, nullptr, wide, &scans[j]));
        sc.a[j] = ScatterArgs{cs[j]->xyzi, cs[j]->hdr, G.count, G.start, G.cell_of, G.order, G.sxyzi};
        nb = std::max(nb, (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (cs[j]->cap + 255) / 256)));
    }
    {
        StageTimer timer(launch, GCA_STAGE_GRID);
        GCA_TRY(scan_launch(launch->stream, scans, nj, false, wide));
        hipLaunchKernelGGL(k_grid_scatter<kMaxJobs>, dim3(nb, (unsigned)nj), dim3(256), 0, launch->stream, sc);
        GCA_HIP(hipGetLastError());
    }
    for (int j = 0; j < nj; ++j) {
        Grid &G = cs[j]->grid;
        G.desc = ds[j];
        G.valid = true;
        G.attrs_valid = false;
    }
    return GCA_OK;
}

__global__ __launch_bounds__(256) void k_gather_sorted(const CloudHdr *__restrict__ hdr,
                                                       const int32_t *__restrict__ order,
                                                       const float4 *__restrict__ nrm,
                                                       const float4 *__restrict__ grad,
                                                       float4 *__restrict__ snrm, float4 *__restrict__ sgrad)
{
    const int n = hdr->n;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const int32_t j = order[s];
        snrm[s] = nrm[j];
        sgrad[s] = grad[j];
    }
}

int32_t grid_gather_attributes(gca_ctx ctx, gca_cloud c)
{
    Grid &G = c->grid;
    if (!G.valid) return GCA_ERR_NOT_READY;
    if (G.attrs_valid) return GCA_OK;
    if (!G.snrm) {
        GCA_HIP(hipMalloc(&G.snrm, sizeof(float4) * G.cap_pts));
        GCA_HIP(hipMalloc(&G.sgrad, sizeof(float4) * G.cap_pts));
    }
    const unsigned nb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (c->cap + 255) / 256));
    hipLaunchKernelGGL(k_gather_sorted, dim3(nb), dim3(256), 0, ctx->stream, c->hdr, G.order, c->nrm, c->grad,
                       G.snrm, G.sgrad);
    GCA_HIP(hipGetLastError());
    G.attrs_valid = true;
    return GCA_OK;
}

// ---- queries ---------------------------------------------------------------------------------

struct GridView {
    GridDesc g;
    const uint32_t *start;
    const int32_t *order;
    const float4 *sxyzi;
};

__device__ __forceinline__ GridView make_view(const GridDesc &g, const uint32_t *start, const int32_t *order,
                                              const float4 *sxyzi)
{
    GridView v;
    v.g = g;
    v.start = start;
    v.order = order;
    v.sxyzi = sxyzi;
    return v;
}

// Nearest neighbour of q within g.radius: sorted slot of argmin (d2, index) or -1.
__device__ __forceinline__ int32_t nn_query(const GridView &v, float qx, float qy, float qz, double *best_d2)
{
    int32_t lo[3], hi[3];
    *best_d2 = DBL_MAX;
    if (!grid_range(v.g, qx, qy, qz, lo, hi)) return -1;
    const double r2 = (double)v.g.radius * (double)v.g.radius;
    double bd = DBL_MAX;
    int32_t bi = INT_MAX, bs = -1;
    for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
        for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
            const int64_t row = ((int64_t)ix * v.g.dims[1] + iy) * v.g.dims[2];
            const uint32_t s0 = v.start[row + lo[2]], s1 = v.start[row + hi[2] + 1];
            for (uint32_t s = s0; s < s1; ++s) {
                const float4 p = v.sxyzi[s];
                const double d2 = dist2(p.x, p.y, p.z, qx, qy, qz);
                if (d2 < r2 && d2 <= bd) {
                    const int32_t idx = v.order[s];
                    if (key_less(d2, idx, bd, bi)) {
                        bd = d2;
                        bi = idx;
                        bs = (int32_t)s;
                    }
                }
            }
        }
    *best_d2 = bd;
    return bs;
}

__global__ __launch_bounds__(256) void k_nn(const float4 *__restrict__ q, const CloudHdr *__restrict__ qh,
                                            GridDesc g, const uint32_t *__restrict__ start,
                                            const int32_t *__restrict__ order, const float4 *__restrict__ sxyzi,
                                            int32_t *__restrict__ out)
{
    const GridView v = make_view(g, start, order, sxyzi);
    const int n = qh->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = q[i];
        double d2;
        int32_t s = nn_query(v, p.x, p.y, p.z, &d2);
        out[i] = s >= 0 ? order[s] : -1;
    }
}

int32_t launch_nn(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t *out)
{
    const Grid &G = ref->grid;
    const unsigned nb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (q->cap + 255) / 256));
    hipLaunchKernelGGL(k_nn, dim3(nb), dim3(256), 0, ctx->stream, q->xyzi, q->hdr, query_desc(G, radius), G.start, G.order,
                       G.sxyzi, out);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// ---- radius search ------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_radius_count(const float4 *__restrict__ q,
                                                      const CloudHdr *__restrict__ qh, GridDesc g,
                                                      const uint32_t *__restrict__ start,
                                                      const float4 *__restrict__ sxyzi,
                                                      uint32_t *__restrict__ cnt)
{
    const int n = qh->n;
    const double r2 = (double)g.radius * (double)g.radius;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = q[i];
        int32_t lo[3], hi[3];
        uint32_t k = 0;
        if (grid_range(g, p.x, p.y, p.z, lo, hi))
            for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
                for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
                    const int64_t row = ((int64_t)ix * g.dims[1] + iy) * g.dims[2];
                    for (uint32_t s = start[row + lo[2]]; s < start[row + hi[2] + 1]; ++s) {
                        float4 c = sxyzi[s];
                        k += dist2(c.x, c.y, c.z, p.x, p.y, p.z) < r2 ? 1u : 0u;
                    }
                }
        cnt[i] = k;
    }
}

// Fills each query's segment sorted by (d2, index) by insertion.
__global__ __launch_bounds__(256) void k_radius_fill(const float4 *__restrict__ q,
                                                     const CloudHdr *__restrict__ qh, GridDesc g,
                                                     const uint32_t *__restrict__ start,
                                                     const int32_t *__restrict__ order,
                                                     const float4 *__restrict__ sxyzi,
                                                     const uint32_t *__restrict__ off,
                                                     double *__restrict__ td2, int32_t *__restrict__ tidx)
{
    const int n = qh->n;
    const double r2 = (double)g.radius * (double)g.radius;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = q[i];
        int32_t lo[3], hi[3];
        const uint32_t base = off[i];
        uint32_t k = 0;
        if (grid_range(g, p.x, p.y, p.z, lo, hi))
            for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
                for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
                    const int64_t row = ((int64_t)ix * g.dims[1] + iy) * g.dims[2];
                    for (uint32_t s = start[row + lo[2]]; s < start[row + hi[2] + 1]; ++s) {
                        float4 c = sxyzi[s];
                        double d2 = dist2(c.x, c.y, c.z, p.x, p.y, p.z);
                        if (!(d2 < r2)) continue;
                        int32_t idx = order[s];
                        uint32_t j = k;
                        while (j > 0 && key_less(d2, idx, td2[base + j - 1], tidx[base + j - 1])) {
                            td2[base + j] = td2[base + j - 1];
                            tidx[base + j] = tidx[base + j - 1];
                            --j;
                        }
                        td2[base + j] = d2;
                        tidx[base + j] = idx;
                        ++k;
                    }
                }
    }
}

__global__ __launch_bounds__(256) void k_radius_cap(const uint32_t *__restrict__ cnt,
                                                    const CloudHdr *__restrict__ qh, uint32_t cap,
                                                    uint32_t *__restrict__ capped)
{
    const int n = qh->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        capped[i] = cap ? min(cnt[i], cap) : cnt[i];
}

__global__ __launch_bounds__(256) void k_radius_copy(const CloudHdr *__restrict__ qh,
                                                     const uint32_t *__restrict__ off_unc,
                                                     const int32_t *__restrict__ tidx,
                                                     const uint32_t *__restrict__ off_cap,
                                                     int32_t *__restrict__ out)
{
    const int n = qh->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t a = off_unc[i], b = off_cap[i], m = off_cap[i + 1] - off_cap[i];
        for (uint32_t k = 0; k < m; ++k) out[b + k] = tidx[a + k];
    }
}

int32_t launch_radius(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t max_nn, int32_t *offsets,
                      int32_t *idx, int64_t idx_cap, int64_t *total)
{
    GCA_TRY(cloud_sync_header(q));
    const int64_t nq = q->host_hdr.n;
    Grid G = ref->grid;
    G.desc.radius = radius;
    const unsigned nb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (nq + 255) / 256));
    uint32_t *cnt = (uint32_t *)ctx_scratch(ctx, 8, sizeof(uint32_t) * (nq + 1), false);
    uint32_t *off = (uint32_t *)ctx_scratch(ctx, 9, sizeof(uint32_t) * (nq + 1), false);
    uint32_t *tot_dev = (uint32_t *)ctx_scratch(ctx, 10, sizeof(uint32_t) * 4, false);
    if (!cnt || !off || !tot_dev) return GCA_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL(k_radius_count, dim3(nb), dim3(256), 0, ctx->stream, q->xyzi, q->hdr, G.desc, G.start,
                       G.sxyzi, cnt);
    GCA_HIP(hipGetLastError());
    GCA_TRY(exclusive_scan_u32(ctx, cnt, off, nq, tot_dev));
    uint32_t unc = 0;
    GCA_HIP(hipMemcpyAsync(&unc, tot_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    double *td2 = (double *)ctx_scratch(ctx, 11, sizeof(double) * (unc + 1), false);
    int32_t *tidx = (int32_t *)ctx_scratch(ctx, 12, sizeof(int32_t) * (unc + 1), false);
    if (!td2 || !tidx) return GCA_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL(k_radius_fill, dim3(nb), dim3(256), 0, ctx->stream, q->xyzi, q->hdr, G.desc, G.start,
                       G.order, G.sxyzi, off, td2, tidx);
    hipLaunchKernelGGL(k_radius_cap, dim3(nb), dim3(256), 0, ctx->stream, cnt, q->hdr, (uint32_t)max_nn, cnt);
    GCA_HIP(hipGetLastError());
    GCA_TRY(exclusive_scan_u32(ctx, cnt, (uint32_t *)offsets, nq, tot_dev));
    uint32_t capped = 0;
    GCA_HIP(hipMemcpyAsync(&capped, tot_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    *total = capped;
    if (!idx) return GCA_OK;  // count-only query
    if ((int64_t)capped > idx_cap) return GCA_ERR_INVALID_VALUE;
    hipLaunchKernelGGL(k_radius_copy, dim3(nb), dim3(256), 0, ctx->stream, q->hdr, off, tidx,
                       (const uint32_t *)offsets, idx);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// ---- cell binning (exposed for parity checks) --------------------------------------------

__global__ __launch_bounds__(256) void k_cells(const float4 *__restrict__ xyzi, const CloudHdr *__restrict__ h,
                                               GridDesc g, int32_t *__restrict__ cells)
{
    const int n = h->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = xyzi[i];
        cells[3 * i] = grid_cell(p.x, g.origin[0], g.cell, g.dims[0]);
        cells[3 * i + 1] = grid_cell(p.y, g.origin[1], g.cell, g.dims[1]);
        cells[3 * i + 2] = grid_cell(p.z, g.origin[2], g.cell, g.dims[2]);
    }
}

int32_t launch_grid_cells(gca_ctx ctx, gca_cloud c, int32_t *cells)
{
    const unsigned nb = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, (c->cap + 255) / 256));
    hipLaunchKernelGGL(k_cells, dim3(nb), dim3(256), 0, ctx->stream, c->xyzi, c->hdr, c->grid.desc, cells);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// ---- radius outlier removal (cuda_nn_search.cu:416-542,725-800) ---------------------------
// keep iff (# points within radius, itself included) > min_nn; stable compaction.

__global__ __launch_bounds__(256) void k_outlier_flag(const float4 *__restrict__ xyzi,
                                                      const CloudHdr *__restrict__ h, GridDesc g,
                                                      const uint32_t *__restrict__ start,
                                                      const float4 *__restrict__ sxyzi, int32_t min_nn,
                                                      uint8_t *__restrict__ keep, uint32_t *__restrict__ bcnt)
{
    __shared__ uint32_t wc[4];
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double r2 = (double)g.radius * (double)g.radius;
    bool k = false;
    if (i < n) {
        float4 p = xyzi[i];
        int32_t lo[3], hi[3];
        int64_t c = 0;
        if (grid_range(g, p.x, p.y, p.z, lo, hi))
            for (int32_t ix = lo[0]; ix <= hi[0] && c <= min_nn; ++ix)
                for (int32_t iy = lo[1]; iy <= hi[1] && c <= min_nn; ++iy) {
                    const int64_t row = ((int64_t)ix * g.dims[1] + iy) * g.dims[2];
                    for (uint32_t s = start[row + lo[2]]; s < start[row + hi[2] + 1] && c <= min_nn; ++s) {
                        float4 q = sxyzi[s];
                        c += dist2(q.x, q.y, q.z, p.x, p.y, p.z) < r2 ? 1 : 0;
                    }
                }
        k = c > (int64_t)min_nn;
        keep[i] = k ? 1 : 0;
    }
    uint64_t m = __ballot(k);
    if (lane_id() == 0) wc[threadIdx.x / kWave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(256) void k_outlier_compact(const float4 *__restrict__ xyzi,
                                                         const float4 *__restrict__ rgb,
                                                         const CloudHdr *__restrict__ h,
                                                         const uint8_t *__restrict__ keep,
                                                         const uint32_t *__restrict__ boff,
                                                         float4 *__restrict__ oxyzi, float4 *__restrict__ orgb)
{
    __shared__ uint32_t wo[4];
    const int n = h->n;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool k = i < n && keep[i];
    uint64_t m = __ballot(k);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) wo[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = wo[0], b = wo[1], c = wo[2];
        wo[0] = 0;
        wo[1] = a;
        wo[2] = a + b;
        wo[3] = a + b + c;
    }
    __syncthreads();
    if (k) {
        uint32_t dst = boff[blockIdx.x] + wo[w] + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
        oxyzi[dst] = xyzi[i];
        orgb[dst] = rgb[i];
    }
}

int32_t launch_outlier(gca_ctx ctx, gca_cloud in, float radius, int32_t min_nn, gca_cloud out)
{
    GCA_TRY(cloud_reserve(out, in->cap));
    Grid G = in->grid;
    G.desc.radius = radius;
    const int64_t nblk = std::max<int64_t>(1, (in->cap + 255) / 256);
    uint8_t *keep = (uint8_t *)ctx_scratch(ctx, 2, (size_t)in->cap, false);
    uint32_t *bcnt = (uint32_t *)ctx_scratch(ctx, 3, sizeof(uint32_t) * (nblk + 1), false);
    uint32_t *boff = (uint32_t *)ctx_scratch(ctx, 4, sizeof(uint32_t) * (nblk + 1), false);
    if (!keep || !bcnt || !boff) return GCA_ERR_OUT_OF_MEMORY;
    hipLaunchKernelGGL(k_outlier_flag, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, in->xyzi, in->hdr, G.desc,
                       G.start, G.sxyzi, min_nn, keep, bcnt);
    GCA_HIP(hipGetLastError());
    GCA_TRY(exclusive_scan_u32(ctx, bcnt, boff, nblk, (uint32_t *)&out->hdr->n));
    hipLaunchKernelGGL(k_outlier_compact, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, in->xyzi, in->rgb,
                       in->hdr, keep, boff, out->xyzi, out->rgb);
    GCA_HIP(hipGetLastError());
    GCA_TRY(launch_bounds(ctx, out));
    return GCA_OK;
}

}  // namespace gca
