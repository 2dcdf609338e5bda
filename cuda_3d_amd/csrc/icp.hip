// icp.hip -- colored ICP on the device (K5 + K9 + K10).
//
// Replaces, per iteration of color_icp::align (src/registration/color_icp.cu:151-234):
//   point_cloud::nn_search (grid rebuild + sort + nn_search_functor, cuda_nn_search.cu:545-619)
//   cuda_build_gauss_newton_color_icp (thrust::transform_reduce of 172-byte tuples,
//       cuda_color_icp_build_least_square.cu:134-266,431-474) + device->host copy
//   solve_JTJ_JTr<true> (host Eigen LDLT, eigen_solver.hpp:44-85)
//   point_cloud::transform (in-place FP32 transform + bounds reduce + device->host copy)
// with ONE kernel launch per iteration and no host round trip:
//   - every thread maps its source points by the current pose (FP64 -> FP32), finds the
//     nearest target point in the cached target grid, and forms the residuals and 6-D
//     Jacobians in FP32 (reference formulas);
//   - the 21 + 6 + 1 sums are accumulated per thread with TwoSum-compensated FP32, reduced
//     across the wavefront and the block with compensated merges, and written as FP64 block
//     partials;
//   - the last block to finish (atomic ticket) sums the partials in a fixed order in FP64,
//     solves the 6x6 system by LDL^T in FP64, applies the Euler increment to the FP64 pose,
//     evaluates the stopping rule and re-arms the ticket.  Later launches see `done` and exit.
// The same kernel runs many (source, target) pairs at once: a block -> pair table gives every
// pair its own range of blocks, its own partials and its own ticket.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

constexpr int kIcpBlock = 256;
constexpr int kNSys = 29;        // 21 JTJ + 6 JTr + r^2 + n_corr
constexpr int kPartStride = 32;  // doubles per block partial

constexpr int kMaxClusters = 64;  // block partials are reduced in clusters, then clusters
constexpr int kTicketStride = 64; // uint32 per ticket: one 256-byte line each, so the atomics
                                  // of different clusters do not serialise on a shared line

struct IcpState {
    double T[16];
    double sys[kNSys];
    double conv_rot, conv_t2;
    int32_t iter, max_iter, done, converged, solver_failures, pad0;
    alignas(256) uint32_t ticket;
    alignas(256) uint32_t cl_ticket[kMaxClusters * kTicketStride];
};

struct PairDesc {
    const float4 *src;
    const CloudHdr *src_hdr;
    GridDesc g;
    const uint32_t *start;
    const int32_t *order;
    const float4 *sxyzi;
    const float4 *nrm;   // target normals / gradients by original index
    const float4 *grad;
    const float4 *snrm;  // ... and in grid (slot) order
    const float4 *sgrad;
    IcpState *st;
    double *partials;   // nblk block partials, then nclust cluster partials (kPartStride each)
    int32_t nblk;
    int32_t first_block;
    int32_t cluster;    // blocks per cluster
    int32_t nclust;
};

// ---- per-correspondence residuals and Jacobians (FP32, reference formulas) ------------------
struct Terms {
    float Jg[6], Jc[6], rg, rc;
};

__device__ __forceinline__ void corr_terms(float px, float py, float pz, float Ip, float4 q, float4 n, float4 d,
                                           float sl, float sc, Terms &t)
{
    const float dqx = px - q.x, dqy = py - q.y, dqz = pz - q.z;
    const float rg_ = (dqx * n.x + dqy * n.y) + dqz * n.z;  // .cu:193
    t.Jg[0] = sl * (py * n.z - pz * n.y);                   // sqrt(l) p x n  (.cu:197)
    t.Jg[1] = sl * (pz * n.x - px * n.z);
    t.Jg[2] = sl * (px * n.y - py * n.x);
    t.Jg[3] = sl * n.x;
    t.Jg[4] = sl * n.y;
    t.Jg[5] = sl * n.z;
    const float dn = (d.x * n.x + d.y * n.y) + d.z * n.z;   // d^T (I - n n^T)  (.cu:214-220)
    const float gx = d.x - dn * n.x, gy = d.y - dn * n.y, gz = d.z - dn * n.z;
    t.Jc[0] = sc * (py * gz - pz * gy);                     // .cu:240-243
    t.Jc[1] = sc * (pz * gx - px * gz);
    t.Jc[2] = sc * (px * gy - py * gx);
    t.Jc[3] = sc * gx;
    t.Jc[4] = sc * gy;
    t.Jc[5] = sc * gz;
    const float prx = px - rg_ * n.x - q.x, pry = py - rg_ * n.y - q.y, prz = pz - rg_ * n.z - q.z;
    const float Iproj = q.w + ((d.x * prx + d.y * pry) + d.z * prz);  // .cu:248-251
    t.rc = sc * (Iproj - Ip);
    t.rg = sl * rg_;
}

// ---- accumulators ---------------------------------------------------------------------------
template <int MODE>
struct Acc;

template <>
struct Acc<GCA_ACCUM_COMPENSATED> {
    float s[28], c[28];
    __device__ void zero()
    {
#pragma unroll
        for (int k = 0; k < 28; ++k) s[k] = c[k] = 0.f;
    }
    __device__ __forceinline__ void add(int k, float x)
    {
        const float t = s[k] + x;
        const float bp = t - s[k];
        c[k] += (s[k] - (t - bp)) + (x - bp);
        s[k] = t;
    }
    __device__ __forceinline__ void merge_from_lane(int mask)
    {
#pragma unroll
        for (int k = 0; k < 28; ++k) {
            const float os = __shfl_xor(s[k], mask, kWave), oc = __shfl_xor(c[k], mask, kWave);
            const float t = s[k] + os;
            const float bp = t - s[k];
            const float e = (s[k] - (t - bp)) + (os - bp);
            s[k] = t;
            c[k] = (c[k] + oc) + e;
        }
    }
    __device__ __forceinline__ double value(int k) const { return (double)s[k] + (double)c[k]; }
};

template <>
struct Acc<GCA_ACCUM_NAIVE_FP32> {
    float s[28];
    __device__ void zero()
    {
#pragma unroll
        for (int k = 0; k < 28; ++k) s[k] = 0.f;
    }
    __device__ __forceinline__ void add(int k, float x) { s[k] += x; }
    __device__ __forceinline__ void merge_from_lane(int mask)
    {
#pragma unroll
        for (int k = 0; k < 28; ++k) s[k] += __shfl_xor(s[k], mask, kWave);
    }
    __device__ __forceinline__ double value(int k) const { return (double)s[k]; }
};

template <>
struct Acc<GCA_ACCUM_FP64> {
    double s[28];
    __device__ void zero()
    {
#pragma unroll
        for (int k = 0; k < 28; ++k) s[k] = 0.0;
    }
    __device__ __forceinline__ void add(int k, float x) { s[k] += (double)x; }
    __device__ __forceinline__ void merge_from_lane(int mask)
    {
#pragma unroll
        for (int k = 0; k < 28; ++k) s[k] += __shfl_xor(s[k], mask, kWave);
    }
    __device__ __forceinline__ double value(int k) const { return s[k]; }
};

template <int MODE>
__device__ __forceinline__ void acc_terms(Acc<MODE> &A, const Terms &t)
{
    int k = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = a; b < 6; ++b) A.add(k++, t.Jg[a] * t.Jg[b] + t.Jc[a] * t.Jc[b]);
#pragma unroll
    for (int a = 0; a < 6; ++a) A.add(21 + a, t.Jg[a] * t.rg + t.Jc[a] * t.rc);
    A.add(27, t.rg * t.rg + t.rc * t.rc);
}

__device__ __forceinline__ void xform_f32(const double *T, float4 p, float &x, float &y, float &z)
{
    const double a = p.x, b = p.y, c = p.z;
    x = (float)(((T[0] * a + T[1] * b) + T[2] * c) + T[3]);
    y = (float)(((T[4] * a + T[5] * b) + T[6] * c) + T[7]);
    z = (float)(((T[8] * a + T[9] * b) + T[10] * c) + T[11]);
}

// NN of (qx,qy,qz): sorted slot of argmin (d2, index) within g.radius, or -1.
__device__ __forceinline__ int32_t nn_slot(const PairDesc &P, float qx, float qy, float qz)
{
    int32_t lo[3], hi[3];
    if (!grid_range(P.g, qx, qy, qz, lo, hi)) return -1;
    const double r2 = (double)P.g.radius * (double)P.g.radius;
    double bd = DBL_MAX;
    int32_t bi = INT_MAX, bs = -1;
    for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
        for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
            const int64_t row = ((int64_t)ix * P.g.dims[1] + iy) * P.g.dims[2];
            const uint32_t s0 = P.start[row + lo[2]], s1 = P.start[row + hi[2] + 1];
            for (uint32_t s = s0; s < s1; ++s) {
                const float4 c = P.sxyzi[s];
                const double d2 = dist2(c.x, c.y, c.z, qx, qy, qz);
                if (d2 < r2 && d2 <= bd) {
                    const int32_t idx = P.order[s];
                    if (key_less(d2, idx, bd, bi)) {
                        bd = d2;
                        bi = idx;
                        bs = (int32_t)s;
                    }
                }
            }
        }
    return bs;
}

// Block reduction of the accumulator into FP64 partial `out[0..27]`, plus the correspondence
// count into out[28].  Wave: butterfly merges; block: the 4 wave results merged in wave order.
// LEAD: only lanes that are multiples of LEAD hold data (group leaders), so the butterfly runs
// over masks LEAD .. 32 only.
template <int MODE, int LEAD = 1>
__device__ __forceinline__ void block_partial(Acc<MODE> &A, int ncorr, double *out)
{
    __shared__ Acc<MODE> wacc[kIcpBlock / kWave];
    __shared__ int wcnt[kIcpBlock / kWave];
#pragma unroll
    for (int m = kWave / 2; m >= LEAD; m >>= 1) A.merge_from_lane(m);
#pragma unroll
    for (int m = kWave / 2; m >= LEAD; m >>= 1) ncorr += __shfl_xor(ncorr, m, kWave);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        wacc[w] = A;
        wcnt[w] = ncorr;
    }
    __syncthreads();
    if (threadIdx.x < 28) {
        const int k = threadIdx.x;
        double v = 0.0;
        for (int i = 0; i < kIcpBlock / kWave; ++i) v += wacc[i].value(k);
        out[k] = v;
    } else if (threadIdx.x == 28) {
        int c = 0;
        for (int i = 0; i < kIcpBlock / kWave; ++i) c += wcnt[i];
        out[28] = (double)c;
    }
}

// Last-block: fixed-order FP64 sum of the pair's partials into sys[29] (in LDS).
__device__ __forceinline__ void sum_partials(const double *partials, int nblk, double *sys)
{
    __shared__ double part[8][kPartStride];
    const int v = threadIdx.x % kPartStride, grp = threadIdx.x / kPartStride;  // 8 groups of 32
    if (v < kNSys) {
        double a = 0.0;
        for (int b = grp; b < nblk; b += 8) a += partials[(int64_t)b * kPartStride + v];
        part[grp][v] = a;
    }
    __syncthreads();
    if (threadIdx.x < kNSys) {
        double a = 0.0;
        for (int g = 0; g < 8; ++g) a += part[g][threadIdx.x];
        sys[threadIdx.x] = a;
    }
    __syncthreads();
}

__device__ __forceinline__ void solve_and_update(IcpState *st, const double *sys)
{
    double xi[6], dT[16], T[16];
    const bool ok = solve6(sys, sys + 21, xi);
    euler_to_T(xi, dT);
    for (int k = 0; k < 16; ++k) T[k] = st->T[k];
    matmul4(dT, T, T);
    for (int k = 0; k < 16; ++k) st->T[k] = T[k];
    for (int k = 0; k < kNSys; ++k) st->sys[k] = sys[k];
    if (!ok) st->solver_failures += 1;
    const int it = st->iter + 1;
    st->iter = it;
    // stopping rule of color_icp.cu:216-228 on the increment
    const double cos_t = 0.5 * (dT[0] + dT[5] + dT[10] - 1.0);
    const double tt = (dT[3] * dT[3] + dT[7] * dT[7]) + dT[11] * dT[11];
    if (cos_t >= 1.0 - st->conv_rot && tt <= st->conv_t2) {
        st->converged = 1;
        st->done = 1;
    }
    if (it >= st->max_iter) st->done = 1;
}

// Nearest neighbour searched by a group of G lanes (G divides 64): lane g of the group scans
// candidates s0 + g, s0 + g + G, ... of every run, then the group reduces (d2, index) by
// shuffles.  All lanes of the group return the same slot.
//
// Latency structure: the bounds of all (ix, iy) runs (at most 3 x 3 when cell >= radius) are
// loaded up front, then the lane walks one flattened candidate space with kUnroll independent
// loads in flight, so a query costs a few memory round trips instead of one per candidate.
constexpr int kMaxRuns = 9;
constexpr int kUnroll = 4;

template <int G>
__device__ __forceinline__ int32_t nn_slot_group(const PairDesc &P, float qx, float qy, float qz, int g)
{
    int32_t lo[3], hi[3];
    double bd = DBL_MAX;
    int32_t bi = INT_MAX, bs = -1;
    if (grid_range(P.g, qx, qy, qz, lo, hi)) {
        const double r2 = (double)P.g.radius * (double)P.g.radius;
        const int nry = hi[1] - lo[1] + 1;
        const int nruns = (hi[0] - lo[0] + 1) * nry;
        uint32_t rs[kMaxRuns], cum[kMaxRuns + 1];
        if (nruns <= kMaxRuns) {
#pragma unroll
            for (int r = 0; r < kMaxRuns; ++r) {
                uint32_t a = 0, b = 0;
                if (r < nruns) {
                    const int ix = lo[0] + r / nry, iy = lo[1] + r % nry;
                    const int64_t row = ((int64_t)ix * P.g.dims[1] + iy) * P.g.dims[2];
                    a = P.start[row + lo[2]];
                    b = P.start[row + hi[2] + 1];
                }
                rs[r] = a;
                cum[r + 1] = b - a;
            }
            cum[0] = 0;
#pragma unroll
            for (int r = 0; r < kMaxRuns; ++r) cum[r + 1] += cum[r];
            const uint32_t total = cum[kMaxRuns];
            for (uint32_t f0 = g; f0 < total; f0 += G * kUnroll) {
                uint32_t sl[kUnroll];
                float4 c[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const uint32_t f = f0 + u * G;
                    uint32_t s = 0xffffffffu;
                    if (f < total) {
                        s = rs[0] + f;
#pragma unroll
                        for (int r = 1; r < kMaxRuns; ++r)
                            if (f >= cum[r]) s = rs[r] + (f - cum[r]);
                        c[u] = P.sxyzi[s];
                    }
                    sl[u] = s;
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    if (sl[u] == 0xffffffffu) continue;
                    const double d2 = dist2(c[u].x, c[u].y, c[u].z, qx, qy, qz);
                    if (d2 < r2 && d2 <= bd) {
                        const int32_t idx = P.order[sl[u]];
                        if (key_less(d2, idx, bd, bi)) {
                            bd = d2;
                            bi = idx;
                            bs = (int32_t)sl[u];
                        }
                    }
                }
            }
        } else {
            // cell smaller than the radius (not produced by grid_build): plain run walk
            for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
                for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
                    const int64_t row = ((int64_t)ix * P.g.dims[1] + iy) * P.g.dims[2];
                    const uint32_t s0 = P.start[row + lo[2]], s1 = P.start[row + hi[2] + 1];
                    for (uint32_t s = s0 + g; s < s1; s += G) {
                        const float4 c = P.sxyzi[s];
                        const double d2 = dist2(c.x, c.y, c.z, qx, qy, qz);
                        if (d2 < r2 && d2 <= bd) {
                            const int32_t idx = P.order[s];
                            if (key_less(d2, idx, bd, bi)) {
                                bd = d2;
                                bi = idx;
                                bs = (int32_t)s;
                            }
                        }
                    }
                }
        }
    }
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
        const double od = __shfl_xor(bd, m, kWave);
        const int32_t oi = __shfl_xor(bi, m, kWave), os = __shfl_xor(bs, m, kWave);
        if (key_less(od, oi, bd, bi)) {
            bd = od;
            bi = oi;
            bs = os;
        }
    }
    return bs;
}

template <int MODE, int G>
__global__ __launch_bounds__(kIcpBlock) void k_icp_iter(const PairDesc *__restrict__ pairs,
                                                        const int32_t *__restrict__ block_pair,
                                                        float sl, float sc)
{
    const int pid = block_pair ? block_pair[blockIdx.x] : 0;
    const PairDesc P = pairs[pid];
    IcpState *st = P.st;
    if (st->done) return;  // converged or out of iterations: the rest of the launches are no-ops
    const int lb = blockIdx.x - P.first_block;
    double T[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) T[k] = st->T[k];
    const int n = P.src_hdr->n;
    Acc<MODE> A;
    A.zero();
    int ncorr = 0;
    const int g = threadIdx.x % G;
    constexpr int kGroups = kIcpBlock / G;
    // one query per group of G lanes; the group leader accumulates the 28 sums
    for (int i0 = lb * kGroups; i0 < n; i0 += P.nblk * kGroups) {
        const int i = i0 + (int)threadIdx.x / G;
        const bool active = i < n;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        float x = 0.f, y = 0.f, z = 0.f;
        if (active) {
            p = P.src[i];
            xform_f32(T, p, x, y, z);
        }
        const int32_t s = nn_slot_group<G>(P, active ? x : NAN, y, z, g);
        if (s >= 0 && g == 0) {
            Terms t;
            corr_terms(x, y, z, p.w, P.sxyzi[s], P.snrm[s], P.sgrad[s], sl, sc, t);
            acc_terms(A, t);
            ++ncorr;
        }
    }
    __shared__ double sys[kPartStride];
    __shared__ bool last;
    double *my = P.partials + (int64_t)lb * kPartStride;
    block_partial<MODE, G>(A, ncorr, sys);
    if (threadIdx.x < kNSys) my[threadIdx.x] = sys[threadIdx.x];
    // two-level deterministic finish: the last block of each cluster sums the cluster's block
    // partials, the last cluster sums the cluster partials and solves
    fence_release_agent();
    __syncthreads();
    const int cl = lb / P.cluster;
    const int csize = min(P.cluster, P.nblk - cl * P.cluster);
    if (threadIdx.x == 0) last = atomicAdd(&st->cl_ticket[cl * kTicketStride], 1u) == (uint32_t)(csize - 1);
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    double *cpart = P.partials + (int64_t)P.nblk * kPartStride;
    sum_partials(P.partials + (int64_t)cl * P.cluster * kPartStride, csize, sys);
    if (threadIdx.x < kNSys) cpart[(int64_t)cl * kPartStride + threadIdx.x] = sys[threadIdx.x];
    if (threadIdx.x == 0) st->cl_ticket[cl * kTicketStride] = 0u;
    fence_release_agent();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&st->ticket, 1u) == (uint32_t)(P.nclust - 1);
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    sum_partials(cpart, P.nclust, sys);
    if (threadIdx.x == 0) {
        solve_and_update(st, sys);
        st->ticket = 0u;
    }
}

// Gauss-Newton system for given correspondences (no solve): writes sys into st->sys.
template <int MODE>
__global__ __launch_bounds__(kIcpBlock) void k_build_gn(PairDesc P, const float4 *__restrict__ tgt_xyzi,
                                                        const int32_t *__restrict__ nn, double t00, double t01,
                                                        double t02, double t03, double t10, double t11,
                                                        double t12, double t13, double t20, double t21,
                                                        double t22, double t23, float sl, float sc)
{
    IcpState *st = P.st;
    const double T[12] = {t00, t01, t02, t03, t10, t11, t12, t13, t20, t21, t22, t23};
    const int n = P.src_hdr->n;
    Acc<MODE> A;
    A.zero();
    int ncorr = 0;
    for (int i = blockIdx.x * kIcpBlock + threadIdx.x; i < n; i += P.nblk * kIcpBlock) {
        const int32_t j = nn[i];
        if (j < 0) continue;
        const float4 p = P.src[i];
        float x, y, z;
        xform_f32(T, p, x, y, z);
        Terms t;
        corr_terms(x, y, z, p.w, tgt_xyzi[j], P.nrm[j], P.grad[j], sl, sc, t);
        acc_terms(A, t);
        ++ncorr;
    }
    __shared__ double sys[kPartStride];
    __shared__ bool last;
    block_partial(A, ncorr, sys);
    if (threadIdx.x < kNSys) P.partials[(int64_t)blockIdx.x * kPartStride + threadIdx.x] = sys[threadIdx.x];
    fence_release_agent();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(&st->ticket, 1u) == (uint32_t)(P.nblk - 1);
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    sum_partials(P.partials, P.nblk, sys);
    if (threadIdx.x < kNSys) st->sys[threadIdx.x] = sys[threadIdx.x];
    if (threadIdx.x == 0) st->ticket = 0u;
}

// ---- generic FP64 block sums with last-block finish (residuals, objective) -----------------
template <int NV>
__device__ __forceinline__ void sum_finish(double (&v)[NV], double *partials, int nblk, uint32_t *ticket,
                                           double *result)
{
    __shared__ double wv[kIcpBlock / kWave][NV];
    __shared__ bool last;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane_id() == 0)
        for (int k = 0; k < NV; ++k) wv[threadIdx.x / kWave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        double a = 0.0;
        for (int w = 0; w < kIcpBlock / kWave; ++w) a += wv[w][threadIdx.x];
        partials[(int64_t)blockIdx.x * kPartStride + threadIdx.x] = a;
    }
    fence_release_agent();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == (uint32_t)(nblk - 1);
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    if (threadIdx.x < NV) {
        double a = 0.0;
        for (int b = 0; b < nblk; ++b) a += partials[(int64_t)b * kPartStride + threadIdx.x];
        result[threadIdx.x] = a;
    }
    if (threadIdx.x == 0) *ticket = 0u;
}

__global__ __launch_bounds__(kIcpBlock) void k_residuals(PairDesc P, const float4 *__restrict__ tgt_xyzi,
                                                         const int32_t *__restrict__ nn, const double *__restrict__ Tm,
                                                         float lam, float sl, float sc, int kind,
                                                         float *__restrict__ out, double *__restrict__ result)
{
    double T[12];
    for (int k = 0; k < 12; ++k) T[k] = Tm[k];
    const int n = P.src_hdr->n;
    double v[1] = {0.0};
    for (int i = blockIdx.x * kIcpBlock + threadIdx.x; i < n; i += P.nblk * kIcpBlock) {
        const int32_t j = nn[i];
        float rg = 0.f, rc = 0.f, rg_ = 0.f, rc_ = 0.f;
        if (j >= 0) {
            const float4 p = P.src[i];
            float x, y, z;
            xform_f32(T, p, x, y, z);
            Terms t;
            corr_terms(x, y, z, p.w, tgt_xyzi[j], P.nrm[j], P.grad[j], sl, sc, t);
            rg = t.rg;
            rc = t.rc;
            // unweighted residuals for the objective form (.cu:117-130)
            Terms u;
            corr_terms(x, y, z, p.w, tgt_xyzi[j], P.nrm[j], P.grad[j], 1.0f, 1.0f, u);
            rg_ = u.rg;
            rc_ = u.rc;
        }
        if (kind == GCA_RESIDUAL_PAIR) {
            if (out) {
                out[2 * i] = rg;
                out[2 * i + 1] = rc;
            }
            v[0] += (double)rg + (double)rc;
        } else if (kind == GCA_RESIDUAL_ABS) {
            const float a = fabsf(rg) + fabsf(rc);
            if (out) out[i] = a;
            v[0] += a;
        } else {
            const float a = lam * rg_ * rg_ + (1.0f - lam) * rc_ * rc_;
            if (out) out[i] = a;
            v[0] += a;
        }
    }
    sum_finish<1>(v, P.partials, P.nblk, &P.st->ticket, result);
}

// Objective E and its adjoint at pose dT(xi) * T with frozen correspondences, all in FP64.
__global__ __launch_bounds__(kIcpBlock) void k_objective_grad(PairDesc P, const float4 *__restrict__ tgt_xyzi,
                                                              const int32_t *__restrict__ nn,
                                                              const double *__restrict__ prm, double lam,
                                                              float4 *__restrict__ dpts, double *__restrict__ result)
{
    // prm: T[16] | dT[16] | RzRy[9] | Rx[9] | M[9]
    __shared__ double s[59];
    if (threadIdx.x < 59) s[threadIdx.x] = prm[threadIdx.x];
    __syncthreads();
    const double *T = s, *dT = s + 16, *RzRy = s + 32, *Rx = s + 41, *M = s + 50;
    const int n = P.src_hdr->n;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * kIcpBlock + threadIdx.x; i < n; i += P.nblk * kIcpBlock) {
        const int32_t j = nn[i];
        if (j < 0) {
            if (dpts) dpts[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        const float4 pf = P.src[i];
        double y[3], z[3], p[3];
        for (int a = 0; a < 3; ++a)
            y[a] = ((T[4 * a] * (double)pf.x + T[4 * a + 1] * (double)pf.y) + T[4 * a + 2] * (double)pf.z) + T[4 * a + 3];
        for (int a = 0; a < 3; ++a) z[a] = (dT[4 * a] * y[0] + dT[4 * a + 1] * y[1]) + dT[4 * a + 2] * y[2];
        for (int a = 0; a < 3; ++a) p[a] = z[a] + dT[4 * a + 3];
        const float4 q4 = tgt_xyzi[j], n4 = P.nrm[j], d4 = P.grad[j];
        const double q[3] = {q4.x, q4.y, q4.z}, nn3[3] = {n4.x, n4.y, n4.z}, d[3] = {d4.x, d4.y, d4.z};
        const double dq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
        const double rg = (dq[0] * nn3[0] + dq[1] * nn3[1]) + dq[2] * nn3[2];
        const double dn = (d[0] * nn3[0] + d[1] * nn3[1]) + d[2] * nn3[2];
        const double g[3] = {d[0] - dn * nn3[0], d[1] - dn * nn3[1], d[2] - dn * nn3[2]};
        const double pr[3] = {p[0] - rg * nn3[0] - q[0], p[1] - rg * nn3[1] - q[1], p[2] - rg * nn3[2] - q[2]};
        const double rc = ((double)q4.w + ((d[0] * pr[0] + d[1] * pr[1]) + d[2] * pr[2])) - (double)pf.w;
        v[0] += lam * rg * rg + (1.0 - lam) * rc * rc;
        double u[3];
        for (int a = 0; a < 3; ++a) u[a] = 2.0 * lam * rg * nn3[a] + 2.0 * (1.0 - lam) * rc * g[a];
        double w[3], w2[3], xy[3];
        for (int a = 0; a < 3; ++a) w[a] = (dT[a] * u[0] + dT[4 + a] * u[1]) + dT[8 + a] * u[2];
        v[1] += y[1] * w[2] - y[2] * w[1];
        for (int a = 0; a < 3; ++a) w2[a] = (RzRy[a] * u[0] + RzRy[3 + a] * u[1]) + RzRy[6 + a] * u[2];
        for (int a = 0; a < 3; ++a) xy[a] = (Rx[3 * a] * y[0] + Rx[3 * a + 1] * y[1]) + Rx[3 * a + 2] * y[2];
        v[2] += xy[2] * w2[0] - xy[0] * w2[2];
        v[3] += z[0] * u[1] - z[1] * u[0];
        v[4] += u[0];
        v[5] += u[1];
        v[6] += u[2];
        if (dpts) {
            float4 o;
            o.x = (float)((M[0] * u[0] + M[3] * u[1]) + M[6] * u[2]);
            o.y = (float)((M[1] * u[0] + M[4] * u[1]) + M[7] * u[2]);
            o.z = (float)((M[2] * u[0] + M[5] * u[1]) + M[8] * u[2]);
            o.w = (float)(-2.0 * (1.0 - lam) * rc);
            dpts[i] = o;
        }
    }
    sum_finish<7>(v, P.partials, P.nblk, &P.st->ticket, result);
}

__global__ void k_state_init(IcpState *st, int n_states, const double *__restrict__ T0, int max_iter,
                             double conv_rot, double conv_t2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_states) return;
    IcpState &s = st[p];
    for (int k = 0; k < 16; ++k) s.T[k] = T0[16 * p + k];
    for (int k = 0; k < kNSys; ++k) s.sys[k] = 0.0;
    s.conv_rot = conv_rot;
    s.conv_t2 = conv_t2;
    s.iter = 0;
    s.max_iter = max_iter;
    s.done = max_iter <= 0 ? 1 : 0;
    s.converged = 0;
    s.solver_failures = 0;
    s.ticket = 0u;
    for (int c = 0; c < kMaxClusters; ++c) s.cl_ticket[c * kTicketStride] = 0u;
}

// ---------------------------------------------------------------------------------------------
// host side

static int32_t icp_blocks(int64_t n) { return (int32_t)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kIcpBlock - 1) / kIcpBlock)); }

// Lanes per correspondence query: wide groups shorten each query's dependent load chain,
// which is what bounds small clouds; narrow groups keep large clouds' wave count reasonable.
static int icp_group_lanes(gca_ctx ctx, int64_t max_n)
{
    const int key = max_n < 16384 ? GCA_TUNE_ICP_LANES_SMALL
                                  : (max_n < 131072 ? GCA_TUNE_ICP_LANES_MEDIUM : GCA_TUNE_ICP_LANES_LARGE);
    if (ctx->tune[key]) return ctx->tune[key];
    if (max_n < 16384) return 16;
    if (max_n < 131072) return 8;
    return 4;
}

static int32_t icp_blocks_g(gca_ctx ctx, int64_t n, int G)
{
    const int64_t groups_per_block = kIcpBlock / G;
    const int64_t cap = ctx->tune[GCA_TUNE_ICP_MAX_BLOCKS] ? ctx->tune[GCA_TUNE_ICP_MAX_BLOCKS] : 1024;
    return (int32_t)std::min<int64_t>(cap, std::max<int64_t>(1, (n + groups_per_block - 1) / groups_per_block));
}

struct IcpWork {
    IcpState *states = nullptr;
    double *partials = nullptr;
    PairDesc *pairs = nullptr;
    int32_t *block_pair = nullptr;
    double *aux = nullptr;  // T0 upload / results
};

constexpr int kIcpCluster = 32;  // block partials per cluster of the two-level finish

static int32_t clusters_of(int32_t nblk) { return (nblk + kIcpCluster - 1) / kIcpCluster; }

// partial slots: every pair's block partials followed by its cluster partials (<= 2 per block)
static int32_t icp_work(gca_ctx ctx, int n_pairs, int64_t total_blocks, IcpWork &w)
{
    w.states = (IcpState *)ctx_scratch(ctx, 13, sizeof(IcpState) * n_pairs, false);
    w.partials = (double *)ctx_scratch(ctx, 14, sizeof(double) * kPartStride * (2 * total_blocks + 1), false);
    if (!w.states || !w.partials) return GCA_ERR_OUT_OF_MEMORY;
    const size_t aux_doubles = std::max<size_t>(16 * (size_t)n_pairs, 64);
    const size_t need = sizeof(PairDesc) * n_pairs + sizeof(int32_t) * total_blocks + sizeof(double) * aux_doubles + 1024;
    if (!ctx->icp_state || ctx->icp_state_bytes < need) {
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        hipFree(ctx->icp_state);
        ctx->icp_state = nullptr;
        size_t want = need * 2 + 4096;
        GCA_HIP(hipMalloc(&ctx->icp_state, want));
        ctx->icp_state_bytes = want;
    }
    char *base = (char *)ctx->icp_state;
    w.pairs = (PairDesc *)base;
    base += ((sizeof(PairDesc) * n_pairs + 255) & ~size_t(255));
    w.aux = (double *)base;
    base += ((sizeof(double) * aux_doubles + 255) & ~size_t(255));
    w.block_pair = (int32_t *)base;
    return GCA_OK;
}

static PairDesc make_pair(gca_cloud src, gca_cloud tgt, IcpState *st, double *partials, int32_t nblk,
                          int32_t first, float radius)
{
    PairDesc P{};
    P.src = src->xyzi;
    P.src_hdr = src->hdr;
    if (tgt) {
        P.g = tgt->grid.desc;
        P.g.radius = radius;
        P.start = tgt->grid.start;
        P.order = tgt->grid.order;
        P.sxyzi = tgt->grid.sxyzi;
        P.nrm = tgt->nrm;
        P.grad = tgt->grad;
        P.snrm = tgt->grid.snrm;
        P.sgrad = tgt->grid.sgrad;
    }
    P.st = st;
    P.partials = partials;
    P.nblk = nblk;
    P.first_block = first;
    P.cluster = kIcpCluster;
    P.nclust = clusters_of(nblk);
    return P;
}

template <int MODE>
static void launch_iter_mode(gca_ctx ctx, int G, const PairDesc *pairs, const int32_t *bp, dim3 g, float sl, float sc)
{
    dim3 b(kIcpBlock);
    switch (G) {
    case 64: hipLaunchKernelGGL((k_icp_iter<MODE, 64>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    case 32: hipLaunchKernelGGL((k_icp_iter<MODE, 32>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    case 16: hipLaunchKernelGGL((k_icp_iter<MODE, 16>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    case 8: hipLaunchKernelGGL((k_icp_iter<MODE, 8>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    case 4: hipLaunchKernelGGL((k_icp_iter<MODE, 4>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    case 2: hipLaunchKernelGGL((k_icp_iter<MODE, 2>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    default: hipLaunchKernelGGL((k_icp_iter<MODE, 1>), g, b, 0, ctx->stream, pairs, bp, sl, sc); break;
    }
}

static void launch_iter(gca_ctx ctx, int mode, int G, const PairDesc *pairs, const int32_t *bp, int64_t nblocks,
                        float sl, float sc)
{
    dim3 g((unsigned)nblocks);
    StageTimer timer(ctx, GCA_STAGE_ICP_ITER);
    if (mode == GCA_ACCUM_NAIVE_FP32)
        launch_iter_mode<GCA_ACCUM_NAIVE_FP32>(ctx, G, pairs, bp, g, sl, sc);
    else if (mode == GCA_ACCUM_FP64)
        launch_iter_mode<GCA_ACCUM_FP64>(ctx, G, pairs, bp, g, sl, sc);
    else
        launch_iter_mode<GCA_ACCUM_COMPENSATED>(ctx, G, pairs, bp, g, sl, sc);
}

static int32_t prepare_target(gca_ctx ctx, gca_cloud tgt, float corr, float grad_r)
{
    if (!tgt->has_normals) return GCA_ERR_NOT_READY;
    if (!tgt->has_gradient) {
        if (!(grad_r > 0.0f)) return GCA_ERR_NOT_READY;
        GCA_TRY(gca_compute_color_gradient(ctx, tgt, grad_r, 0));
    }
    GCA_TRY(grid_build(ctx, tgt, corr));
    GCA_TRY(grid_gather_attributes(ctx, tgt));
    return GCA_OK;
}

static int32_t validate_params(const gca_icp_params *p)
{
    if (!p) return GCA_ERR_INVALID_VALUE;
    if (!(p->max_correspondence_distance > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!(p->lambda_geometric >= 0.0f) || p->lambda_geometric > 1.0f) return GCA_ERR_INVALID_VALUE;  // .cu:438-441
    if (p->max_iterations < 0 || p->accumulate_mode < 0 || p->accumulate_mode > 2) return GCA_ERR_INVALID_VALUE;
    return GCA_OK;
}

// Runs the batched loop.  T0/T_out: n x 16, reports n.
static int32_t run_batch(gca_ctx ctx, int n, const gca_cloud *srcs, const gca_cloud *tgts, const gca_icp_params *p,
                         float grad_r, const double *T0, double *T_out, gca_icp_report *reps)
{
    GCA_TRY(validate_params(p));
    std::vector<int64_t> ns(n);
    std::vector<int32_t> nblk(n), first(n);
    int64_t total = 0, max_n = 0;
    for (int k = 0; k < n; ++k) {
        if (!srcs[k] || !tgts[k]) return GCA_ERR_INVALID_VALUE;
        GCA_TRY(prepare_target(ctx, tgts[k], p->max_correspondence_distance, grad_r));
        GCA_TRY(cloud_sync_header(srcs[k]));
        ns[k] = srcs[k]->host_hdr.n;
        max_n = std::max(max_n, ns[k]);
    }
    const int G = icp_group_lanes(ctx, max_n);
    for (int k = 0; k < n; ++k) {
        nblk[k] = icp_blocks_g(ctx, ns[k], G);
        first[k] = (int32_t)total;
        total += nblk[k];
    }
    IcpWork w;
    GCA_TRY(icp_work(ctx, n, total, w));
    std::vector<PairDesc> P(n);
    std::vector<int32_t> bp((size_t)total);
    int64_t slot = 0;
    for (int k = 0; k < n; ++k) {
        P[k] = make_pair(srcs[k], tgts[k], w.states + k, w.partials + slot * kPartStride, nblk[k], first[k],
                         p->max_correspondence_distance);
        slot += nblk[k] + P[k].nclust;
        for (int b = 0; b < nblk[k]; ++b) bp[first[k] + b] = k;
    }
    GCA_HIP(hipMemcpyAsync(w.pairs, P.data(), sizeof(PairDesc) * n, hipMemcpyHostToDevice, ctx->stream));
    GCA_HIP(hipMemcpyAsync(w.block_pair, bp.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice, ctx->stream));
    GCA_HIP(hipMemcpyAsync(w.aux, T0, sizeof(double) * 16 * n, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_state_init, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, w.states, n, (const double *)w.aux,
                       p->max_iterations, p->conv_rotation, p->conv_translation_sq);
    GCA_HIP(hipGetLastError());
    const float sl = sqrtf(p->lambda_geometric), sc = sqrtf(1.0f - p->lambda_geometric);
    for (int it = 0; it < p->max_iterations; ++it)
        launch_iter(ctx, p->accumulate_mode, G, w.pairs, n > 1 ? w.block_pair : nullptr, total, sl, sc);
    GCA_HIP(hipGetLastError());
    std::vector<IcpState> S(n);
    // results only (the ticket lines after them stay on the device)
    GCA_HIP(hipMemcpy2DAsync(S.data(), sizeof(IcpState), w.states, sizeof(IcpState), offsetof(IcpState, ticket), n,
                             hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    int32_t status = GCA_OK;
    for (int k = 0; k < n; ++k) {
        memcpy(T_out + 16 * k, S[k].T, sizeof(double) * 16);
        if (reps) {
            gca_icp_report &r = reps[k];
            r.iterations = S[k].iter;
            r.converged = S[k].converged;
            r.solver_failures = S[k].solver_failures;
            r.n_correspondences = (int32_t)S[k].sys[28];
            r.rmse = ns[k] > 0 ? S[k].sys[27] / (double)ns[k] : 0.0;  // color_icp.cu:231
            r.fitness = ns[k] > 0 ? S[k].sys[28] / (double)ns[k] : 0.0;
        }
    }
    return status;
}

static int32_t single_pair(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn, IcpWork &w, PairDesc &P,
                           int64_t &n)
{
    GCA_TRY(cloud_sync_header(src));
    n = src->host_hdr.n;
    const int32_t nb = icp_blocks(n);
    GCA_TRY(icp_work(ctx, 1, nb, w));
    P = make_pair(src, nullptr, w.states, w.partials, nb, 0, 0.f);
    P.nrm = tgt->nrm;
    P.grad = tgt->grad;
    (void)nn;
    static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    GCA_HIP(hipMemcpyAsync(w.aux, I, sizeof(I), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, ctx->stream, w.states, 1, (const double *)w.aux, 0, 0.0, 0.0);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

}  // namespace gca

using namespace gca;

extern "C" {

int32_t gca_build_gauss_newton(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn_dev, float lambda,
                               const double T[16], int32_t mode, double out[29])
{
    if (!ctx || !src || !tgt || !nn_dev || !T || !out) return GCA_ERR_INVALID_VALUE;
    if (lambda > 1.0f || !(lambda >= 0.0f) || mode < 0 || mode > 2) return GCA_ERR_INVALID_VALUE;  // .cu:438-441
    if (!tgt->has_normals || !tgt->has_gradient) return GCA_ERR_NOT_READY;                       // .cu:449-452
    IcpWork w;
    PairDesc P;
    int64_t n;
    GCA_TRY(single_pair(ctx, src, tgt, nn_dev, w, P, n));
    const float sl = sqrtf(lambda), sc = sqrtf(1.0f - lambda);
    dim3 g((unsigned)P.nblk), b(kIcpBlock);
    if (mode == GCA_ACCUM_NAIVE_FP32)
        hipLaunchKernelGGL(k_build_gn<GCA_ACCUM_NAIVE_FP32>, g, b, 0, ctx->stream, P, tgt->xyzi, nn_dev, T[0], T[1],
                           T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11], sl, sc);
    else if (mode == GCA_ACCUM_FP64)
        hipLaunchKernelGGL(k_build_gn<GCA_ACCUM_FP64>, g, b, 0, ctx->stream, P, tgt->xyzi, nn_dev, T[0], T[1], T[2],
                           T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11], sl, sc);
    else
        hipLaunchKernelGGL(k_build_gn<GCA_ACCUM_COMPENSATED>, g, b, 0, ctx->stream, P, tgt->xyzi, nn_dev, T[0], T[1],
                           T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11], sl, sc);
    GCA_HIP(hipGetLastError());
    IcpState S;
    GCA_HIP(hipMemcpyAsync(&S, w.states, offsetof(IcpState, ticket), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(out, S.sys, sizeof(double) * kNSys);
    return GCA_OK;
}

int32_t gca_compute_residuals(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn_dev, float lambda,
                              const double T[16], int32_t kind, float *out_dev, double *sum)
{
    if (!ctx || !src || !tgt || !nn_dev || !T || kind < 0 || kind > 2) return GCA_ERR_INVALID_VALUE;
    if (lambda > 1.0f || !(lambda >= 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!tgt->has_normals || !tgt->has_gradient) return GCA_ERR_NOT_READY;
    IcpWork w;
    PairDesc P;
    int64_t n;
    GCA_TRY(single_pair(ctx, src, tgt, nn_dev, w, P, n));
    GCA_HIP(hipMemcpyAsync(w.aux + 16, T, sizeof(double) * 16, hipMemcpyHostToDevice, ctx->stream));
    double *res = w.aux + 32;
    const float sl = sqrtf(lambda), sc = sqrtf(1.0f - lambda);
    hipLaunchKernelGGL(k_residuals, dim3((unsigned)P.nblk), dim3(kIcpBlock), 0, ctx->stream, P, tgt->xyzi, nn_dev,
                       (const double *)(w.aux + 16), lambda, sl, sc, kind, out_dev, res);
    GCA_HIP(hipGetLastError());
    double s = 0.0;
    GCA_HIP(hipMemcpyAsync(&s, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    if (sum) *sum = s;
    return GCA_OK;
}

int32_t gca_solve_gauss_newton(const double jtj[21], const double jtr[6], double xi[6], double T_step[16])
{
    if (!jtj || !jtr || !T_step) return GCA_ERR_INVALID_VALUE;
    // FP64 LDL^T on the host (same factorisation as the device solve)
    double A[6][6], L[6][6] = {{0}}, D[6], x[6];
    int k = 0;
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) A[a][b] = A[b][a] = jtj[k++];
    bool ok = true;
    for (int j = 0; j < 6 && ok; ++j) {
        double dj = A[j][j];
        for (int m = 0; m < j; ++m) dj -= L[j][m] * L[j][m] * D[m];
        if (!(dj > 0.0) || !std::isfinite(dj)) {
            ok = false;
            break;
        }
        D[j] = dj;
        for (int i = j + 1; i < 6; ++i) {
            double v = A[i][j];
            for (int m = 0; m < j; ++m) v -= L[i][m] * L[j][m] * D[m];
            L[i][j] = v / dj;
        }
    }
    if (ok) {
        double y[6];
        for (int i = 0; i < 6; ++i) {
            double v = -jtr[i];
            for (int m = 0; m < i; ++m) v -= L[i][m] * y[m];
            y[i] = v;
        }
        for (int i = 0; i < 6; ++i) y[i] /= D[i];
        for (int i = 5; i >= 0; --i) {
            double v = y[i];
            for (int m = i + 1; m < 6; ++m) v -= L[m][i] * x[m];
            x[i] = v;
        }
        for (int i = 0; i < 6; ++i)
            if (!std::isfinite(x[i])) ok = false;
    }
    if (!ok)
        for (int i = 0; i < 6; ++i) x[i] = 0.0;
    const double ca = cos(x[0]), sa = sin(x[0]), cb = cos(x[1]), sb = sin(x[1]), cc = cos(x[2]), sc = sin(x[2]);
    const double T[16] = {cc * cb, cc * sb * sa - sc * ca, cc * sb * ca + sc * sa, x[3],
                          sc * cb, sc * sb * sa + cc * ca, sc * sb * ca - cc * sa, x[4],
                          -sb,     cb * sa,                cb * ca,                x[5],
                          0.0,     0.0,                    0.0,                    1.0};
    memcpy(T_step, T, sizeof(T));
    if (xi) memcpy(xi, x, sizeof(x));
    return ok ? GCA_OK : GCA_ERR_NOT_SPD;
}

int32_t gca_color_icp_align(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const gca_icp_params *p, float grad_r,
                            const double T0[16], double T_out[16], gca_icp_report *rep)
{
    if (!ctx || !src || !tgt || !T_out) return GCA_ERR_INVALID_VALUE;
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    return run_batch(ctx, 1, &src, &tgt, p, grad_r, T0 ? T0 : I, T_out, rep);
}

int32_t gca_color_icp_batch(gca_ctx ctx, int32_t n_pairs, const gca_cloud *srcs, const gca_cloud *tgts,
                            const gca_icp_params *p, float grad_r, const double *T0, double *T_out,
                            gca_icp_report *reps)
{
    if (!ctx || n_pairs < 0 || (n_pairs > 0 && (!srcs || !tgts || !T0 || !T_out))) return GCA_ERR_INVALID_VALUE;
    if (n_pairs == 0) return GCA_OK;
    return run_batch(ctx, n_pairs, srcs, tgts, p, grad_r, T0, T_out, reps);
}

int32_t gca_multiscale_icp(gca_ctx ctx, gca_cloud src_full, gca_cloud tgt_full, int32_t n_levels,
                           const gca_level_params *levels, const double T0[16], double T_out[16],
                           gca_icp_report *reports)
{
    if (!ctx || !src_full || !tgt_full || !levels || !T_out || n_levels < 1 || n_levels > 8)
        return GCA_ERR_INVALID_VALUE;
    double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (T0) memcpy(T, T0, sizeof(T));
    for (int l = 0; l < n_levels; ++l) {
        const gca_level_params &L = levels[l];
        gca_cloud *cs = &ctx->level_clouds[2 * l], *ct = &ctx->level_clouds[2 * l + 1];
        if (!*cs) GCA_TRY(gca_cloud_create(ctx, 1, cs));
        if (!*ct) GCA_TRY(gca_cloud_create(ctx, 1, ct));
        GCA_TRY(gca_voxel_downsample(ctx, src_full, L.voxel_size, *cs));
        GCA_TRY(gca_voxel_downsample(ctx, tgt_full, L.voxel_size, *ct));
        GCA_TRY(gca_estimate_normals(ctx, *ct, L.normal_radius, L.normal_max_nn));
        GCA_TRY(gca_compute_color_gradient(ctx, *ct, L.gradient_radius, L.gradient_max_nn));
        gca_icp_report rep{};
        GCA_TRY(gca_color_icp_align(ctx, *cs, *ct, &L.icp, L.gradient_radius, T, T, &rep));
        if (reports) reports[l] = rep;
    }
    memcpy(T_out, T, sizeof(T));
    return GCA_OK;
}

int32_t gca_color_icp_objective_grad(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn_dev,
                                     float lambda, const double T[16], const double xi[6], double *E,
                                     double dE_dxi[6], float *dpts_dev)
{
    if (!ctx || !src || !tgt || !nn_dev || !T || !xi) return GCA_ERR_INVALID_VALUE;
    if (lambda > 1.0f || !(lambda >= 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!tgt->has_normals || !tgt->has_gradient) return GCA_ERR_NOT_READY;
    IcpWork w;
    PairDesc P;
    int64_t n;
    GCA_TRY(single_pair(ctx, src, tgt, nn_dev, w, P, n));
    double prm[59];
    memcpy(prm, T, sizeof(double) * 16);
    const double ca = cos(xi[0]), sa = sin(xi[0]), cb = cos(xi[1]), sb = sin(xi[1]), cc = cos(xi[2]), sc = sin(xi[2]);
    double *dT = prm + 16;
    const double dTv[16] = {cc * cb, cc * sb * sa - sc * ca, cc * sb * ca + sc * sa, xi[3],
                            sc * cb, sc * sb * sa + cc * ca, sc * sb * ca - cc * sa, xi[4],
                            -sb,     cb * sa,                cb * ca,                xi[5],
                            0.0,     0.0,                    0.0,                    1.0};
    memcpy(dT, dTv, sizeof(dTv));
    const double RzRy[9] = {cc * cb, -sc, cc * sb, sc * cb, cc, sc * sb, -sb, 0.0, cb};
    const double Rx[9] = {1, 0, 0, 0, ca, -sa, 0, sa, ca};
    memcpy(prm + 32, RzRy, sizeof(RzRy));
    memcpy(prm + 41, Rx, sizeof(Rx));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            prm[50 + 3 * i + j] = (dT[4 * i] * T[j] + dT[4 * i + 1] * T[4 + j]) + dT[4 * i + 2] * T[8 + j];
    double *dprm = (double *)ctx_scratch(ctx, 10, sizeof(double) * 80, false);
    if (!dprm) return GCA_ERR_OUT_OF_MEMORY;
    GCA_HIP(hipMemcpyAsync(dprm, prm, sizeof(prm), hipMemcpyHostToDevice, ctx->stream));
    double *res = dprm + 64;
    hipLaunchKernelGGL(k_objective_grad, dim3((unsigned)P.nblk), dim3(kIcpBlock), 0, ctx->stream, P, tgt->xyzi, nn_dev,
                       (const double *)dprm, (double)lambda, (float4 *)dpts_dev, res);
    GCA_HIP(hipGetLastError());
    double r[7];
    GCA_HIP(hipMemcpyAsync(r, res, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    if (E) *E = r[0];
    if (dE_dxi) memcpy(dE_dxi, r + 1, sizeof(double) * 6);
    return GCA_OK;
}

}  // extern "C"
