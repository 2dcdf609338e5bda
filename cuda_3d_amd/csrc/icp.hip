// icp.hip -- colored ICP on the device.
//
// Replaces color_icp::align (src/registration/color_icp.cu:151-234), whose every iteration is
//   point_cloud::nn_search (grid rebuild + sort + nn_search_functor, cuda_nn_search.cu:545-619)
//   cuda_build_gauss_newton_color_icp (thrust::transform_reduce of 172-byte tuples,
//       cuda_color_icp_build_least_square.cu:134-266,431-474) + device->host copy
//   solve_JTJ_JTr<true> (host Eigen LDLT, eigen_solver.hpp:44-85)
//   point_cloud::transform (in-place FP32 transform + bounds reduce + device->host copy),
// by ONE persistent launch that runs every iteration of every registration state (a source /
// target pair at one pyramid level):
//   - an iteration of a state is split into `nvb` work units (fixed strided sets of source
//     points), dealt round-robin into kSubQ sub-queues.  Workgroups claim units from a
//     sub-queue's word (epoch << 32 | units claimed) with one atomic, so any number of resident
//     workgroups serve any number of states and a workgroup never waits on work that nobody has
//     claimed;
//   - a unit maps its source points by the current pose (FP64 -> FP32), finds the nearest
//     target point in the cached target grid (exact; the search box is shrunk by the previous
//     iteration's match), forms residuals and 6-D Jacobians in FP32 (reference formulas),
//     sums the 28 products over its points (TwoSum-compensated FP32 within a pass of up to 256
//     points, FP64 across passes) and writes one FP64 partial;
//   - the workgroup that completes the last unit of an epoch (ticket) sums the partials in a
//     fixed order in FP64, solves the 6x6 system by LDL^T, applies the Euler increment to the
//     FP64 pose, evaluates the stopping rule and opens the next epoch -- or, when the state is
//     done, hands its pose to the next level of its chain and opens that.
// Units are a fixed partition of the points, so the result does not depend on which workgroup
// ran which unit: the one-launch-per-iteration variant (k_icp_iter) gives bit-identical poses.
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
constexpr int kPartStride = 32;  // doubles per partial
constexpr int kMaxUnits = 1024;  // work units (partials) per state and iteration
static_assert(kNSys <= kWave, "a partial is written by wavefront 0, which also takes the ticket");

constexpr uint32_t kEpochWait = 0xfffffffeu;  // chained level whose predecessor still runs
constexpr uint32_t kEpochDone = 0xffffffffu;
// Idle polls after which a workgroup of the persistent launch gives up and flags an error.  A
// poll is at least one device-scope memory round trip (~1 us), so this is seconds: it only
// ends a launch whose states can no longer make progress.
constexpr uint32_t kSpinLimit = 1u << 22;

// Claim words of a state: (epoch << 32) | units claimed, one per sub-queue, each on its own
// 256-byte line.  Every workgroup that finds an epoch open attempts a claim; with one word those
// atomics serialise (~11 ns each, ~6 us for 512 workgroups), with kSubQ words they spread.
constexpr int kSubQ = 8;

struct alignas(256) QWord {
    unsigned long long w;
};

// Units of sub-queue x of a state with nvb units.
__host__ __device__ inline uint32_t subq_units(int nvb, int x)
{
    return nvb > x ? (uint32_t)((nvb - x + kSubQ - 1) / kSubQ) : 0u;
}

// Unit c of sub-queue x: the sub-queues hold consecutive ranges of units (in the order above:
// sizes q + 1 for the first nvb mod kSubQ, q after), so the workgroups of one XCD, which serve
// their own sub-queue first, register neighbouring source points against one target region.
__host__ __device__ inline int subq_unit(int nvb, int x, int c)
{
    const int q = nvb / kSubQ, r = nvb % kSubQ;
    return x * q + (x < r ? x : r) + c;
}

struct IcpState {
    double T[16];
    double sys[kNSys];
    double conv_rot, conv_t2;
    int32_t iter, max_iter, done, converged, solver_failures, n_src;
    int32_t G, nvb, n_tgt, next;  // lanes per query, work units, target size, next level (-1)
    long long t_init, t_done;     // device wall clock at set-up and when the state finished
    long long t_start;            // ... and when its first iteration's work started (0 = not yet)
    QWord q[kSubQ];                        // sub-queue x holds units x, x + kSubQ, ...
    alignas(256) uint32_t ticket;          // units completed in the current epoch
    QWord bar[kSubQ];                      // resident launch: arrivals of workgroup group x (b % kSubQ)
    alignas(256) uint32_t bar_top;         // ... and of completed groups
    alignas(256) uint32_t pad_end;
};

struct PairDesc {
    const float4 *src;
    const CloudHdr *src_hdr;
    const CloudHdr *tgt_hdr;
    GridDesc g;
    const uint32_t *start;
    const int32_t *order;
    const float4 *sxyzi;
    const float4 *nrm;   // target normals / gradients by original index
    const float4 *grad;
    const float4 *snrm;  // ... and in grid (slot) order
    const float4 *sgrad;
    double *partials;    // one FP64 partial (kPartStride doubles) per work unit
    int32_t *nnprev;     // matched target slot of each source point in the previous iteration
    IcpState *st;        // single-pair kernels (Gauss-Newton build, residuals, objective)
    float sl, sc;        // sqrt(lambda), sqrt(1 - lambda)
    int32_t nblk;        // single-pair kernels: workgroups
    int32_t unit_cap;    // cap on work units per iteration
    int32_t lanes[3];    // lanes per query for n < 16384, n < 49152, larger
};

// Launch-shape rules shared by host and device: the unit partition fixes the summation order, so
// it must come out the same wherever it is computed.
__host__ __device__ inline int icp_lanes_for(const int32_t lanes[3], int64_t n)
{
    return n < 16384 ? lanes[0] : (n < 49152 ? lanes[1] : lanes[2]);
}

__host__ __device__ inline int32_t icp_units_for(int64_t n, int G, int32_t cap)
{
    const int64_t q = kIcpBlock / G;  // queries per pass
    int64_t u = (n + q - 1) / q;
    if (u < 1) u = 1;
    if (u > cap) u = cap;
    return (int32_t)u;
}

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

// The 28 FP32 products one correspondence adds: JTJ upper triangle (21), JTr (6), r^2.
__device__ __forceinline__ void term_products(const Terms &t, float v[28])
{
    int k = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = a; b < 6; ++b) v[k++] = t.Jg[a] * t.Jg[b] + t.Jc[a] * t.Jc[b];
#pragma unroll
    for (int a = 0; a < 6; ++a) v[21 + a] = t.Jg[a] * t.rg + t.Jc[a] * t.rc;
    v[27] = t.rg * t.rg + t.rc * t.rc;
}

template <int MODE>
__device__ __forceinline__ void acc_terms(Acc<MODE> &A, const Terms &t)
{
    float v[28];
    term_products(t, v);
#pragma unroll
    for (int k = 0; k < 28; ++k) A.add(k, v[k]);
}

// One sum of the accumulator kind MODE (the scalar form of Acc<MODE>).  An accumulator holding
// a single term is that term exactly, in every mode.
template <int MODE>
struct Sum1;

template <>
struct Sum1<GCA_ACCUM_COMPENSATED> {
    float s, c;
    __device__ void zero() { s = c = 0.f; }
    __device__ __forceinline__ void add(float x)
    {
        const float t = s + x;
        const float bp = t - s;
        c += (s - (t - bp)) + (x - bp);
        s = t;
    }
    __device__ __forceinline__ void merge(const Sum1 &o)
    {
        const float t = s + o.s;
        const float bp = t - s;
        const float e = (s - (t - bp)) + (o.s - bp);
        s = t;
        c = (c + o.c) + e;
    }
    __device__ double value() const { return (double)s + (double)c; }
};

template <>
struct Sum1<GCA_ACCUM_NAIVE_FP32> {
    float s;
    __device__ void zero() { s = 0.f; }
    __device__ __forceinline__ void add(float x) { s += x; }
    __device__ __forceinline__ void merge(const Sum1 &o) { s += o.s; }
    __device__ double value() const { return (double)s; }
};

template <>
struct Sum1<GCA_ACCUM_FP64> {
    double s;
    __device__ void zero() { s = 0.0; }
    __device__ __forceinline__ void add(float x) { s += (double)x; }
    __device__ __forceinline__ void merge(const Sum1 &o) { s += o.s; }
    __device__ double value() const { return s; }
};

// a + b in the precision of the accumulator kind: FP32 for the naive ablation (every sum of the
// reduction rounds to FP32, as a plain FP32 reduction would), FP64 otherwise.
template <int MODE>
__device__ __forceinline__ double acc_add(double a, double b)
{
    if constexpr (MODE == GCA_ACCUM_NAIVE_FP32) return (double)((float)a + (float)b);
    else return a + b;
}

__device__ __forceinline__ void xform_f32(const double *T, float4 p, float &x, float &y, float &z)
{
    const double a = p.x, b = p.y, c = p.z;
    x = (float)(((T[0] * a + T[1] * b) + T[2] * c) + T[3]);
    y = (float)(((T[4] * a + T[5] * b) + T[6] * c) + T[7]);
    z = (float)(((T[8] * a + T[9] * b) + T[10] * c) + T[11]);
}

// Device-scope coherent accesses: loads that cannot hit a stale cached copy, and stores that
// write through to the device coherence point (no L2 write-back of unrelated data needed).
__device__ __forceinline__ double load_coherent(const double *p)
{
    const unsigned long long u =
        __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)u);
}

__device__ __forceinline__ int32_t load_coherent(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_coherent(double *p, double v)
{
    __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_coherent(int32_t *p, int32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block reduction of the accumulator into FP64 partial `out[0..27]`, plus the correspondence
// count into out[28].  Only lanes whose thread index is below `span` and a multiple of `lead`
// hold data: each such wavefront merges them by butterflies over masks min(32, span/2) .. lead,
// then the wave results are merged in wave order.
template <int MODE>
__device__ __forceinline__ void block_partial(Acc<MODE> &A, int ncorr, int lead, int span, double *out)
{
    __shared__ Acc<MODE> wacc[kIcpBlock / kWave];
    __shared__ int wcnt[kIcpBlock / kWave];
    const int w = threadIdx.x / kWave;
    const int nw = (span + kWave - 1) / kWave;
    if (w < nw) {
        const int top = span >= kWave ? kWave / 2 : span / 2;
        for (int m = top; m >= lead; m >>= 1) A.merge_from_lane(m);
        for (int m = top; m >= lead; m >>= 1) ncorr += __shfl_xor(ncorr, m, kWave);
        if (lane_id() == 0) {
            wacc[w] = A;
            wcnt[w] = ncorr;
        }
    }
    __syncthreads();
    if (threadIdx.x < 28) {
        const int k = threadIdx.x;
        double v = 0.0;
        for (int i = 0; i < nw; ++i) v = acc_add<MODE>(v, wacc[i].value(k));
        out[k] = v;
    } else if (threadIdx.x == 28) {
        int c = 0;
        for (int i = 0; i < nw; ++i) c += wcnt[i];
        out[28] = (double)c;
    }
}

// Sum over the lanes of one pass of a work unit, where every lane j < Q holds at most one
// correspondence: rows[k][j] is value k of lane j (k < 28: the products of term_products, zero
// without a correspondence; k = 28: 1 or 0).  Thread (k = t % 32, chunk = t / 32) sums lanes
// [chunk Q / 8, (chunk + 1) Q / 8) of value k in lane order with the MODE accumulator; thread k
// then merges the 8 chunk sums in chunk order into out[k] (FP64).  Starts with the barrier that
// completes rows[]; out[k] is written by thread k (the caller's next barrier publishes it).
// Transposing through LDS costs one write and one read per value and lane, where butterflies
// over the 28 accumulators would take log2(Q) rounds of 28 (or 56) cross-lane moves per lane.
constexpr int kRowStride = kIcpBlock + 1;  // odd: a wavefront reading 32 rows hits 32 banks

template <int MODE>
__device__ __forceinline__ void pass_partial(const float (*rows)[kRowStride], int Q, double *out)
{
    __shared__ Sum1<MODE> chunk[8][32];
    const int k = (int)threadIdx.x & 31, ch = (int)threadIdx.x >> 5;
    __syncthreads();
    if (k < kNSys) {
        const int lo = (ch * Q) >> 3, hi = ((ch + 1) * Q) >> 3;
        Sum1<MODE> a;
        a.zero();
#pragma unroll 4
        for (int j = lo; j < hi; ++j) a.add(rows[k][j]);
        chunk[ch][k] = a;
    }
    __syncthreads();
    if (threadIdx.x < kNSys) {
        Sum1<MODE> a = chunk[0][threadIdx.x];
#pragma unroll
        for (int c = 1; c < 8; ++c) a.merge(chunk[c][threadIdx.x]);
        out[threadIdx.x] = a.value();
    }
}

// Fixed-order FP64 sum of `n` partials into sys[29] (in LDS): 8 interleaved groups, then the
// groups in order.  The loads go out in batches of kSumBatch per thread before any add (a batch
// costs one memory round trip), and padding a batch with +0.0 leaves every sum as the plain
// loop's.  The caller has made the partials visible (kernel boundary, or release / acquire).
// Ends with a barrier.
constexpr int kSumBatch = 32;

// naive: the GCA_ACCUM_NAIVE_FP32 ablation, whose sums across units are FP32 too (the
// correspondence count, entry 28, stays exact).
__device__ __forceinline__ void sum_partials(const double *partials, int n, double *sys, bool naive = false)
{
    __shared__ double part[8][kPartStride];
    const int v = threadIdx.x % kPartStride, grp = threadIdx.x / kPartStride;  // 8 groups of 32
    const bool f32 = naive && v < kNSys - 1;
    if (v < kNSys) {
        double a = 0.0;
        float af = 0.f;
        for (int b0 = grp; b0 < n; b0 += 8 * kSumBatch) {
            double x[kSumBatch];
#pragma unroll
            for (int u = 0; u < kSumBatch; ++u) {
                const int b = b0 + 8 * u;
                x[u] = b < n ? partials[(int64_t)b * kPartStride + v] : 0.0;
            }
            if (f32) {
#pragma unroll
                for (int u = 0; u < kSumBatch; ++u) af += (float)x[u];
            } else {
#pragma unroll
                for (int u = 0; u < kSumBatch; ++u) a += x[u];
            }
        }
        part[grp][v] = f32 ? (double)af : a;
    }
    __syncthreads();
    if (threadIdx.x < kNSys) {
        double a = 0.0;
        float af = 0.f;
        for (int g = 0; g < 8; ++g) {
            if (f32) af += (float)part[g][threadIdx.x];
            else a += part[g][threadIdx.x];
        }
        sys[threadIdx.x] = f32 ? (double)af : a;
    }
    __syncthreads();
}

// ---- nearest neighbour ----------------------------------------------------------------------
// Lanes g = 0..G-1 of a group search one query: lane g scans candidates g, g + G, ... of the
// flattened candidate space, then the group reduces by shuffles.  The result is the sorted slot
// of argmin (d2, original index) over target points with d2 < r^2 (FP64 d2), or -1.
//
// Warm start: `sp` is the slot matched in the previous iteration.  Its distance under the
// current pose bounds the answer, so only cells overlapping the box of half-width sqrt(d2_sp)
// around the query can hold it.  The slack kWarmRel / kWarmAbs covers FP64 rounding of d2 and
// of the cell arithmetic, so the box always contains every point with d2 <= d2_sp and the
// result is exactly that of the full search.  Ties in d2 compare original indices, which are
// loaded only then.
constexpr int kMaxRuns = 9;
constexpr int kUnroll = 4;
constexpr double kWarmRel = 1e-9, kWarmAbs = 1e-9;

__device__ __forceinline__ void nn_consider(const PairDesc &P, double d2, int32_t s, double r2, double &bd,
                                            int32_t &bs)
{
    if (!(d2 < r2)) return;
    if (d2 < bd) {
        bd = d2;
        bs = s;
    } else if (d2 == bd && s != bs && P.order[s] < P.order[bs]) {
        bs = s;
    }
}

__device__ __forceinline__ int32_t nn_group(const PairDesc &P, float qx, float qy, float qz, int g, int G,
                                            int32_t sp)
{
    const double r2 = (double)P.g.radius * (double)P.g.radius;
    double bd = DBL_MAX;
    int32_t bs = -1;
    double box = (double)P.g.radius;
    if (sp >= 0) {
        const float4 c = P.sxyzi[sp];
        const double d2 = dist2(c.x, c.y, c.z, qx, qy, qz);
        if (d2 < r2) {
            bd = d2;
            bs = sp;
            box = fmin(box, sqrt(d2) * (1.0 + kWarmRel) + kWarmAbs);
        }
    }
    int32_t lo[3], hi[3];
    if (grid_range_r(P.g, qx, qy, qz, box, lo, hi)) {
        const int nry = hi[1] - lo[1] + 1;
        const int nruns = (hi[0] - lo[0] + 1) * nry;
        if (nruns <= kMaxRuns) {
            // bounds of all (ix, iy) runs up front, then one flattened walk with kUnroll
            // independent loads in flight per lane
            uint32_t rs[kMaxRuns], cum[kMaxRuns + 1];
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
                cum[r + 1] = b > a ? b - a : 0u;  // a start table never decreases
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
                    nn_consider(P, dist2(c[u].x, c[u].y, c[u].z, qx, qy, qz), (int32_t)sl[u], r2, bd, bs);
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
                        nn_consider(P, dist2(c.x, c.y, c.z, qx, qy, qz), (int32_t)s, r2, bd, bs);
                    }
                }
        }
    }
    for (int m = 1; m < G; m <<= 1) {
        const double od = __shfl_xor(bd, m, kWave);
        const int32_t os = __shfl_xor(bs, m, kWave);
        if (od < bd) {
            bd = od;
            bs = os;
        } else if (od == bd && os != bs && os >= 0 && P.order[os] < P.order[bs]) {
            bs = os;
        }
    }
    return bs;
}

// ---- one work unit --------------------------------------------------------------------------
// With Q = 256 / G queries per pass, unit `vb` of a state covers the passes of source points
// [b * Q, b * Q + Q) for b = vb, vb + nvb, ...  Per pass:
//   1. nearest neighbours: G lanes per query; the group leader parks the matched slot in LDS;
//   2. lanes 0 .. Q-1 form the residuals, Jacobians and the 28 products of one query each and
//      park them in LDS; pass_partial sums them (MODE accumulator within the pass);
//   3. the pass's block partial is added, in pass order and FP64, to the unit's.
// No accumulator is live during the search, which keeps register use (and so the number of
// resident wavefronts hiding the search's load latency) in check.  Leaves the unit's partial in
// sys[0..28] (LDS; entry k written by thread k).
template <int MODE>
__device__ __forceinline__ void icp_unit(const PairDesc &P, int n, int G, int nvb, int ntgt, int vb, bool warm,
                                         const double *T, double *sys)
{
    __shared__ int32_t slot[kIcpBlock];
    __shared__ double bsys[kPartStride];
    __shared__ float rows[kNSys][kRowStride];
    const int lg = __builtin_ctz((unsigned)G);
    const int g = (int)threadIdx.x & (G - 1);
    const int Q = kIcpBlock >> lg;
    bool first = true;
    for (int b0 = vb * Q; b0 < n; b0 += nvb * Q) {
        {
            const int j = (int)threadIdx.x >> lg;
            const int i = b0 + j;
            float x = NAN, y = 0.f, z = 0.f;
            int32_t sp = -1;
            if (i < n) {
                if (warm) sp = P.nnprev[i];
                xform_f32(T, P.src[i], x, y, z);
            }
            // the warm-start slot is only a bound: any in-range value is safe, stale or not
            if ((uint32_t)sp >= (uint32_t)ntgt) sp = -1;
            const int32_t s = nn_group(P, x, y, z, g, G, sp);
            if (g == 0) slot[j] = s;
        }
        __syncthreads();
        const int i = b0 + (int)threadIdx.x;
        if ((int)threadIdx.x < Q) {
            float v[28];
#pragma unroll
            for (int k = 0; k < 28; ++k) v[k] = 0.f;
            float c = 0.f;
            if (i < n) {
                const int32_t s = slot[threadIdx.x];
                store_coherent(&P.nnprev[i], s);
                if (s >= 0) {
                    const float4 p = P.src[i];
                    float x, y, z;
                    xform_f32(T, p, x, y, z);
                    Terms t;
                    corr_terms(x, y, z, p.w, P.sxyzi[s], P.snrm[s], P.sgrad[s], P.sl, P.sc, t);
                    term_products(t, v);
                    c = 1.f;
                }
            }
#pragma unroll
            for (int k = 0; k < 28; ++k) rows[k][threadIdx.x] = v[k];
            rows[28][threadIdx.x] = c;
        }
        pass_partial<MODE>(rows, Q, first ? sys : bsys);
        if (!first && threadIdx.x < kNSys) sys[threadIdx.x] = acc_add<MODE>(sys[threadIdx.x], bsys[threadIdx.x]);
        first = false;
        __syncthreads();  // slot[] and the block-partial scratch are reused by the next pass
    }
    if (first && threadIdx.x < kNSys) sys[threadIdx.x] = 0.0;  // no point in this unit
}

// Element t of the Euler increment Rz(xi2) Ry(xi1) Rx(xi0) | xi3..5 (same expressions as
// euler_to_T).  trig = sin/cos of xi0, xi1, xi2.
__device__ __forceinline__ double euler_elem(int t, const double *trig, const double *xi)
{
    const double sa = trig[0], ca = trig[1], sb = trig[2], cb = trig[3], sc = trig[4], cc = trig[5];
    switch (t) {
    case 0: return cc * cb;
    case 1: return cc * sb * sa - sc * ca;
    case 2: return cc * sb * ca + sc * sa;
    case 3: return xi[3];
    case 4: return sc * cb;
    case 5: return sc * sb * sa + cc * ca;
    case 6: return sc * sb * ca - cc * sa;
    case 7: return xi[4];
    case 8: return -sb;
    case 9: return cb * sa;
    case 10: return cb * ca;
    case 11: return xi[5];
    case 15: return 1.0;
    default: return 0.0;
    }
}

// Epoch finish of state p (whole workgroup, run by the workgroup that took the last ticket):
// fixed-order FP64 sum of the unit partials, LDL^T solve, Euler increment, T <- dT T, the
// stopping rule of color_icp.cu:216-228, then the hand-over: the next epoch of this state, or
// the pose to the next level of the chain.  A failed solve takes the identity step (which also
// satisfies the stopping rule), as the reference does.
__device__ __noinline__ void icp_finish(IcpState *states, int p, uint32_t epoch, const double *partials, int nvb, double *sys,
                                        bool naive)
{
    IcpState *st = states + p;
    __shared__ double Tc[16], Tn[16], dT[16], xi[6], trig[6];
    __shared__ double f_crot, f_ct2;
    __shared__ int f_iter, f_max, f_fail, f_next, f_ok, f_done;
    const int t = threadIdx.x;
    // state fields load alongside the partial sums (coherent: the previous epoch's finisher may
    // have run on another XCD)
    if (t >= 64 && t < 80) {
        Tc[t - 64] = load_coherent(&st->T[t - 64]);
    } else if (t == 80) {
        f_iter = load_coherent(&st->iter);
        f_max = load_coherent(&st->max_iter);
        f_fail = load_coherent(&st->solver_failures);
        f_next = load_coherent(&st->next);
        f_crot = load_coherent(&st->conv_rot);
        f_ct2 = load_coherent(&st->conv_t2);
    }
    sum_partials(partials, nvb, sys, naive);  // after the acquire in unit_commit
    if (t == 0) f_ok = solve6(sys, sys + 21, xi) ? 1 : 0;
    __syncthreads();
    if (t < 3) {
        double s, c;
        sincos(xi[t], &s, &c);
        trig[2 * t] = s;
        trig[2 * t + 1] = c;
    }
    __syncthreads();
    if (t < 16) dT[t] = euler_elem(t, trig, xi);
    __syncthreads();
    if (t < 16) {
        const int i = t >> 2, j = t & 3;
        const double v =
            ((dT[4 * i] * Tc[j] + dT[4 * i + 1] * Tc[4 + j]) + dT[4 * i + 2] * Tc[8 + j]) + dT[4 * i + 3] * Tc[12 + j];
        Tn[t] = v;
        store_coherent(&st->T[t], v);
    } else if (t >= 32 && t < 32 + kNSys) {
        st->sys[t - 32] = sys[t - 32];
    } else if (t == 64) {
        const int it = f_iter + 1;
        const double cos_t = 0.5 * (dT[0] + dT[5] + dT[10] - 1.0);
        const double tt = (dT[3] * dT[3] + dT[7] * dT[7]) + dT[11] * dT[11];
        const bool conv = cos_t >= 1.0 - f_crot && tt <= f_ct2;
        const bool done = conv || it >= f_max;
        store_coherent(&st->iter, it);
        if (!f_ok) store_coherent(&st->solver_failures, f_fail + 1);
        if (conv) st->converged = 1;
        if (done) {
            st->done = 1;
            st->t_done = wall_clock64();
        }
        __hip_atomic_store(&st->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        f_done = done ? 1 : 0;
    }
    __syncthreads();
    if (f_done && f_next >= 0 && t < 16) store_coherent(&states[f_next].T[t], Tn[t]);
    // Release before the queue word opens the next epoch or level: readers on other XCDs load
    // the pose with coherent loads once they see the word.  (A vmcnt wait after the stores is
    // not enough: it does not order them before the word for readers on other XCDs.)
    fence_release_agent();
    __syncthreads();
    if (t < kSubQ) {
        if (!f_done) {
            __hip_atomic_store(&st->q[t].w, (unsigned long long)(epoch + 1) << 32, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&st->q[t].w, (unsigned long long)kEpochDone << 32, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            if (f_next >= 0)
                __hip_atomic_store(&states[f_next].q[t].w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Unit partial -> global, ticket, and the epoch finish by the workgroup completing the epoch.
// The ticket increment is a release (threads 0..28, which store the partial, and the
// ticket-taking thread 0 are all in wavefront 0) and the finisher acquires after the last
// ticket: the canonical hand-over, which makes every unit's partial visible to the finisher on
// whichever XCD it runs.  The partial and the matched slots are written through, so the
// release has little left to write back.
template <int MODE>
__device__ __forceinline__ void unit_commit(const PairDesc &P, IcpState *states, int p, int vb, int nvb,
                                            uint32_t epoch, double *sys)
{
    __shared__ bool last;
    if (threadIdx.x < kNSys) store_coherent(&P.partials[(int64_t)vb * kPartStride + threadIdx.x], sys[threadIdx.x]);
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(&states[p].ticket, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) ==
               (uint32_t)(nvb - 1);
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    icp_finish(states, p, epoch, P.partials, nvb, sys, MODE == GCA_ACCUM_NAIVE_FP32);
}

// ---- persistent launch: every iteration of every state --------------------------------------
// Scheduling loop of one workgroup: read the claim words (one coherent load per state and
// sub-queue, a window at a time), pick the first state with unclaimed units in round-robin order
// from the state last served and, in it, the sub-queue nearest to the workgroup's own, claim a
// unit with one atomic, run it.  A workgroup only waits (sleep + re-poll)
// when every unit of every open epoch is claimed, i.e. held by running workgroups, so progress
// never depends on how many workgroups are resident.  Exits when every state is done.
// (Claiming first and scanning only on failure was measured slower: every workgroup that ends
// a unit after the epoch is drained adds a failing atomic on a hot claim word.)
// 4 wavefronts per SIMD (<= 128 VGPRs): resident wavefronts are what hides the latency of the
// neighbour search's dependent loads.
#define GCA_ICP_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4, 8)))
constexpr int kScanWindow = 64;  // claim words read per scan round (one wavefront's loads)

// refresh (launches over one state only): the state's sizes are taken from its clouds' headers
// at the start, by every workgroup for itself (workgroup 0 also stores them in the state), for a
// launch enqueued before its source cloud's size was final.
template <int MODE>
__global__ __launch_bounds__(kIcpBlock) GCA_ICP_OCCUPANCY void k_icp_persist(const PairDesc *__restrict__ pairs, IcpState *states,
                                                           int s_first, int n_states, int refresh, uint32_t *err)
{
    __shared__ double T[12];
    __shared__ double sys[kPartStride];
    __shared__ int s_pick, s_alive, s_vb;
    __shared__ uint32_t s_epoch, s_pick_epoch;
    __shared__ int s_nsrc, s_ntgt, s_G, s_nvb;
    if (refresh) {
        if (threadIdx.x == 0) {
            const PairDesc &P = pairs[s_first];
            const int ns = P.src_hdr->n, nt = P.tgt_hdr->n;
            const int G = icp_lanes_for(P.lanes, ns);
            const int nvb = icp_units_for(ns, G, P.unit_cap);
            s_nsrc = ns;
            s_ntgt = nt;
            s_G = G;
            s_nvb = nvb;
            if (blockIdx.x == 0) {
                IcpState &S = states[s_first];
                S.n_src = ns;
                S.n_tgt = nt;
                S.G = G;
                S.nvb = nvb;
            }
        }
        __syncthreads();
    }
    // A launch over one state needs no more workgroups than the state has units per iteration:
    // the others could only poll (and their polling competes with whatever else runs).
    if (n_states == 1 && (int)blockIdx.x >= max(refresh ? s_nvb : states[s_first].nvb, 1)) return;
    // Highest issue priority among the wavefronts sharing a SIMD: a registration iteration is a
    // chain of dependent memory round trips, and in a frame pipeline it runs beside the kernels
    // that build the finer levels, whose wavefronts would otherwise be issued first (measured:
    // the coarsest level's 4 iterations beside those builds, 125 -> 87 us).
    __builtin_amdgcn_s_setprio(3);
    // serves states [s_first, s_first + n_states); `home` (the state last served) is relative
    int home = (int)(blockIdx.x % (unsigned)n_states);
    const int myq = (int)(blockIdx.x % (unsigned)kSubQ);  // workgroups are dealt to XCDs round-robin
    uint32_t idle = 0;
    const int n_words = n_states * kSubQ;
    for (;;) {
        // The claim words are read a window of kScanWindow words at a time, from the state last
        // served on; the next window is read only when this one has no unclaimed unit.  (Reading
        // every word for every claim put so much polling traffic on a few hot lines that it
        // bounded a launch with many states: 64 words per round run the 64-pair batch 17 %
        // faster than all 1536, 256 words 12 % faster.)
        int my_rot = INT_MAX;
        uint32_t my_epoch = 0;
        int pick = INT_MAX, alive = 0;
        for (int w0 = 0; w0 < n_words; w0 += kScanWindow) {
            if (threadIdx.x == 0) {
                s_pick = INT_MAX;
                s_alive = 0;
            }
            __syncthreads();
            const int r = w0 + (int)threadIdx.x;  // rotation index: state (home + r / kSubQ), slot r % kSubQ
            if ((int)threadIdx.x < kScanWindow && r < n_words) {
                const int p = s_first + (home + r / kSubQ) % n_states, x = (myq + r % kSubQ) % kSubQ;
                const unsigned long long w =
                    __hip_atomic_load(&states[p].q[x].w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t e = (uint32_t)(w >> 32), c = (uint32_t)w;
                if (e != kEpochDone) {
                    s_alive = 1;
                    // first state in round-robin order from the last one served, then the
                    // sub-queue nearest to this workgroup's own
                    if (e != kEpochWait && c < subq_units(refresh ? s_nvb : states[p].nvb, x)) {
                        my_rot = r;
                        my_epoch = e;
                        atomicMin(&s_pick, r);
                    }
                }
            }
            __syncthreads();
            pick = s_pick;
            alive |= s_alive;
            __syncthreads();
            if (pick != INT_MAX) break;
        }
        if (my_rot == pick) s_pick_epoch = my_epoch;  // the one thread that saw the picked state
        __syncthreads();
        if (!alive) break;
        if (pick == INT_MAX) {
            if (++idle > kSpinLimit) {
                if (threadIdx.x == 0) atomicOr(err, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        const int p = __builtin_amdgcn_readfirstlane(s_first + (home + pick / kSubQ) % n_states);
        const int x = __builtin_amdgcn_readfirstlane((myq + pick % kSubQ) % kSubQ);
        IcpState *st = states + p;
        // The claim (wavefront 0) and the pose load (wavefront 1) go out together.  The scan saw
        // epoch e open, so the pose read now is epoch e's or later; if the claim lands in epoch
        // e, epoch e cannot finish (and write a later pose) before this unit does, so it is
        // exactly epoch e's.  A claim in a later epoch reloads the pose after the claim.  The pose
        // is the only data a unit reads that this launch writes (besides the matched slots,
        // which only bound the search), and its loads are coherent ones behind the finisher's
        // release; an acquire fence per claim would also invalidate the target data cached in
        // L2, which costs about a third of the ICP time.
        if (threadIdx.x == 0) {
            const unsigned long long old =
                __hip_atomic_fetch_add(&st->q[x].w, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t e = (uint32_t)(old >> 32), c = (uint32_t)old;
            const bool ok = e < kEpochWait && c < subq_units(refresh ? s_nvb : st->nvb, x);
            s_vb = ok ? subq_unit(refresh ? s_nvb : st->nvb, x, (int)c) : -1;
            s_epoch = e;
            if (ok && e == 0 && c == 0 && x == 0) st->t_start = wall_clock64();  // unit 0 of the first epoch
        } else if (threadIdx.x >= kWave && threadIdx.x < kWave + 12) {
            T[threadIdx.x - kWave] = load_coherent(&st->T[threadIdx.x - kWave]);
        }
        __syncthreads();
        const int vb = s_vb;
        const uint32_t epoch = s_epoch;
        if (vb < 0) continue;  // lost the race for the last unit: poll again
        if (epoch != s_pick_epoch) {
            if (threadIdx.x < 12) T[threadIdx.x] = load_coherent(&st->T[threadIdx.x]);
            __syncthreads();
        }
        idle = 0;
        home = p - s_first;
        const PairDesc &P = pairs[p];
        const int nvb = refresh ? s_nvb : st->nvb;
        if (refresh)
            icp_unit<MODE>(P, s_nsrc, s_G, nvb, s_ntgt, vb, epoch > 0, T, sys);
        else
            icp_unit<MODE>(P, st->n_src, st->G, nvb, st->n_tgt, vb, epoch > 0, T, sys);
        unit_commit<MODE>(P, states, p, vb, nvb, epoch, sys);
    }
}

// ---- one launch per iteration (same units, same finish) -------------------------------------
// Workgroup b runs unit b - unit_first[s] of state s = unit_state[b]; states that are done exit.
template <int MODE>
__global__ __launch_bounds__(kIcpBlock) GCA_ICP_OCCUPANCY void k_icp_iter(const PairDesc *__restrict__ pairs, IcpState *states,
                                                        const int32_t *__restrict__ unit_state,
                                                        const int32_t *__restrict__ unit_first)
{
    __shared__ double T[12];
    __shared__ double sys[kPartStride];
    const int p = __builtin_amdgcn_readfirstlane(unit_state[blockIdx.x]);
    IcpState *st = states + p;
    if (st->done) return;
    const int vb = (int)blockIdx.x - unit_first[p];
    const uint32_t epoch = (uint32_t)st->iter;
    if (threadIdx.x < 12) T[threadIdx.x] = st->T[threadIdx.x];
    if (threadIdx.x == 0 && vb == 0 && epoch == 0) st->t_start = wall_clock64();
    __syncthreads();
    const PairDesc &P = pairs[p];
    const int nvb = st->nvb;
    icp_unit<MODE>(P, st->n_src, st->G, nvb, st->n_tgt, vb, epoch > 0, T, sys);
    unit_commit<MODE>(P, states, p, vb, nvb, epoch, sys);
}

// ---- resident launch: one state, every iteration, no scheduler ---------------------------------
// Workgroup b < nvb runs unit b of every iteration.  After its unit it publishes the partial
// (write-through) and arrives at the iteration's barrier: a release increment of its group's
// counter (group b % kSubQ, each counter on its own line), and the arrival that completes a group
// increments the top counter.  Every workgroup then waits for the top counter, reads all nvb
// partials with coherent loads, sums them in the finisher's fixed order (sum_partials), solves and
// updates the pose itself: every workgroup computes the same bits, so no pose is handed over
// inside the launch, and the result is bit-identical to the persistent and per-iteration launches
// of the same shape.  Partials alternate between two banks by iteration parity: a bank is
// overwritten only after every workgroup has passed the barrier of the iteration after the one
// that read it.  The workgroups keep their source points and previous matches from one
// iteration to the next only through memory they alone touch (the matched slots).
// Progress needs all nvb workgroups resident at once, so the launch is used only when nvb is at
// most two per CU (the occupancy allows four); kernels running beside it only delay it, as their
// workgroups end and free the slots the rest need.  A wait beyond the spin limit flags the error
// word instead of hanging.
__device__ __forceinline__ uint32_t load_coherent_u32(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sum_partials over coherent loads (partials written in this launch by other workgroups, read
// without an acquire so that the target data cached in L2 survives).  Same order and arithmetic.
__device__ __forceinline__ void sum_partials_coherent(const double *partials, int n, double *sys, bool naive)
{
    __shared__ double part[8][kPartStride];
    const int v = threadIdx.x % kPartStride, grp = threadIdx.x / kPartStride;
    const bool f32 = naive && v < kNSys - 1;
    if (v < kNSys) {
        double a = 0.0;
        float af = 0.f;
        for (int b0 = grp; b0 < n; b0 += 8 * kSumBatch) {
            double x[kSumBatch];
#pragma unroll
            for (int u = 0; u < kSumBatch; ++u) {
                const int b = b0 + 8 * u;
                x[u] = b < n ? load_coherent(partials + (int64_t)b * kPartStride + v) : 0.0;
            }
            if (f32) {
#pragma unroll
                for (int u = 0; u < kSumBatch; ++u) af += (float)x[u];
            } else {
#pragma unroll
                for (int u = 0; u < kSumBatch; ++u) a += x[u];
            }
        }
        part[grp][v] = f32 ? (double)af : a;
    }
    __syncthreads();
    if (threadIdx.x < kNSys) {
        double a = 0.0;
        float af = 0.f;
        for (int g = 0; g < 8; ++g) {
            if (f32) af += (float)part[g][threadIdx.x];
            else a += part[g][threadIdx.x];
        }
        sys[threadIdx.x] = f32 ? (double)af : a;
    }
    __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(kIcpBlock) GCA_ICP_OCCUPANCY void k_icp_resident(const PairDesc *__restrict__ pairs,
                                                                               IcpState *states, int s, int refresh,
                                                                               uint32_t *err)
{
    __shared__ double Tc[16], Tn[16], dT[16], xi[6], trig[6];
    __shared__ double sys[kPartStride];
    __shared__ int f_ok, f_done, f_stop;
    __shared__ double f_crot, f_ct2;
    __shared__ int f_max, f_next, f_iter0, f_fail0;
    __shared__ int s_nsrc, s_ntgt, s_G, s_nvb;
    IcpState *st = states + s;
    const PairDesc &P = pairs[s];
    const int t = threadIdx.x;
    if (t == 0) {
        int ns, nt, G, nvb;
        if (refresh) {
            ns = P.src_hdr->n;
            nt = P.tgt_hdr->n;
            G = icp_lanes_for(P.lanes, ns);
            nvb = icp_units_for(ns, G, P.unit_cap);
        } else {
            ns = st->n_src;
            nt = st->n_tgt;
            G = st->G;
            nvb = st->nvb;
        }
        s_nsrc = ns;
        s_ntgt = nt;
        s_G = G;
        s_nvb = nvb;
        f_max = st->max_iter;
        f_next = st->next;
        f_iter0 = st->iter;
        f_fail0 = st->solver_failures;
        f_crot = st->conv_rot;
        f_ct2 = st->conv_t2;
        f_stop = st->done;
        if (refresh && blockIdx.x == 0) {
            st->n_src = ns;
            st->n_tgt = nt;
            st->G = G;
            st->nvb = nvb;
        }
    } else if (t >= 64 && t < 80) {
        Tc[t - 64] = st->T[t - 64];
    }
    __syncthreads();
    const int nvb = s_nvb;
    if ((int)blockIdx.x >= nvb || f_stop) return;
    // runs of 4 units (neighbouring source points) per XCD: the runs' target reads share an L2
    // (the partition, and so the sums, do not change)
    const int vb = xcd_runs<4>((int)blockIdx.x, nvb);
    if (vb == 0 && t == 0) st->t_start = wall_clock64();
    __builtin_amdgcn_s_setprio(3);  // a chain of dependent round trips; see k_icp_persist
    const int grp = vb % kSubQ;
    const uint64_t grp_size = subq_units(nvb, grp);
    const uint32_t n_groups = (uint32_t)min(nvb, kSubQ);
    int iter = f_iter0, fails = f_fail0;
    for (uint32_t e = 0;; ++e) {
        double *bank = P.partials + (size_t)(e & 1u) * (size_t)P.unit_cap * kPartStride;
        icp_unit<MODE>(P, s_nsrc, s_G, nvb, s_ntgt, vb, iter > 0, Tc, sys);
        if (t < kNSys) store_coherent(&bank[(int64_t)vb * kPartStride + t], sys[t]);
        if (t == 0) {
            const unsigned long long old =
                __hip_atomic_fetch_add(&st->bar[grp].w, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1 == grp_size * (e + 1))
                __hip_atomic_fetch_add(&st->bar_top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t want = n_groups * (e + 1);
            uint32_t spins = 0;
            while (load_coherent_u32(&st->bar_top) < want) {
                if (++spins > kSpinLimit) {
                    atomicOr(err, 2u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            f_stop = spins > kSpinLimit ? 1 : 0;
        }
        __syncthreads();
        if (f_stop) return;
        sum_partials_coherent(bank, nvb, sys, MODE == GCA_ACCUM_NAIVE_FP32);
        if (t == 0) f_ok = solve6(sys, sys + 21, xi) ? 1 : 0;
        __syncthreads();
        if (t < 3) {
            double sn, cs;
            sincos(xi[t], &sn, &cs);
            trig[2 * t] = sn;
            trig[2 * t + 1] = cs;
        }
        __syncthreads();
        if (t < 16) dT[t] = euler_elem(t, trig, xi);
        __syncthreads();
        if (t < 16) {
            const int i = t >> 2, j = t & 3;
            Tn[t] = ((dT[4 * i] * Tc[j] + dT[4 * i + 1] * Tc[4 + j]) + dT[4 * i + 2] * Tc[8 + j]) + dT[4 * i + 3] * Tc[12 + j];
        } else if (t == 64) {
            const double cos_t = 0.5 * (dT[0] + dT[5] + dT[10] - 1.0);
            const double tt = (dT[3] * dT[3] + dT[7] * dT[7]) + dT[11] * dT[11];
            const bool conv = cos_t >= 1.0 - f_crot && tt <= f_ct2;
            f_done = (conv || iter + 1 >= f_max) ? (conv ? 2 : 1) : 0;
        }
        __syncthreads();
        ++iter;
        if (!f_ok) ++fails;
        if (t < 16) Tc[t] = Tn[t];
        const int done = f_done;
        if (done && vb == 0) {
            if (t < 16) {
                st->T[t] = Tn[t];
                if (f_next >= 0) states[f_next].T[t] = Tn[t];
            } else if (t >= 32 && t < 32 + kNSys) {
                st->sys[t - 32] = sys[t - 32];
            } else if (t == 64) {
                st->iter = iter;
                st->solver_failures = fails;
                st->converged = done == 2 ? 1 : 0;
                st->done = 1;
                st->t_done = wall_clock64();
            }
        }
        __syncthreads();  // Tc is read by the next unit; f_done is rewritten by the next finish
        if (done) return;
    }
}

// Gauss-Newton system for given correspondences (no solve).  ticket: zero at the launch, left at
// zero; out: the 29 sums (FP64).
template <int MODE>
__global__ __launch_bounds__(kIcpBlock) void k_build_gn(PairDesc P, const float4 *__restrict__ tgt_xyzi,
                                                        const int32_t *__restrict__ nn, double t00, double t01,
                                                        double t02, double t03, double t10, double t11,
                                                        double t12, double t13, double t20, double t21,
                                                        double t22, double t23, float sl, float sc,
                                                        uint32_t *ticket, double *out)
{
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
    block_partial<MODE>(A, ncorr, 1, kIcpBlock, sys);
    if (threadIdx.x < kNSys) P.partials[(int64_t)blockIdx.x * kPartStride + threadIdx.x] = sys[threadIdx.x];
    fence_release_agent();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == (uint32_t)(P.nblk - 1);
    __syncthreads();
    if (!last) return;
    fence_acquire_agent();
    sum_partials(P.partials, P.nblk, sys, MODE == GCA_ACCUM_NAIVE_FP32);
    if (threadIdx.x < kNSys) out[threadIdx.x] = sys[threadIdx.x];
    if (threadIdx.x == 0) *ticket = 0u;
}

__global__ void k_ticket_clear(uint32_t *ticket) { *ticket = 0u; }

// k_build_gn over given partials (nblk workgroups' worth), ticket and output, on ctx's stream.
static void launch_build_gn(gca_ctx ctx, const PairDesc &P, const float4 *tgt_xyzi, const int32_t *nn, const double *T,
                            float lambda, int32_t mode, uint32_t *ticket, double *out)
{
    const float sl = sqrtf(lambda), sc = sqrtf(1.0f - lambda);
    const dim3 g((unsigned)P.nblk), b(kIcpBlock);
    if (mode == GCA_ACCUM_NAIVE_FP32)
        hipLaunchKernelGGL(k_build_gn<GCA_ACCUM_NAIVE_FP32>, g, b, 0, ctx->stream, P, tgt_xyzi, nn, T[0], T[1], T[2],
                           T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11], sl, sc, ticket, out);
    else if (mode == GCA_ACCUM_FP64)
        hipLaunchKernelGGL(k_build_gn<GCA_ACCUM_FP64>, g, b, 0, ctx->stream, P, tgt_xyzi, nn, T[0], T[1], T[2], T[3],
                           T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11], sl, sc, ticket, out);
    else
        hipLaunchKernelGGL(k_build_gn<GCA_ACCUM_COMPENSATED>, g, b, 0, ctx->stream, P, tgt_xyzi, nn, T[0], T[1], T[2],
                           T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11], sl, sc, ticket, out);
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

// ---------------------------------------------------------------------------------------------
// state set-up

struct StateInit {
    double conv_rot, conv_t2;
    int32_t max_iter;
    int32_t next;  // next level of the chain (-1)
    int32_t open;  // 1 = first level of its chain (runs from the start)
    int32_t t0;    // row of the initial pose in the T0 table
};

// Sizes come from the clouds' device headers, so set-up needs no host round trip.
__global__ void k_state_init(IcpState *states, const PairDesc *__restrict__ pairs, const StateInit *__restrict__ init,
                             const double *__restrict__ T0, int n, uint32_t *err)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) *err = 0u;
    if (p >= n) return;
    IcpState &s = states[p];
    const PairDesc &P = pairs[p];
    const StateInit &I = init[p];
    for (int k = 0; k < 16; ++k) s.T[k] = T0[16 * I.t0 + k];
    for (int k = 0; k < kNSys; ++k) s.sys[k] = 0.0;
    s.conv_rot = I.conv_rot;
    s.conv_t2 = I.conv_t2;
    s.iter = 0;
    s.max_iter = I.max_iter;
    s.converged = 0;
    s.solver_failures = 0;
    s.n_src = P.src_hdr->n;
    s.n_tgt = P.tgt_hdr->n;
    s.G = icp_lanes_for(P.lanes, s.n_src);
    s.nvb = icp_units_for(s.n_src, s.G, P.unit_cap);
    s.next = I.next;
    s.done = I.max_iter <= 0 ? 1 : 0;
    s.t_init = wall_clock64();
    s.t_done = s.t_init;
    s.t_start = 0;
    const unsigned long long w0 = I.max_iter <= 0 ? ((unsigned long long)kEpochDone << 32)
                                                  : (I.open ? 0ull : ((unsigned long long)kEpochWait << 32));
    for (int x = 0; x < kSubQ; ++x) {
        s.q[x].w = w0;
        s.bar[x].w = 0ull;
    }
    s.ticket = 0u;
    s.bar_top = 0u;
}

// State of the single-pair kernels (Gauss-Newton build, residuals, objective).
__global__ void k_state_simple(IcpState *st, const double *__restrict__ T0)
{
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 16; ++k) st->T[k] = T0[k];
    for (int k = 0; k < kNSys; ++k) st->sys[k] = 0.0;
    st->iter = 0;
    st->done = 0;
    st->ticket = 0u;
}

// ---------------------------------------------------------------------------------------------
// host side

static int32_t icp_blocks(int64_t n) { return (int32_t)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kIcpBlock - 1) / kIcpBlock)); }

// Lanes per correspondence query by source size.  A call with one pair is latency-bound (the
// GPU is wider than a small cloud's queries): wide lane groups shorten each query's dependent
// load chain.  A call with many pairs is throughput-bound: one lane per query wastes no lane on
// short candidate lists.  Defaults: the autotuner's choice on MI355X for the 640x480 frame
// pipeline (levels of ~5k, ~20k and ~74k points) and for 64 uneven pairs; a ctx's tuning
// overrides them.
static void icp_shape(gca_ctx ctx, int n_pairs, int32_t lanes[3], int32_t *unit_cap)
{
    static const int32_t def[3] = {8, 4, 1};
    for (int k = 0; k < 3; ++k) {
        if (n_pairs > 1) lanes[k] = ctx->tune[GCA_TUNE_ICP_LANES_BATCH] ? ctx->tune[GCA_TUNE_ICP_LANES_BATCH] : 1;
        else lanes[k] = ctx->tune[GCA_TUNE_ICP_LANES_SMALL + k] ? ctx->tune[GCA_TUNE_ICP_LANES_SMALL + k] : def[k];
    }
    // units per iteration: one pair needs them to spread over the device, many pairs already
    // do, and fewer units mean fewer partials to sum per iteration
    const int32_t def_units = n_pairs > 1 ? 128 : 512;
    *unit_cap = std::min(kMaxUnits, ctx->tune[GCA_TUNE_ICP_MAX_UNITS] ? ctx->tune[GCA_TUNE_ICP_MAX_UNITS] : def_units);
}

struct IcpWork {
    IcpState *states = nullptr;
    double *partials = nullptr;
    int32_t *nnprev = nullptr;
    PairDesc *pairs = nullptr;
    StateInit *init = nullptr;
    double *T0 = nullptr;
    int32_t *unit_state = nullptr;
    int32_t *unit_first = nullptr;
    uint32_t *err = nullptr;
};

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static int32_t icp_work(gca_ctx ctx, int n_states, int n_t0, int64_t partial_slots, int64_t nnprev_n,
                        int64_t table_units, IcpWork &w)
{
    w.states = (IcpState *)ctx_scratch(ctx, 13, sizeof(IcpState) * n_states, false);
    w.partials = (double *)ctx_scratch(ctx, 14, sizeof(double) * kPartStride * std::max<int64_t>(partial_slots, 1), false);
    w.nnprev = (int32_t *)ctx_scratch(ctx, 16, sizeof(int32_t) * std::max<int64_t>(nnprev_n, 1), false);
    if (!w.states || !w.partials || !w.nnprev) return GCA_ERR_OUT_OF_MEMORY;
    const size_t b_pairs = align256(sizeof(PairDesc) * n_states), b_init = align256(sizeof(StateInit) * n_states),
                 b_t0 = align256(sizeof(double) * 16 * std::max(n_t0, 4)),
                 b_us = align256(sizeof(int32_t) * std::max<int64_t>(table_units, 1)),
                 b_uf = align256(sizeof(int32_t) * n_states), b_err = 256;
    const size_t need = b_pairs + b_init + b_t0 + b_us + b_uf + b_err;
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
    base += b_pairs;
    w.init = (StateInit *)base;
    base += b_init;
    w.T0 = (double *)base;
    base += b_t0;
    w.unit_state = (int32_t *)base;
    base += b_us;
    w.unit_first = (int32_t *)base;
    base += b_uf;
    w.err = (uint32_t *)base;
    return GCA_OK;
}

static PairDesc make_pair(gca_cloud src, gca_cloud tgt, double *partials, int32_t *nnprev, float lambda,
                          float radius)
{
    PairDesc P{};
    P.src = src->xyzi;
    P.src_hdr = src->hdr;
    P.tgt_hdr = tgt->hdr;
    P.g = tgt->grid.desc;
    P.g.radius = radius;
    P.start = tgt->grid.start;
    P.order = tgt->grid.order;
    P.sxyzi = tgt->grid.sxyzi;
    P.nrm = tgt->nrm;
    P.grad = tgt->grad;
    P.snrm = tgt->grid.snrm;
    P.sgrad = tgt->grid.sgrad;
    P.partials = partials;
    P.nnprev = nnprev;
    P.sl = sqrtf(lambda);
    P.sc = sqrtf(1.0f - lambda);
    return P;
}

// keep_grid: any valid grid with cell >= corr is used as it is (the neighbour results do not
// depend on it); otherwise a grid too coarse for corr is rebuilt for speed.
static int32_t prepare_target(gca_ctx ctx, gca_cloud tgt, float corr, float grad_r, bool keep_grid)
{
    if (!tgt->has_normals) return GCA_ERR_NOT_READY;
    if (!tgt->has_gradient) {
        if (!(grad_r > 0.0f)) return GCA_ERR_NOT_READY;
        GCA_TRY(gca_compute_color_gradient(ctx, tgt, grad_r, 0));
    }
    if (keep_grid && tgt->grid.valid && tgt->grid.desc.cell >= corr) return GCA_OK;
    return grid_build(ctx, tgt, corr);
}

static int32_t validate_params(const gca_icp_params *p)
{
    if (!p) return GCA_ERR_INVALID_VALUE;
    if (!(p->max_correspondence_distance > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!(p->lambda_geometric >= 0.0f) || p->lambda_geometric > 1.0f) return GCA_ERR_INVALID_VALUE;  // .cu:438-441
    if (p->max_iterations < 0 || p->accumulate_mode < 0 || p->accumulate_mode > 2) return GCA_ERR_INVALID_VALUE;
    return GCA_OK;
}

template <int MODE>
static void launch_iter(gca_ctx ctx, const IcpWork &w, int64_t off, int64_t units)
{
    hipLaunchKernelGGL(k_icp_iter<MODE>, dim3((unsigned)units), dim3(kIcpBlock), 0, ctx->stream, w.pairs, w.states,
                       w.unit_state + off, w.unit_first);
}

// ---- registration states: shared set-up ----------------------------------------------------
// State k * n_levels + l = pair k at level l (coarse to fine); level l + 1 of a pair starts from
// the pose level l ends with; levels with max_iterations == 0 are skipped.

// Validation, target preparation (gradient if absent, grid, grid-ordered attributes) and the
// launch shape.  Every target's grid is built before any descriptor is made: a later build for
// another state of the same target may replace arrays an earlier one set up.
static int32_t states_prepare(gca_ctx ctx, int n_pairs, int n_levels, const gca_cloud *srcs, const gca_cloud *tgts,
                              const gca_icp_params *params, const float *grad_r, bool keep_grid, int32_t lanes[3],
                              int32_t *unit_cap, std::vector<int64_t> &nn_off)
{
    const int n_states = n_pairs * n_levels;
    for (int l = 0; l < n_levels; ++l) {
        GCA_TRY(validate_params(&params[l]));
        if (params[l].accumulate_mode != params[0].accumulate_mode) return GCA_ERR_INVALID_VALUE;
    }
    for (int s = 0; s < n_states; ++s)
        if (!srcs[s] || !tgts[s] || srcs[s]->ctx != ctx || tgts[s]->ctx != ctx) return GCA_ERR_INVALID_VALUE;
    for (int s = 0; s < n_states; ++s)
        GCA_TRY(prepare_target(ctx, tgts[s], params[s % n_levels].max_correspondence_distance, grad_r[s % n_levels],
                               keep_grid));
    for (int s = 0; s < n_states; ++s) GCA_TRY(grid_gather_attributes(ctx, tgts[s]));
    icp_shape(ctx, n_pairs, lanes, unit_cap);
    nn_off.assign(n_states + 1, 0);
    for (int s = 0; s < n_states; ++s) nn_off[s + 1] = nn_off[s] + srcs[s]->cap;
    return GCA_OK;
}

// Descriptors and chain links of the states over the given partial / previous-match buffers.
static void states_describe(int n_pairs, int n_levels, const gca_cloud *srcs, const gca_cloud *tgts,
                            const gca_icp_params *params, const int32_t lanes[3], int32_t unit_cap, double *partials,
                            int32_t *nnprev, const std::vector<int64_t> &nn_off, std::vector<PairDesc> &P,
                            std::vector<StateInit> &I, std::vector<int32_t> &tail)
{
    const int n_states = n_pairs * n_levels;
    P.assign(n_states, PairDesc{});
    I.assign(n_states, StateInit{});
    tail.assign(n_pairs, -1);  // last state of each chain
    for (int k = 0; k < n_pairs; ++k) {
        int prev = -1;
        for (int l = 0; l < n_levels; ++l) {
            const int s = k * n_levels + l;
            const gca_icp_params &pl = params[l];
            P[s] = make_pair(srcs[s], tgts[s], partials + (int64_t)s * 2 * unit_cap * kPartStride, nnprev + nn_off[s],
                             pl.lambda_geometric, pl.max_correspondence_distance);
            memcpy(P[s].lanes, lanes, sizeof(int32_t) * 3);
            P[s].unit_cap = unit_cap;
            I[s] = StateInit{pl.conv_rotation, pl.conv_translation_sq, pl.max_iterations, -1, 0, k};
            if (pl.max_iterations > 0) {
                if (prev < 0) I[s].open = 1;
                else I[prev].next = s;
                prev = s;
            }
        }
        tail[k] = prev;
    }
}

template <int MODE>
static void launch_persist_mode(gca_ctx ctx, const PairDesc *pairs, IcpState *states, int s_first, int n_states,
                                int refresh, uint32_t *err)
{
    const int blocks = ctx->tune[GCA_TUNE_ICP_PERSIST_BLOCKS] ? ctx->tune[GCA_TUNE_ICP_PERSIST_BLOCKS] : 2 * ctx->n_cu;
    hipLaunchKernelGGL(k_icp_persist<MODE>, dim3((unsigned)blocks), dim3(kIcpBlock), 0, ctx->stream, pairs, states,
                       s_first, n_states, refresh, err);
}

// The persistent launch over states [s_first, s_first + n_states) (the chain links of their
// finishers may open states after the range: those run in a later launch).  refresh: one state,
// whose sizes the launch takes from its clouds' headers.
static void launch_persist_any(gca_ctx ctx, int mode, const PairDesc *pairs, IcpState *states, int n_states,
                               uint32_t *err, int s_first = 0, int refresh = 0)
{
    StageTimer timer(ctx, GCA_STAGE_ICP);
    if (mode == GCA_ACCUM_NAIVE_FP32)
        launch_persist_mode<GCA_ACCUM_NAIVE_FP32>(ctx, pairs, states, s_first, n_states, refresh, err);
    else if (mode == GCA_ACCUM_FP64)
        launch_persist_mode<GCA_ACCUM_FP64>(ctx, pairs, states, s_first, n_states, refresh, err);
    else
        launch_persist_mode<GCA_ACCUM_COMPENSATED>(ctx, pairs, states, s_first, n_states, refresh, err);
}

// The resident launch (k_icp_resident) over state s: one workgroup per unit, so it is used only
// while a state's units fit two per CU (unit_cap <= 2 CUs).  Launch strategy 0 (default) takes it
// for one-pair calls, 3 for every call.
static bool resident_fits(gca_ctx ctx, int32_t unit_cap) { return unit_cap <= 2 * std::max(ctx->n_cu, 1); }

static bool use_resident(gca_ctx ctx, int n_pairs, int32_t unit_cap)
{
    const int32_t L = ctx->tune[GCA_TUNE_ICP_LAUNCH];
    if (!resident_fits(ctx, unit_cap)) return false;
    return L == 3 || (L == 0 && n_pairs == 1);
}

static void launch_resident_any(gca_ctx ctx, int mode, const PairDesc *pairs, IcpState *states, int s, int refresh,
                                int32_t unit_cap, uint32_t *err)
{
    StageTimer timer(ctx, GCA_STAGE_ICP);
    const dim3 g((unsigned)unit_cap), b(kIcpBlock);
    if (mode == GCA_ACCUM_NAIVE_FP32)
        hipLaunchKernelGGL(k_icp_resident<GCA_ACCUM_NAIVE_FP32>, g, b, 0, ctx->stream, pairs, states, s, refresh, err);
    else if (mode == GCA_ACCUM_FP64)
        hipLaunchKernelGGL(k_icp_resident<GCA_ACCUM_FP64>, g, b, 0, ctx->stream, pairs, states, s, refresh, err);
    else
        hipLaunchKernelGGL(k_icp_resident<GCA_ACCUM_COMPENSATED>, g, b, 0, ctx->stream, pairs, states, s, refresh, err);
}

// Poses and reports from a host copy of the states.
static void states_report(gca_ctx ctx, const IcpState *S, int n_pairs, int n_states, const std::vector<int32_t> &tail,
                          const double *T0, double *T_out, gca_icp_report *reps)
{
    for (int k = 0; k < n_pairs; ++k)
        memcpy(T_out + 16 * k, tail[k] >= 0 ? S[tail[k]].T : T0 + 16 * k, sizeof(double) * 16);
    if (!reps) return;
    long long t0 = S[0].t_init;
    for (int s = 1; s < n_states; ++s) t0 = std::min(t0, S[s].t_init);
    const double ticks_per_us = ctx->wall_khz > 0 ? ctx->wall_khz * 1e-3 : 0.0;
    for (int s = 0; s < n_states; ++s) {
        gca_icp_report &r = reps[s];
        const double n = (double)S[s].n_src;
        r.iterations = S[s].iter;
        r.converged = S[s].converged;
        r.solver_failures = S[s].solver_failures;
        r.n_correspondences = (int32_t)S[s].sys[28];
        r.rmse = n > 0 ? S[s].sys[27] / n : 0.0;  // color_icp.cu:231
        r.fitness = n > 0 ? S[s].sys[28] / n : 0.0;
        r.latency_us = ticks_per_us > 0 ? (double)(S[s].t_done - t0) / ticks_per_us : 0.0;
        r.active_us = ticks_per_us > 0 && S[s].t_start > 0 ? (double)(S[s].t_done - S[s].t_start) / ticks_per_us : 0.0;
    }
}

// One-shot registration (every call sets up, runs and reads back).  T0 / T_out: n_pairs x 16;
// reports: n_states (may be NULL).
static int32_t run_states(gca_ctx ctx, int n_pairs, int n_levels, const gca_cloud *srcs, const gca_cloud *tgts,
                          const gca_icp_params *params, const float *grad_r, const double *T0, double *T_out,
                          gca_icp_report *reps)
{
    const int n_states = n_pairs * n_levels;
    int32_t lanes[3], unit_cap;
    std::vector<int64_t> nn_off;
    GCA_TRY(states_prepare(ctx, n_pairs, n_levels, srcs, tgts, params, grad_r, false, lanes, &unit_cap, nn_off));
    const bool per_iter = ctx->tune[GCA_TUNE_ICP_LAUNCH] == 2;
    std::vector<int32_t> nvb(n_states, 0);
    int64_t table_units = 0;
    if (per_iter) {  // the launch grids need the sizes on the host
        for (int s = 0; s < n_states; ++s) {
            GCA_TRY(cloud_sync_header(srcs[s]));
            const int64_t n = srcs[s]->host_hdr.n;
            nvb[s] = icp_units_for(n, icp_lanes_for(lanes, n), unit_cap);
            table_units += nvb[s];
        }
    }
    IcpWork w;
    GCA_TRY(icp_work(ctx, n_states, n_pairs, (int64_t)n_states * 2 * unit_cap, nn_off[n_states], table_units, w));
    std::vector<PairDesc> P;
    std::vector<StateInit> I;
    std::vector<int32_t> tail;
    states_describe(n_pairs, n_levels, srcs, tgts, params, lanes, unit_cap, w.partials, w.nnprev, nn_off, P, I, tail);
    GCA_HIP(hipMemcpyAsync(w.pairs, P.data(), sizeof(PairDesc) * n_states, hipMemcpyHostToDevice, ctx->stream));
    GCA_HIP(hipMemcpyAsync(w.init, I.data(), sizeof(StateInit) * n_states, hipMemcpyHostToDevice, ctx->stream));
    GCA_HIP(hipMemcpyAsync(w.T0, T0, sizeof(double) * 16 * n_pairs, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_state_init, dim3((n_states + 63) / 64), dim3(64), 0, ctx->stream, w.states,
                       (const PairDesc *)w.pairs, (const StateInit *)w.init, (const double *)w.T0, n_states, w.err);
    GCA_HIP(hipGetLastError());

    const int mode = params[0].accumulate_mode;
    std::vector<int32_t> us, uf(n_states, 0);
    if (!per_iter && use_resident(ctx, n_pairs, unit_cap)) {
        // one launch per state, pair by pair and coarse to fine: a state's launch hands its pose to
        // the next level of its chain before that level's launch starts
        for (int s = 0; s < n_states; ++s) launch_resident_any(ctx, mode, w.pairs, w.states, s, 0, unit_cap, w.err);
        GCA_HIP(hipGetLastError());
    } else if (!per_iter) {
        launch_persist_any(ctx, mode, w.pairs, w.states, n_states, w.err);
        GCA_HIP(hipGetLastError());
    } else {
        // unit tables level by level: level l's launches cover the units of its n_pairs states
        std::vector<int64_t> loff(n_levels + 1, 0);
        us.reserve((size_t)table_units);
        for (int l = 0; l < n_levels; ++l) {
            loff[l] = (int64_t)us.size();
            for (int k = 0; k < n_pairs; ++k) {
                const int s = k * n_levels + l;
                uf[s] = (int32_t)((int64_t)us.size() - loff[l]);
                for (int b = 0; b < nvb[s]; ++b) us.push_back(s);
            }
        }
        loff[n_levels] = (int64_t)us.size();
        GCA_HIP(hipMemcpyAsync(w.unit_state, us.data(), sizeof(int32_t) * us.size(), hipMemcpyHostToDevice, ctx->stream));
        GCA_HIP(hipMemcpyAsync(w.unit_first, uf.data(), sizeof(int32_t) * n_states, hipMemcpyHostToDevice, ctx->stream));
        for (int l = 0; l < n_levels; ++l) {
            const int64_t units = loff[l + 1] - loff[l];
            for (int it = 0; it < params[l].max_iterations && units > 0; ++it) {
                StageTimer timer(ctx, GCA_STAGE_ICP);
                if (mode == GCA_ACCUM_NAIVE_FP32) launch_iter<GCA_ACCUM_NAIVE_FP32>(ctx, w, loff[l], units);
                else if (mode == GCA_ACCUM_FP64) launch_iter<GCA_ACCUM_FP64>(ctx, w, loff[l], units);
                else launch_iter<GCA_ACCUM_COMPENSATED>(ctx, w, loff[l], units);
            }
        }
        GCA_HIP(hipGetLastError());
    }

    std::vector<IcpState> S(n_states);
    uint32_t err = 0;
    // one contiguous copy (a strided copy into pageable memory costs a transfer per row)
    GCA_HIP(hipMemcpyAsync(S.data(), w.states, sizeof(IcpState) * n_states, hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipMemcpyAsync(&err, w.err, sizeof(err), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    if (err) return GCA_ERR_DEVICE;
    states_report(ctx, S.data(), n_pairs, n_states, tail, T0, T_out, reps);
    return GCA_OK;
}

// ---- registration plans: set up once, launched many times without host round trips --------
struct IcpPlan {
    int n_pairs = 0, n_levels = 0, n_states = 0, mode = 0;
    int32_t unit_cap = 0;
    IcpState *states = nullptr;
    double *partials = nullptr;
    int32_t *nnprev = nullptr;
    PairDesc *pairs = nullptr;
    StateInit *init = nullptr;
    double *T0 = nullptr;
    uint32_t *err = nullptr;
    IcpState *host_states = nullptr;  // pinned: written by the launch's copy, read after a sync
    uint32_t *host_err = nullptr;
    std::vector<int32_t> tail;
    std::vector<double> T0h;
    std::vector<PairDesc> desc;  // host copy: the target grids the plan reads
};

bool grid_same_geometry(const GridDesc &a, const GridDesc &b)
{
    for (int k = 0; k < 3; ++k)
        if (a.origin[k] != b.origin[k] || a.dims[k] != b.dims[k]) return false;
    return a.cell == b.cell && a.ncells == b.ncells;
}

bool icp_plan_target_current(const IcpPlan *p, int s, gca_cloud tgt)
{
    if (!p || s < 0 || s >= p->n_states) return false;
    const PairDesc &P = p->desc[(size_t)s];
    const Grid &G = tgt->grid;
    return G.valid && P.start == G.start && P.order == G.order && P.sxyzi == G.sxyzi && P.snrm == G.snrm &&
           P.sgrad == G.sgrad && grid_same_geometry(P.g, G.desc);
}

void icp_plan_destroy(IcpPlan *p)
{
    if (!p) return;
    hipFree(p->states);
    hipFree(p->partials);
    hipFree(p->nnprev);
    hipFree(p->pairs);
    hipFree(p->init);
    hipFree(p->T0);
    hipFree(p->err);
    hipHostFree(p->host_states);
    hipHostFree(p->host_err);
    delete p;
}

int32_t icp_plan_create(gca_ctx ctx, int n_pairs, int n_levels, const gca_cloud *srcs, const gca_cloud *tgts,
                        const gca_icp_params *params, const float *grad_r, const double *T0, IcpPlan **out)
{
    *out = nullptr;
    const int n_states = n_pairs * n_levels;
    int32_t lanes[3], unit_cap;
    std::vector<int64_t> nn_off;
    // a plan keeps the targets' grids: later steps rebuild them in place with the same geometry
    GCA_TRY(states_prepare(ctx, n_pairs, n_levels, srcs, tgts, params, grad_r, true, lanes, &unit_cap, nn_off));
    IcpPlan *p = new (std::nothrow) IcpPlan();
    if (!p) return GCA_ERR_OUT_OF_MEMORY;
    p->n_pairs = n_pairs;
    p->n_levels = n_levels;
    p->n_states = n_states;
    p->mode = params[0].accumulate_mode;
    p->unit_cap = unit_cap;
    p->T0h.assign(T0, T0 + 16 * n_pairs);
    auto fail = [&](hipError_t e) {
        icp_plan_destroy(p);
        return hip_status(e);
    };
    hipError_t e;
    if ((e = hipMalloc(&p->states, sizeof(IcpState) * n_states)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->partials, sizeof(double) * kPartStride * n_states * 2 * unit_cap)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->nnprev, sizeof(int32_t) * std::max<int64_t>(nn_off[n_states], 1))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->pairs, sizeof(PairDesc) * n_states)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->init, sizeof(StateInit) * n_states)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->T0, sizeof(double) * 16 * n_pairs)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&p->err, 256)) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc(&p->host_states, sizeof(IcpState) * n_states)) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc(&p->host_err, 256)) != hipSuccess) return fail(e);
    std::vector<PairDesc> P;
    std::vector<StateInit> I;
    states_describe(n_pairs, n_levels, srcs, tgts, params, lanes, unit_cap, p->partials, p->nnprev, nn_off, P, I,
                    p->tail);
    p->desc = P;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess ||
        (e = hipMemcpy(p->pairs, P.data(), sizeof(PairDesc) * n_states, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->init, I.data(), sizeof(StateInit) * n_states, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(p->T0, T0, sizeof(double) * 16 * n_pairs, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e);
    *out = p;
    return GCA_OK;
}

// Plan launch in parts, for a caller that overlaps each state's registration with the
// building of the next state's source: states set up (sizes are provisional), then per state
// its sizes and one persistent launch over it alone (each on the stream that built its source,
// after the launch of the state chained before it), then the results.
int32_t icp_plan_init(gca_ctx ctx, IcpPlan *p)
{
    hipLaunchKernelGGL(k_state_init, dim3((p->n_states + 63) / 64), dim3(64), 0, ctx->stream, p->states,
                       (const PairDesc *)p->pairs, (const StateInit *)p->init, (const double *)p->T0, p->n_states,
                       p->err);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

int32_t icp_plan_persist(gca_ctx ctx, IcpPlan *p, int s)
{
    if (s < 0 || s >= p->n_states) return GCA_ERR_INVALID_VALUE;
    if (use_resident(ctx, p->n_pairs, p->unit_cap))
        launch_resident_any(ctx, p->mode, p->pairs, p->states, s, 1, p->unit_cap, p->err);
    else
        launch_persist_any(ctx, p->mode, p->pairs, p->states, 1, p->err, s, 1);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// The states, the launch's error word and an optional caller word, stored straight into pinned
// host memory (device-accessible) by a grid-stride loop: one launch instead of a copy each.
__global__ __launch_bounds__(256) void k_results_to_host(const uint4 *__restrict__ states, int n16,
                                                         const uint32_t *__restrict__ err,
                                                         const uint32_t *__restrict__ extra, uint4 *host_states,
                                                         uint32_t *host_err, uint32_t *host_extra)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += gridDim.x * blockDim.x)
        host_states[k] = states[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *host_err = *err;
        if (extra) *host_extra = *extra;
    }
}

int32_t icp_plan_copy_results(gca_ctx ctx, IcpPlan *p, const uint32_t *extra, uint32_t *host_extra)
{
    static_assert(sizeof(IcpState) % 16 == 0, "states are copied as 16-byte words");
    const int n16 = (int)(sizeof(IcpState) * p->n_states / 16);
    const unsigned nb = (unsigned)std::min(64, (n16 + 255) / 256);
    hipLaunchKernelGGL(k_results_to_host, dim3(nb), dim3(256), 0, ctx->stream, (const uint4 *)p->states, n16,
                       (const uint32_t *)p->err, extra, (uint4 *)p->host_states, p->host_err, host_extra);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// One whole run of a plan on the ctx stream: states set up, every pair and level registered
// (one resident launch per state for a single pair, else one persistent launch over all), the
// states and the error word stored into the plan's pinned host memory.  No host round trip and
// no allocation, so it can be captured.
int32_t icp_plan_run(gca_ctx ctx, IcpPlan *p)
{
    GCA_TRY(icp_plan_init(ctx, p));
    if (use_resident(ctx, p->n_pairs, p->unit_cap))
        for (int s = 0; s < p->n_states; ++s)
            launch_resident_any(ctx, p->mode, p->pairs, p->states, s, 0, p->unit_cap, p->err);
    else
        launch_persist_any(ctx, p->mode, p->pairs, p->states, p->n_states, p->err);
    GCA_HIP(hipGetLastError());
    return icp_plan_copy_results(ctx, p, nullptr, nullptr);
}

// Whether every cloud of a plan still has the buffers (and, for a target, the neighbour grid
// and its gathered attributes) the plan's descriptors point at.
bool icp_plan_clouds_current(const IcpPlan *p, const gca_cloud *srcs, const int64_t *src_cap, const gca_cloud *tgts)
{
    for (int s = 0; s < p->n_states; ++s) {
        const PairDesc &P = p->desc[(size_t)s];
        const gca_cloud src = srcs[s], tgt = tgts[s];
        if (P.src != src->xyzi || P.src_hdr != src->hdr || src->cap != src_cap[s]) return false;
        if (!tgt->has_normals || !tgt->has_gradient || !tgt->grid.attrs_valid) return false;
        if (P.tgt_hdr != tgt->hdr || P.nrm != tgt->nrm || P.grad != tgt->grad) return false;
        if (!icp_plan_target_current(p, s, tgt)) return false;
    }
    return true;
}

// After the ctx stream has been synchronised.
int32_t icp_plan_results(gca_ctx ctx, const IcpPlan *p, double *T_out, gca_icp_report *reps)
{
    if (*p->host_err) return GCA_ERR_DEVICE;
    states_report(ctx, p->host_states, p->n_pairs, p->n_states, p->tail, p->T0h.data(), T_out, reps);
    return GCA_OK;
}

// Set-up of the single-pair kernels over given correspondences.
static int32_t single_pair(gca_ctx ctx, gca_cloud src, gca_cloud tgt, IcpWork &w, PairDesc &P, int64_t &n)
{
    GCA_TRY(cloud_sync_header(src));
    n = src->host_hdr.n;
    const int32_t nb = icp_blocks(n);
    GCA_TRY(icp_work(ctx, 1, 1, nb, 1, 0, w));
    P = PairDesc{};
    P.src = src->xyzi;
    P.src_hdr = src->hdr;
    P.tgt_hdr = tgt->hdr;
    P.nrm = tgt->nrm;
    P.grad = tgt->grad;
    P.partials = w.partials;
    P.st = w.states;
    P.nblk = nb;
    static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    GCA_HIP(hipMemcpyAsync(w.T0, I, sizeof(I), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_state_simple, dim3(1), dim3(64), 0, ctx->stream, w.states, (const double *)w.T0);
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
    GCA_TRY(single_pair(ctx, src, tgt, w, P, n));
    launch_build_gn(ctx, P, tgt->xyzi, nn_dev, T, lambda, mode, &w.states->ticket, w.states->sys);
    GCA_HIP(hipGetLastError());
    IcpState S;
    GCA_HIP(hipMemcpyAsync(&S, w.states, offsetof(IcpState, q), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(out, S.sys, sizeof(double) * kNSys);
    return GCA_OK;
}

int32_t gca_build_gauss_newton_noalloc(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn_dev,
                                       float lambda, const double T[16], int32_t mode, double *out_dev)
{
    if (!ctx || !src || !tgt || !nn_dev || !T || !out_dev) return GCA_ERR_INVALID_VALUE;
    if (lambda > 1.0f || !(lambda >= 0.0f) || mode < 0 || mode > 2) return GCA_ERR_INVALID_VALUE;  // .cu:438-441
    if (!tgt->has_normals || !tgt->has_gradient) return GCA_ERR_NOT_READY;                       // .cu:449-452
    if (src->ctx != ctx || tgt->ctx != ctx) return GCA_ERR_INVALID_VALUE;
    // the device's static scratch (set up by gca_ctx_create); workgroups stride over the source
    // points, so any cloud fits the fixed partial slots
    NoAllocScratch *sc = (NoAllocScratch *)ctx->noalloc_scratch;
    if (!sc) return GCA_ERR_DEVICE;
    const int32_t nb = icp_blocks(src->cap);  // capacity, known on the host: no size read-back
    if (nb > kNoAllocMaxBlocks) return GCA_ERR_INVALID_VALUE;
    PairDesc P{};
    P.src = src->xyzi;
    P.src_hdr = src->hdr;
    P.tgt_hdr = tgt->hdr;
    P.nrm = tgt->nrm;
    P.grad = tgt->grad;
    P.partials = sc->partials;
    P.nblk = nb;
    hipLaunchKernelGGL(k_ticket_clear, dim3(1), dim3(1), 0, ctx->stream, &sc->ticket);
    launch_build_gn(ctx, P, tgt->xyzi, nn_dev, T, lambda, mode, &sc->ticket, out_dev);
    GCA_HIP(hipGetLastError());
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
    GCA_TRY(single_pair(ctx, src, tgt, w, P, n));
    GCA_HIP(hipMemcpyAsync(w.T0 + 16, T, sizeof(double) * 16, hipMemcpyHostToDevice, ctx->stream));
    double *res = w.T0 + 32;
    const float sl = sqrtf(lambda), sc = sqrtf(1.0f - lambda);
    hipLaunchKernelGGL(k_residuals, dim3((unsigned)P.nblk), dim3(kIcpBlock), 0, ctx->stream, P, tgt->xyzi, nn_dev,
                       (const double *)(w.T0 + 16), lambda, sl, sc, kind, out_dev, res);
    GCA_HIP(hipGetLastError());
    double s = 0.0;
    GCA_HIP(hipMemcpyAsync(&s, res, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    if (sum) *sum = s;
    return GCA_OK;
}

int32_t gca_abs_residual_host(gca_ctx ctx, gca_cloud src, gca_cloud tgt, float max_corr, float lambda, float grad_r,
                              float *out_host, int64_t capacity, int64_t *n_out, double *mean)
{
    if (!ctx || !src || !tgt || !n_out || capacity < 0) return GCA_ERR_INVALID_VALUE;
    if (!(max_corr > 0.0f) || lambda > 1.0f || !(lambda >= 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!tgt->has_normals) return GCA_ERR_NOT_READY;  // color_icp.cu:104-110
    if (!tgt->has_gradient) {                        // color_icp.cu:112-117
        if (!(grad_r > 0.0f)) return GCA_ERR_NOT_READY;
        GCA_TRY(gca_compute_color_gradient(ctx, tgt, grad_r, 0));
    }
    GCA_TRY(cloud_sync_header(src));
    const int64_t n = src->host_hdr.n;
    *n_out = n;
    if (mean) *mean = 0.0;
    if (n > capacity || (n > 0 && !out_host)) return GCA_ERR_INVALID_VALUE;
    if (n == 0) return GCA_OK;
    int32_t *nn = (int32_t *)ctx_scratch(ctx, 34, sizeof(int32_t) * (size_t)n, false);
    float *res = (float *)ctx_scratch(ctx, 35, sizeof(float) * (size_t)n, false);
    if (!nn || !res) return GCA_ERR_OUT_OF_MEMORY;
    GCA_TRY(gca_nn_search(ctx, src, tgt, max_corr, nn));  // color_icp.cu:120
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double sum = 0.0;
    GCA_TRY(gca_compute_residuals(ctx, src, tgt, nn, lambda, I, GCA_RESIDUAL_ABS, res, &sum));
    GCA_HIP(hipMemcpyAsync(out_host, res, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    if (mean) *mean = sum / (double)n;  // color_icp.cu:134-135
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
    if (!ctx || !src || !tgt || !T_out || !p) return GCA_ERR_INVALID_VALUE;
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    return run_states(ctx, 1, 1, &src, &tgt, p, &grad_r, T0 ? T0 : I, T_out, rep);
}

int32_t gca_color_icp_batch(gca_ctx ctx, int32_t n_pairs, const gca_cloud *srcs, const gca_cloud *tgts,
                            const gca_icp_params *p, float grad_r, const double *T0, double *T_out,
                            gca_icp_report *reps)
{
    if (!ctx || !p || n_pairs < 0 || (n_pairs > 0 && (!srcs || !tgts || !T0 || !T_out))) return GCA_ERR_INVALID_VALUE;
    if (n_pairs == 0) return GCA_OK;
    return run_states(ctx, n_pairs, 1, srcs, tgts, p, &grad_r, T0, T_out, reps);
}

int32_t gca_color_icp_batch_multiscale(gca_ctx ctx, int32_t n_pairs, int32_t n_levels, const gca_cloud *srcs,
                                       const gca_cloud *tgts, const gca_icp_params *params,
                                       const float *gradient_radius, const double *T0, double *T_out,
                                       gca_icp_report *reports)
{
    if (!ctx || n_pairs < 0 || n_levels < 1 || n_levels > 16 || !params || !gradient_radius) return GCA_ERR_INVALID_VALUE;
    if (n_pairs > 0 && (!srcs || !tgts || !T0 || !T_out)) return GCA_ERR_INVALID_VALUE;
    if (n_pairs == 0) return GCA_OK;
    return run_states(ctx, n_pairs, n_levels, srcs, tgts, params, gradient_radius, T0, T_out, reports);
}

}  // extern "C"

// A batch registration set up once: its plan, the clouds it reads, and (use_graph) one run
// captured as a graph.
struct gca_icp_batch_plan_s {
    gca_ctx ctx = nullptr;
    IcpPlan *plan = nullptr;
    std::vector<gca_cloud> src, tgt;
    std::vector<int64_t> src_cap;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t captures = 0;
    bool ran = false;
};

static int32_t batch_plan_capture(gca_icp_batch_plan b)
{
    gca_ctx ctx = b->ctx;
    hipGraph_t g = nullptr;
    ctx->capturing = true;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    int32_t st = e == hipSuccess ? icp_plan_run(ctx, b->plan) : hip_status(e);
    if (e == hipSuccess) {
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (st == GCA_OK && e2 != hipSuccess) st = hip_status(e2);
    }
    ctx->capturing = false;
    if (st == GCA_OK && !g) st = GCA_ERR_DEVICE;
    hipGraphExec_t ex = nullptr;
    if (st == GCA_OK && (e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0)) != hipSuccess) st = hip_status(e);
    if (st != GCA_OK) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    b->graph = g;
    b->exec = ex;
    ++b->captures;
    return GCA_OK;
}

extern "C" {

int32_t gca_icp_batch_plan_create(gca_ctx ctx, int32_t n_pairs, int32_t n_levels, const gca_cloud *srcs,
                                  const gca_cloud *tgts, const gca_icp_params *params, const float *gradient_radius,
                                  const double *T0, int32_t use_graph, gca_icp_batch_plan *out)
{
    if (!out) return GCA_ERR_INVALID_VALUE;
    *out = nullptr;
    if (!ctx || n_pairs < 1 || n_levels < 1 || n_levels > 16 || !params || !gradient_radius || !srcs || !tgts || !T0)
        return GCA_ERR_INVALID_VALUE;
    const int n_states = n_pairs * n_levels;
    gca_icp_batch_plan b = new (std::nothrow) gca_icp_batch_plan_s();
    if (!b) return GCA_ERR_OUT_OF_MEMORY;
    b->ctx = ctx;
    b->use_graph = use_graph != 0;
    const int32_t st = icp_plan_create(ctx, n_pairs, n_levels, srcs, tgts, params, gradient_radius, T0, &b->plan);
    if (st != GCA_OK) {
        delete b;
        return st;
    }
    b->src.assign(srcs, srcs + n_states);
    b->tgt.assign(tgts, tgts + n_states);
    for (int s = 0; s < n_states; ++s) b->src_cap.push_back(srcs[s]->cap);
    *out = b;
    return GCA_OK;
}

int32_t gca_icp_batch_plan_enqueue(gca_icp_batch_plan b)
{
    if (!b) return GCA_ERR_INVALID_VALUE;
    if (!icp_plan_clouds_current(b->plan, b->src.data(), b->src_cap.data(), b->tgt.data()))
        return GCA_ERR_INVALID_VALUE;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    GCA_HIP(hipStreamIsCapturing(b->ctx->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) {  // into the caller's capture: no profiler events in it
        b->ctx->capturing = true;
        const int32_t st = icp_plan_run(b->ctx, b->plan);
        b->ctx->capturing = false;
        GCA_TRY(st);
    } else if (!b->use_graph) {
        GCA_TRY(icp_plan_run(b->ctx, b->plan));
    } else {
        if (!b->exec) GCA_TRY(batch_plan_capture(b));
        GCA_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
    }
    b->ran = true;
    return GCA_OK;
}

int32_t gca_icp_batch_plan_set_initial(gca_icp_batch_plan b, const double *T0)
{
    if (!b || !T0) return GCA_ERR_INVALID_VALUE;
    IcpPlan *p = b->plan;
    GCA_HIP(hipStreamSynchronize(b->ctx->stream));
    GCA_HIP(hipMemcpy(p->T0, T0, sizeof(double) * 16 * p->n_pairs, hipMemcpyHostToDevice));
    p->T0h.assign(T0, T0 + 16 * p->n_pairs);
    return GCA_OK;
}

int32_t gca_icp_batch_plan_results(gca_icp_batch_plan b, double *T_out, gca_icp_report *reports)
{
    if (!b || !T_out) return GCA_ERR_INVALID_VALUE;
    if (!b->ran) return GCA_ERR_NOT_READY;
    GCA_HIP(hipStreamSynchronize(b->ctx->stream));
    return icp_plan_results(b->ctx, b->plan, T_out, reports);
}

int32_t gca_icp_batch_plan_captures(gca_icp_batch_plan b, int64_t *captures)
{
    if (!b || !captures) return GCA_ERR_INVALID_VALUE;
    *captures = b->captures;
    return GCA_OK;
}

int32_t gca_icp_batch_plan_destroy(gca_icp_batch_plan b)
{
    if (!b) return GCA_OK;
    const hipError_t e = hipStreamSynchronize(b->ctx->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    icp_plan_destroy(b->plan);
    delete b;
    return e == hipSuccess ? GCA_OK : hip_status(e);
}

int32_t gca_multiscale_icp(gca_ctx ctx, gca_cloud src_full, gca_cloud tgt_full, int32_t n_levels,
                           const gca_level_params *levels, const double T0[16], double T_out[16],
                           gca_icp_report *reports)
{
    if (!ctx || !src_full || !tgt_full || !levels || !T_out || n_levels < 1 || n_levels > 8)
        return GCA_ERR_INVALID_VALUE;
    const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    gca_cloud cs[8], ct[8];
    gca_icp_params prm[8];
    float gr[8];
    for (int l = 0; l < n_levels; ++l) {
        const gca_level_params &L = levels[l];
        GCA_TRY(validate_params(&L.icp));
        gca_cloud *a = &ctx->level_clouds[2 * l], *b = &ctx->level_clouds[2 * l + 1];
        if (!*a) GCA_TRY(gca_cloud_create(ctx, 1, a));
        if (!*b) GCA_TRY(gca_cloud_create(ctx, 1, b));
        GCA_TRY(gca_voxel_downsample(ctx, src_full, L.voxel_size, *a));
        GCA_TRY(gca_voxel_downsample(ctx, tgt_full, L.voxel_size, *b));
        GCA_TRY(gca_estimate_normals(ctx, *b, L.normal_radius, L.normal_max_nn));
        GCA_TRY(gca_compute_color_gradient(ctx, *b, L.gradient_radius, L.gradient_max_nn));
        cs[l] = *a;
        ct[l] = *b;
        prm[l] = L.icp;
        gr[l] = L.gradient_radius;
    }
    // the levels in order, each starting from the pose the previous one ends with (one resident
    // launch per level, or one persistent launch over the chain)
    std::vector<gca_icp_report> rep(n_levels);
    std::vector<gca_cloud> S(cs, cs + n_levels), Tg(ct, ct + n_levels);
    GCA_TRY(run_states(ctx, 1, n_levels, S.data(), Tg.data(), prm, gr, T0 ? T0 : I, T_out, rep.data()));
    if (reports)
        for (int l = 0; l < n_levels; ++l) reports[l] = rep[l];
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
    GCA_TRY(single_pair(ctx, src, tgt, w, P, n));
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
