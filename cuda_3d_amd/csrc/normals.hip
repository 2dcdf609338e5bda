// normals.hip -- PCA normals and colour gradient: neighbourhood moments, then per-point solves.
//
// Replaces cuda_estimate_normals (src/geometry/cuda_estimate_normals.cu:22-199) and
// cuda_compute_color_gradient (src/registration/cuda_compute_color_gradient.cu:13-127).
//
// The reference materialises every (query, neighbour) pair, reduces 36-byte raw-moment
// records by key in FP32 and solves the 3x3 eigenproblem with an FP32 trigonometric closed
// form.  Here a group of G lanes walks the query's grid runs (all run bounds loaded first, then
// one flattened candidate walk with several loads in flight per lane), and the neighbourhood
// sums are accumulated in 2^-40 fixed point scaled by the radius, centred on the query point:
// integer adds are exact, so neither G nor the reduction order changes the result, which equals
// the FP64 oracle's.  The colour gradient's sums are the raw ones (sum u u^T, sum u dI); its
// tangent projection is applied to the sums in the solve, so they need no normal and can be
// gathered in the same walk as the normals' (k_moments_pair) when the gradient radius is the
// smaller one.  A query with more than K > 0 points within the radius is finished by one
// wavefront that buffers the in-radius candidates in LDS and keeps a candidate iff fewer than K
// others precede it in (d2, index) order.  The covariance is then diagonalised per point with a
// cyclic Jacobi iteration in FP64 (only +, -, *, /, sqrt: identical on host and device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

constexpr int kNbCap = 256;          // LDS candidates per wave
constexpr int kWavesPerBlock = 4;

struct NbBuf {
    double d2[kNbCap];
    int32_t idx[kNbCap];
    int32_t slot[kNbCap];
    float4 p[kNbCap];
};

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct GridRefs {
    GridDesc g;
    const uint32_t *start;
    const int32_t *order;
    const float4 *sxyzi;
};

constexpr int kNbRuns = 9;      // grid runs per query when the cell is at least the radius
constexpr int kOvfUnroll = 2;   // 64-candidate steps in flight per wave in the K-selection walk

// Calls f(slot, idx, point) on the lanes for every selected neighbour of q; returns the
// number of selected neighbours (wave-uniform).  K == 0: all within radius.  With at most
// kNbRuns runs (every grid grid_build makes) all run bounds are loaded at once and the runs are
// walked as one flattened candidate range, kOvfUnroll x 64 candidates in flight; the in-radius
// candidates go to LDS with their points.  The selection (the K smallest (d2, index) keys) and
// the count depend only on the candidate set, not on the visiting order.
template <class F>
__device__ __forceinline__ int32_t for_each_neighbour(const GridRefs &G, float4 q, int32_t K, NbBuf &buf, F &&f)
{
    int32_t lo[3], hi[3];
    if (!grid_range(G.g, q.x, q.y, q.z, lo, hi)) return 0;
    const double r2 = (double)G.g.radius * (double)G.g.radius;
    const int lane = lane_id();
    const int nry = hi[1] - lo[1] + 1;
    const int nruns = (hi[0] - lo[0] + 1) * nry;
    int32_t count = 0;
    wave_lds_sync();
    // stores one in-radius candidate per flagged lane at the next free LDS positions
    auto keep = [&](bool in, uint32_t s, double d2, int32_t ix, float4 p) {
        const uint64_t m = __ballot(in);
        if (in) {
            const int32_t pos = count + __popcll(m & ((1ull << lane) - 1ull));
            if (K == 0) {
                f(s, ix, p);
            } else if (pos < kNbCap) {
                buf.d2[pos] = d2;
                buf.idx[pos] = ix;
                buf.slot[pos] = (int32_t)s;
                buf.p[pos] = p;
            }
        }
        count += __popcll(m);
    };
    if (nruns <= kNbRuns) {
        uint32_t rs[kNbRuns], cum[kNbRuns + 1];
#pragma unroll
        for (int r = 0; r < kNbRuns; ++r) {
            uint32_t a = 0, b = 0;
            if (r < nruns) {
                const int ix = lo[0] + r / nry, iy = lo[1] + r % nry;
                const int64_t row = ((int64_t)ix * G.g.dims[1] + iy) * G.g.dims[2];
                a = G.start[row + lo[2]];
                b = G.start[row + hi[2] + 1];
            }
            rs[r] = a;
            cum[r + 1] = b - a;
        }
        cum[0] = 0;
#pragma unroll
        for (int r = 0; r < kNbRuns; ++r) cum[r + 1] += cum[r];
        const uint32_t total = cum[kNbRuns];
        for (uint32_t f0 = 0; f0 < total; f0 += kWave * kOvfUnroll) {
            float4 c[kOvfUnroll];
            int32_t ci[kOvfUnroll];
            uint32_t cs[kOvfUnroll];
#pragma unroll
            for (int u = 0; u < kOvfUnroll; ++u) {
                const uint32_t fi = f0 + u * kWave + lane;
                cs[u] = 0;
                if (fi < total) {
                    uint32_t sl = rs[0] + fi;
#pragma unroll
                    for (int r = 1; r < kNbRuns; ++r)
                        if (fi >= cum[r]) sl = rs[r] + (fi - cum[r]);
                    cs[u] = sl;
                    c[u] = G.sxyzi[sl];
                    ci[u] = G.order[sl];
                }
            }
#pragma unroll
            for (int u = 0; u < kOvfUnroll; ++u) {
                const bool valid = f0 + u * kWave + lane < total;
                const double d2 = valid ? dist2(c[u].x, c[u].y, c[u].z, q.x, q.y, q.z) : 0.0;
                keep(valid && d2 < r2, cs[u], d2, valid ? ci[u] : 0, valid ? c[u] : make_float4(0.f, 0.f, 0.f, 0.f));
            }
        }
    } else {
        // cell smaller than the radius (not produced by grid_build): run by run
        for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
            for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
                const int64_t row = ((int64_t)ix * G.g.dims[1] + iy) * G.g.dims[2];
                const uint32_t s0 = G.start[row + lo[2]], s1 = G.start[row + hi[2] + 1];
                for (uint32_t b = s0; b < s1; b += kWave) {
                    const uint32_t s = b + lane;
                    bool in = false;
                    double d2 = 0.0;
                    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
                    int32_t id = 0;
                    if (s < s1) {
                        p = G.sxyzi[s];
                        id = G.order[s];
                        d2 = dist2(p.x, p.y, p.z, q.x, q.y, q.z);
                        in = d2 < r2;
                    }
                    keep(in, s, d2, id, p);
                }
            }
    }
    if (K == 0) return count;
    wave_lds_sync();
    if (count <= K) {
        for (int32_t e = lane; e < count; e += kWave) f(buf.slot[e], buf.idx[e], buf.p[e]);
        return count;
    }
    if (count <= kNbCap) {
        for (int32_t e = lane; e < count; e += kWave) {
            const double de = buf.d2[e];
            const int32_t ie = buf.idx[e];
            int32_t rank = 0;
            for (int32_t j = 0; j < count; ++j) rank += key_less(buf.d2[j], buf.idx[j], de, ie) ? 1 : 0;
            if (rank < K) f(buf.slot[e], ie, buf.p[e]);
        }
        return K;
    }
    // Overflow of the LDS buffer (very dense neighbourhood): rank each candidate against a
    // rescan of the runs.  Slow but exact; only reached with > kNbCap points within radius.
    for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
        for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
            const int64_t row = ((int64_t)ix * G.g.dims[1] + iy) * G.g.dims[2];
            const uint32_t s0 = G.start[row + lo[2]], s1 = G.start[row + hi[2] + 1];
            for (uint32_t s = s0 + lane; s < s1; s += kWave) {
                const float4 p = G.sxyzi[s];
                const double de = dist2(p.x, p.y, p.z, q.x, q.y, q.z);
                if (!(de < r2)) continue;
                const int32_t ie = G.order[s];
                int32_t rank = 0;
                for (int32_t jx = lo[0]; jx <= hi[0] && rank < K; ++jx)
                    for (int32_t jy = lo[1]; jy <= hi[1] && rank < K; ++jy) {
                        const int64_t rw = ((int64_t)jx * G.g.dims[1] + jy) * G.g.dims[2];
                        for (uint32_t t = G.start[rw + lo[2]]; t < G.start[rw + hi[2] + 1]; ++t) {
                            const float4 c = G.sxyzi[t];
                            const double dc = dist2(c.x, c.y, c.z, q.x, q.y, q.z);
                            if (dc < r2 && key_less(dc, G.order[t], de, ie)) ++rank;
                        }
                    }
                if (rank < K) f(s, ie, p);
            }
        }
    return K;
}

// Cyclic Jacobi for a symmetric 3x3 (same operation sequence as the oracle's orc_eigen3).
__host__ __device__ inline void jacobi3(double A[3][3], double v[3][3])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = (A[0][1] * A[0][1] + A[0][2] * A[0][2]) + A[1][2] * A[1][2];
        double dg = (A[0][0] * A[0][0] + A[1][1] * A[1][1]) + A[2][2] * A[2][2];
        if (!(off > 1e-36 * dg)) break;
        for (int r = 0; r < 3; ++r) {
            const int p = r == 2 ? 1 : 0;
            const int q = r == 0 ? 1 : 2;
            double apq = A[p][q];
            if (apq == 0.0) continue;
            double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (theta < 0.0) t = -t;
            double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; ++k) {
                double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            A[p][q] = 0.0;
            A[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
}

__host__ __device__ inline float4 normal_from_moments(int32_t cnt, const long long S[9])
{
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (cnt < 3) return out;  // cuda_estimate_normals.cu:89-93
    const double n = (double)cnt;
    double m[3], A[3][3], V[3][3];
    for (int a = 0; a < 3; ++a) m[a] = ((double)S[a] / kFx40) / n;
    const int IA[6] = {0, 0, 0, 1, 1, 2}, IB[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
        double c = ((double)S[3 + k] / kFx40) / n - m[IA[k]] * m[IB[k]];
        A[IA[k]][IB[k]] = c;
        A[IB[k]][IA[k]] = c;
    }
    jacobi3(A, V);
    const double e0 = A[0][0], e1 = A[1][1], e2 = A[2][2];
    int sel;  // strict-< selection (cuda_estimate_normals.cu:116-125)
    if (e0 < e1 && e0 < e2) sel = 0;
    else if (e1 < e0 && e1 < e2) sel = 1;
    else sel = 2;
    double v[3] = {V[0][sel], V[1][sel], V[2][sel]};
    double len = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (!(len > 0.0)) return out;
    for (int a = 0; a < 3; ++a) v[a] /= len;
    if (-v[2] < 0.0)  // orient toward (0,0,-1) (cuda_estimate_normals.cu:131-151)
        for (int a = 0; a < 3; ++a) v[a] = -v[a];
    out.x = (float)v[0];
    out.y = (float)v[1];
    out.z = (float)v[2];
    return out;
}

// ---- group walk ------------------------------------------------------------------------------
// The common case (no K cap, or at most K points within the radius) runs G lanes per query:
// every lane visits its share of the query's flattened candidate space with kNbUnroll loads in
// flight and accumulates integer moments; the group then sums them with shuffles.  Integer
// sums are exact, so the result does not depend on G or on the visiting order.
constexpr int kNbBlock = 256;
constexpr int kNbUnroll = 4;

struct Moments {
    long long S[9];
    int32_t cnt;
    int32_t pad;
};

// Calls f(p, d2) on the lane for every candidate of its share with d2 < r^2; returns the
// group's in-radius count (on every lane of the group).
template <class F>
__device__ __forceinline__ int32_t group_walk(const GridRefs &G, float4 q, int g, int gs, F &&f)
{
    int32_t lo[3], hi[3];
    int32_t cnt = 0;
    if (grid_range(G.g, q.x, q.y, q.z, lo, hi)) {
        const double r2 = (double)G.g.radius * (double)G.g.radius;
        const int nry = hi[1] - lo[1] + 1;
        const int nruns = (hi[0] - lo[0] + 1) * nry;
        if (nruns <= kNbRuns) {
            uint32_t rs[kNbRuns], cum[kNbRuns + 1];
#pragma unroll
            for (int r = 0; r < kNbRuns; ++r) {
                uint32_t a = 0, b = 0;
                if (r < nruns) {
                    const int ix = lo[0] + r / nry, iy = lo[1] + r % nry;
                    const int64_t row = ((int64_t)ix * G.g.dims[1] + iy) * G.g.dims[2];
                    a = G.start[row + lo[2]];
                    b = G.start[row + hi[2] + 1];
                }
                rs[r] = a;
                cum[r + 1] = b - a;
            }
            cum[0] = 0;
#pragma unroll
            for (int r = 0; r < kNbRuns; ++r) cum[r + 1] += cum[r];
            const uint32_t total = cum[kNbRuns];
            for (uint32_t f0 = g; f0 < total; f0 += gs * kNbUnroll) {
                float4 c[kNbUnroll];
                bool ok[kNbUnroll];
#pragma unroll
                for (int u = 0; u < kNbUnroll; ++u) {
                    const uint32_t fi = f0 + u * gs;
                    ok[u] = fi < total;
                    if (ok[u]) {
                        uint32_t sl = rs[0] + fi;
#pragma unroll
                        for (int r = 1; r < kNbRuns; ++r)
                            if (fi >= cum[r]) sl = rs[r] + (fi - cum[r]);
                        c[u] = G.sxyzi[sl];
                    }
                }
#pragma unroll
                for (int u = 0; u < kNbUnroll; ++u) {
                    if (!ok[u]) continue;
                    const double d2 = dist2(c[u].x, c[u].y, c[u].z, q.x, q.y, q.z);
                    if (d2 < r2) {
                        f(c[u], d2);
                        ++cnt;
                    }
                }
            }
        } else {
            // cell smaller than the radius (not produced by grid_build): plain run walk
            for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
                for (int32_t iy = lo[1]; iy <= hi[1]; ++iy) {
                    const int64_t row = ((int64_t)ix * G.g.dims[1] + iy) * G.g.dims[2];
                    const uint32_t s0 = G.start[row + lo[2]], s1 = G.start[row + hi[2] + 1];
                    for (uint32_t sl = s0 + g; sl < s1; sl += gs) {
                        const float4 c = G.sxyzi[sl];
                        const double d2 = dist2(c.x, c.y, c.z, q.x, q.y, q.z);
                        if (d2 < r2) {
                            f(c, d2);
                            ++cnt;
                        }
                    }
                }
        }
    }
    for (int m = 1; m < gs; m <<= 1) cnt += __shfl_xor(cnt, m, kWave);
    return cnt;
}

// Neighbourhood moments (2^-40 fixed point) of one neighbour p of query q, u = (p - q) / r.
// Every neighbour is within the radius, so |u_a| <= 1 (to rounding) and the u and u u^T terms
// are below 2^41 in fixed point: fx_small converts them; the u dI terms use the general fx.
//
//   normals:  S = sum u, sum u u^T (xx xy xz yy yz zz)
//   gradient: S = sum u u^T, sum u dI with dI = I_p - I_q (raw: the projection onto the query's
//             tangent plane is applied to the sums, gradient_from_moments)
// The query itself contributes exactly zero to every sum (and one to the count).
template <int KIND>
__device__ __forceinline__ void add_moments(long long S[9], float4 q, float4 p, double inv_r)
{
    if (KIND == 0) {
        const double u0 = ((double)p.x - (double)q.x) * inv_r;
        const double u1 = ((double)p.y - (double)q.y) * inv_r;
        const double u2 = ((double)p.z - (double)q.z) * inv_r;
        S[0] += fx_small(u0 * kFx40);
        S[1] += fx_small(u1 * kFx40);
        S[2] += fx_small(u2 * kFx40);
        S[3] += fx_small((u0 * u0) * kFx40);
        S[4] += fx_small((u0 * u1) * kFx40);
        S[5] += fx_small((u0 * u2) * kFx40);
        S[6] += fx_small((u1 * u1) * kFx40);
        S[7] += fx_small((u1 * u2) * kFx40);
        S[8] += fx_small((u2 * u2) * kFx40);
    } else {
        const double u0 = ((double)p.x - (double)q.x) * inv_r;
        const double u1 = ((double)p.y - (double)q.y) * inv_r;
        const double u2 = ((double)p.z - (double)q.z) * inv_r;
        const double dI = (double)p.w - (double)q.w;
        S[0] += fx_small((u0 * u0) * kFx40);
        S[1] += fx_small((u0 * u1) * kFx40);
        S[2] += fx_small((u0 * u2) * kFx40);
        S[3] += fx_small((u1 * u1) * kFx40);
        S[4] += fx_small((u1 * u2) * kFx40);
        S[5] += fx_small((u2 * u2) * kFx40);
        S[6] += fx((u0 * dI) * kFx40);
        S[7] += fx((u1 * dI) * kFx40);
        S[8] += fx((u2 * dI) * kFx40);
    }
}

// Pass 1: moments of every query whose neighbourhood needs no K selection; queries with more
// than K > 0 points within the radius are appended to the overflow list instead.
template <int KIND>
__global__ __launch_bounds__(kNbBlock) void k_moments(const float4 *__restrict__ xyzi, const float4 *__restrict__ nrm,
                                                      const CloudHdr *__restrict__ hdr, GridDesc g,
                                                      const uint32_t *__restrict__ start,
                                                      const int32_t *__restrict__ order,
                                                      const float4 *__restrict__ sxyzi, int32_t K, int gs,
                                                      Moments *__restrict__ mom, int32_t *__restrict__ ovf,
                                                      uint32_t *__restrict__ ovf_n)
{
    const GridRefs G{g, start, order, sxyzi};
    const int n = hdr->n;
    const int lg = __builtin_ctz((unsigned)gs);
    const int gl = (int)threadIdx.x & (gs - 1);
    const int per_block = kNbBlock >> lg;
    const double inv_r = 1.0 / (double)g.radius;
    for (int base = blockIdx.x * per_block; base < n; base += gridDim.x * per_block) {
        const int qi = base + ((int)threadIdx.x >> lg);
        if (qi >= n) break;  // whole groups leave together
        const float4 q = xyzi[qi];
        long long S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int32_t cnt = group_walk(G, q, gl, gs, [&](float4 p, double) { add_moments<KIND>(S, q, p, inv_r); });
        if (K > 0 && cnt > K) {
            if (gl == 0) ovf[atomicAdd(ovf_n, 1u)] = qi;
            continue;
        }
#pragma unroll
        for (int c = 0; c < 9; ++c)
            for (int m = 1; m < gs; m <<= 1) S[c] += __shfl_xor(S[c], m, kWave);
        if (gl == 0) {
            Moments &M = mom[qi];
#pragma unroll
            for (int c = 0; c < 9; ++c) M.S[c] = S[c];
            M.cnt = cnt;
        }
    }
}

// Pass 1, normals and colour gradient together (gradient radius rg <= the normals' radius, no K
// cap for the gradient): the walk covers the normals' radius and the candidates within rg also
// add the gradient's raw sums, with the same d2 < rg^2 test as the gradient's own walk, so both
// moment sets equal those of separate walks.  A query with more than K > 0 points within the
// normals' radius goes to the overflow list for its normal moments; its gradient sums are
// complete here.
struct PairArgs {
    const float4 *xyzi;
    const CloudHdr *hdr;
    GridDesc g;
    const uint32_t *start;
    const int32_t *order;
    const float4 *sxyzi;
    int32_t K, gs;
    double rg;
    Moments *mom, *mom_g;
    int32_t *ovf;
    uint32_t *ovf_n;
};

template <int N>
__global__ __launch_bounds__(kNbBlock) void k_moments_pair(Jobs<PairArgs, N> J)
{
    const PairArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    const float4 *__restrict__ xyzi = A.xyzi;
    const GridDesc &g = A.g;
    const int32_t K = A.K;
    const int gs = A.gs;
    const double rg = A.rg;
    Moments *__restrict__ mom = A.mom;
    Moments *__restrict__ mom_g = A.mom_g;
    int32_t *__restrict__ ovf = A.ovf;
    uint32_t *__restrict__ ovf_n = A.ovf_n;
    const GridRefs G{g, A.start, A.order, A.sxyzi};
    const int n = A.hdr->n;
    const int lg = __builtin_ctz((unsigned)gs);
    const int gl = (int)threadIdx.x & (gs - 1);
    const int per_block = kNbBlock >> lg;
    const double inv_r = 1.0 / (double)g.radius;
    const double inv_rg = 1.0 / rg;
    const double rg2 = rg * rg;
    for (int base = blockIdx.x * per_block; base < n; base += gridDim.x * per_block) {
        const int qi = base + ((int)threadIdx.x >> lg);
        if (qi >= n) break;  // whole groups leave together
        const float4 q = xyzi[qi];
        long long S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int32_t cg = 0;
        const int32_t cnt = group_walk(G, q, gl, gs, [&](float4 p, double d2) {
            add_moments<0>(S, q, p, inv_r);
            if (d2 < rg2) {
                add_moments<1>(Sg, q, p, inv_rg);
                ++cg;
            }
        });
        for (int m = 1; m < gs; m <<= 1) cg += __shfl_xor(cg, m, kWave);
#pragma unroll
        for (int c = 0; c < 9; ++c)
            for (int m = 1; m < gs; m <<= 1) Sg[c] += __shfl_xor(Sg[c], m, kWave);
        if (gl == 0) {
            Moments &M = mom_g[qi];
#pragma unroll
            for (int c = 0; c < 9; ++c) M.S[c] = Sg[c];
            M.cnt = cg;
        }
        if (K > 0 && cnt > K) {
            if (gl == 0) ovf[atomicAdd(ovf_n, 1u)] = qi;
            continue;
        }
#pragma unroll
        for (int c = 0; c < 9; ++c)
            for (int m = 1; m < gs; m <<= 1) S[c] += __shfl_xor(S[c], m, kWave);
        if (gl == 0) {
            Moments &M = mom[qi];
#pragma unroll
            for (int c = 0; c < 9; ++c) M.S[c] = S[c];
            M.cnt = cnt;
        }
    }
}

// Pass 2: the overflow queries, one wavefront each, with exact K-nearest selection.
struct OvfArgs {
    const float4 *xyzi, *nrm;
    const CloudHdr *hdr;
    GridDesc g;
    const uint32_t *start;
    const int32_t *order;
    const float4 *sxyzi;
    int32_t K, pad;
    const int32_t *ovf;
    const uint32_t *ovf_n;
    Moments *mom;
};

template <int KIND, int N>
__global__ __launch_bounds__(kWave *kWavesPerBlock) void k_moments_overflow(Jobs<OvfArgs, N> J)
{
    const OvfArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    const float4 *__restrict__ xyzi = A.xyzi;
    const GridDesc &g = A.g;
    const int32_t K = A.K;
    const int32_t *__restrict__ ovf = A.ovf;
    Moments *__restrict__ mom = A.mom;
    __shared__ NbBuf bufs[kWavesPerBlock];
    const int w = threadIdx.x / kWave;
    NbBuf &buf = bufs[w];
    const GridRefs G{g, A.start, A.order, A.sxyzi};
    const int n = A.hdr->n;
    const uint32_t m = *A.ovf_n;
    const double inv_r = 1.0 / (double)g.radius;
    for (uint32_t e = blockIdx.x * kWavesPerBlock + w; e < m; e += gridDim.x * kWavesPerBlock) {
        const int qi = ovf[e];
        if (qi < 0 || qi >= n) continue;
        const float4 q = xyzi[qi];
        long long S[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int32_t cnt = for_each_neighbour(G, q, K, buf, [&](uint32_t, int32_t, float4 p) {
            add_moments<KIND>(S, q, p, inv_r);
        });
#pragma unroll
        for (int c = 0; c < 9; ++c) S[c] = wave_sum(S[c]);
        if (lane_id() == 0) {
            Moments &M = mom[qi];
            for (int c = 0; c < 9; ++c) M.S[c] = S[c];
            M.cnt = cnt;
        }
    }
}

// Colour gradient from the tangent-plane moments: A = sum w w^T r^2 + (k-1)^2 n n^T + 1e-6 I,
// b = sum w dI r, solved by unpivoted LDL^T (cuda_compute_color_gradient.cu:41-90,
// little_3x3_ldlt_solve.cuh:11-60); fewer than 5 neighbours (query included) -> zero.
__device__ __forceinline__ float4 gradient_from_moments(int32_t k, const long long S[9], float4 nf, double radius)
{
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < 5) return out;  // cuda_compute_color_gradient.cu:41-44
    const double nv[3] = {(double)nf.x, (double)nf.y, (double)nf.z};
    const double rr = radius * radius;
    const double kk = (double)(k - 1) * (double)(k - 1);
    const int IA[6] = {0, 0, 0, 1, 1, 2}, IB[6] = {0, 1, 2, 1, 2, 2};
    // sum w w^T = P E P and sum w dI = P F for w = P u, P = I - n n^T (:48-60), E and F the raw
    // sums; the same expressions as the oracle's project_tangent
    double M[3][3], f[3], Mn[3];
    for (int c = 0; c < 6; ++c) {
        const double v = (double)S[c] / kFx40;
        M[IA[c]][IB[c]] = v;
        M[IB[c]][IA[c]] = v;
    }
    for (int a = 0; a < 3; ++a) f[a] = (double)S[6 + a] / kFx40;
    for (int a = 0; a < 3; ++a) Mn[a] = (M[a][0] * nv[0] + M[a][1] * nv[1]) + M[a][2] * nv[2];
    const double nMn = (nv[0] * Mn[0] + nv[1] * Mn[1]) + nv[2] * Mn[2];
    double A[3][3];
    for (int c = 0; c < 6; ++c) {
        const int a = IA[c], bb = IB[c];
        const double pmp = ((M[a][bb] - nv[a] * Mn[bb]) - Mn[a] * nv[bb]) + (nv[a] * nv[bb]) * nMn;
        double v = pmp * rr + kk * (nv[a] * nv[bb]);
        A[a][bb] = v;
        A[bb][a] = v;
    }
    for (int a = 0; a < 3; ++a) A[a][a] += 1.0e-6;
    const double fn = (f[0] * nv[0] + f[1] * nv[1]) + f[2] * nv[2];
    double b[3];
    for (int a = 0; a < 3; ++a) b[a] = (f[a] - fn * nv[a]) * radius;
    double D0 = A[0][0];
    double L10 = A[1][0] / D0, L20 = A[2][0] / D0;
    double D1 = A[1][1] - L10 * L10 * D0;
    double L21 = (A[2][1] - L10 * L20 * D0) / D1;
    double D2 = A[2][2] - L20 * L20 * D0 - L21 * L21 * D1;
    double x0 = b[0];
    double x1 = b[1] - L10 * x0;
    double x2 = b[2] - L20 * x0 - L21 * x1;
    x0 /= D0;
    x1 /= D1;
    x2 /= D2;
    x1 -= L21 * x2;
    x0 -= L10 * x1 + L20 * x2;
    return make_float4((float)x0, (float)x1, (float)x2, 0.0f);
}

// Pass 3: one thread per point (FP64 eigen / LDL^T solves keep every lane busy).  Re-arms the
// overflow counter for the next call.
template <int KIND>
__global__ __launch_bounds__(kNbBlock) void k_solve(const CloudHdr *__restrict__ hdr, const Moments *__restrict__ mom,
                                                    const float4 *__restrict__ nrm, double radius,
                                                    float4 *__restrict__ out, uint32_t *__restrict__ ovf_n)
{
    const int n = hdr->n;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ovf_n = 0u;
    for (int i = blockIdx.x * kNbBlock + threadIdx.x; i < n; i += gridDim.x * kNbBlock) {
        const Moments &M = mom[i];
        long long S[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) S[c] = M.S[c];
        out[i] = KIND == 0 ? normal_from_moments(M.cnt, S) : gradient_from_moments(M.cnt, S, nrm[i], radius);
    }
}

// Pass 3 of the joint walk: the normal, then the colour gradient on that normal (the value
// stored, as a separate gradient call would read it back).
// order != nullptr: thread s solves point order[s] (the grid's sorted slot s) and also writes
// its normal and gradient to snrm[s] / sgrad[s], the grid-ordered copies registration reads.
struct SolveArgs {
    const CloudHdr *hdr;
    const Moments *mom, *mom_g;
    double rg;
    float4 *nrm, *grad;
    const int32_t *order;
    float4 *snrm, *sgrad;
    uint32_t *ovf_n;
};

template <int N>
__global__ __launch_bounds__(kNbBlock) void k_solve_pair(Jobs<SolveArgs, N> J)
{
    const SolveArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    const Moments *__restrict__ mom = A.mom;
    const Moments *__restrict__ mom_g = A.mom_g;
    const double rg = A.rg;
    float4 *__restrict__ nrm = A.nrm;
    float4 *__restrict__ grad = A.grad;
    const int32_t *__restrict__ order = A.order;
    float4 *__restrict__ snrm = A.snrm;
    float4 *__restrict__ sgrad = A.sgrad;
    const int n = A.hdr->n;
    if (blockIdx.x == 0 && threadIdx.x == 0) *A.ovf_n = 0u;
    for (int s = blockIdx.x * kNbBlock + threadIdx.x; s < n; s += gridDim.x * kNbBlock) {
        const int i = order ? order[s] : s;
        long long S[9];
        const Moments &M = mom[i];
#pragma unroll
        for (int c = 0; c < 9; ++c) S[c] = M.S[c];
        const float4 nf = normal_from_moments(M.cnt, S);
        nrm[i] = nf;
        const Moments &Mg = mom_g[i];
#pragma unroll
        for (int c = 0; c < 9; ++c) S[c] = Mg.S[c];
        const float4 gv = gradient_from_moments(Mg.cnt, S, nf, rg);
        grad[i] = gv;
        if (order) {
            snrm[s] = nf;
            sgrad[s] = gv;
        }
    }
}

static unsigned grid_for(int64_t items, int per_block, int64_t cap)
{
    return (unsigned)std::min<int64_t>(cap, std::max<int64_t>(1, (items + per_block - 1) / per_block));
}

static int normals_lanes(gca_ctx ctx) { return ctx->tune[GCA_TUNE_NORMALS_LANES] ? ctx->tune[GCA_TUNE_NORMALS_LANES] : 8; }

template <int KIND>
static int32_t launch_moments_pipeline(gca_ctx ctx, gca_cloud c, float radius, int32_t K, float4 *out)
{
    const Grid &G = c->grid;
    const int gs = normals_lanes(ctx);
    const int64_t cap = std::max<int64_t>(c->cap, 1);
    Moments *mom = (Moments *)ctx_scratch(ctx, 17, sizeof(Moments) * cap, false);
    int32_t *ovf = (int32_t *)ctx_scratch(ctx, 18, sizeof(int32_t) * cap, false);
    uint32_t *ovf_n = (uint32_t *)ctx_scratch(ctx, 19, 256, true);  // zero, re-armed by k_solve
    if (!mom || !ovf || !ovf_n) return GCA_ERR_OUT_OF_MEMORY;
    const GridDesc d = query_desc(G, radius);
    StageTimer timer(ctx, KIND == 0 ? GCA_STAGE_NORMALS : GCA_STAGE_GRADIENT);
    hipLaunchKernelGGL(k_moments<KIND>, dim3(grid_for(cap, kNbBlock / gs, 4096)), dim3(kNbBlock), 0, ctx->stream,
                       c->xyzi, c->nrm, c->hdr, d, G.start, G.order, G.sxyzi, K, gs, mom, ovf, ovf_n);
    if (K > 0) {
        const Jobs<OvfArgs, 1> oj{{{c->xyzi, c->nrm, c->hdr, d, G.start, G.order, G.sxyzi, K, 0, ovf, ovf_n, mom}}};
        hipLaunchKernelGGL((k_moments_overflow<KIND, 1>), dim3(grid_for(cap, kWavesPerBlock, 1024)),
                           dim3(kWave * kWavesPerBlock), 0, ctx->stream, oj);
    }
    hipLaunchKernelGGL(k_solve<KIND>, dim3(grid_for(cap, kNbBlock, 2048)), dim3(kNbBlock), 0, ctx->stream, c->hdr,
                       (const Moments *)mom, (const float4 *)c->nrm, (double)radius, out, ovf_n);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

int32_t launch_normals(gca_ctx ctx, gca_cloud c, float radius, int32_t K)
{
    return launch_moments_pipeline<0>(ctx, c, radius, K, c->nrm);
}

int32_t launch_gradient(gca_ctx ctx, gca_cloud c, float radius, int32_t K)
{
    return launch_moments_pipeline<1>(ctx, c, radius, K, c->grad);
}

// The three kernels of one joint normals + gradient call, as jobs.
struct PairJobs {
    PairArgs pair;
    OvfArgs ovf;
    SolveArgs solve;
    unsigned nb_pair, nb_ovf, nb_solve;
};

static int32_t pair_prepare(gca_ctx ctx, gca_cloud c, float radius, int32_t K, float gradient_radius, bool grid_attrs,
                            PairJobs *pj)
{
    if (!(gradient_radius <= radius)) return GCA_ERR_INVALID_VALUE;
    Grid &G = c->grid;
    if (grid_attrs && !G.snrm) {
        GCA_HIP(hipMalloc(&G.snrm, sizeof(float4) * G.cap_pts));
        GCA_HIP(hipMalloc(&G.sgrad, sizeof(float4) * G.cap_pts));
    }
    const int gs = normals_lanes(ctx);
    const int64_t cap = std::max<int64_t>(c->cap, 1);
    Moments *mom = (Moments *)ctx_scratch(ctx, 17, sizeof(Moments) * cap, false);
    int32_t *ovf = (int32_t *)ctx_scratch(ctx, 18, sizeof(int32_t) * cap, false);
    uint32_t *ovf_n = (uint32_t *)ctx_scratch(ctx, 19, 256, true);  // zero, re-armed by k_solve_pair
    Moments *mom_g = (Moments *)ctx_scratch(ctx, 24, sizeof(Moments) * cap, false);
    if (!mom || !ovf || !ovf_n || !mom_g) return GCA_ERR_OUT_OF_MEMORY;
    const GridDesc d = query_desc(G, radius);
    const double rg = (double)gradient_radius;
    pj->pair = PairArgs{c->xyzi, c->hdr, d, G.start, G.order, G.sxyzi, K, gs, rg, mom, mom_g, ovf, ovf_n};
    pj->ovf = OvfArgs{c->xyzi, c->nrm, c->hdr, d, G.start, G.order, G.sxyzi, K, 0, ovf, ovf_n, mom};
    pj->solve = SolveArgs{c->hdr, mom, mom_g, rg, c->nrm, c->grad, grid_attrs ? (const int32_t *)G.order : nullptr,
                          grid_attrs ? G.snrm : nullptr, grid_attrs ? G.sgrad : nullptr, ovf_n};
    pj->nb_pair = grid_for(cap, kNbBlock / gs, 4096);
    pj->nb_ovf = K > 0 ? grid_for(cap, kWavesPerBlock, 1024) : 0;
    pj->nb_solve = grid_for(cap, kNbBlock, 2048);
    return GCA_OK;
}

// moments (one walk), overflow K-selection (jobs with K > 0), solve: each kernel once for all nj
// calls (blockIdx.y = call), on stream.  A job without K selection contributes no overflow
// workgroups' work: its overflow list stays empty (count zero), so its rows of the overflow
// launch find nothing.
template <int N>
static int32_t pair_launch(hipStream_t stream, const PairJobs *pj, int nj)
{
    Jobs<PairArgs, N> pair{};
    Jobs<OvfArgs, N> ovf{};
    Jobs<SolveArgs, N> solve{};
    unsigned nb_pair = 1, nb_ovf = 0, nb_solve = 1;
    for (int j = 0; j < nj; ++j) {
        pair.a[j] = pj[j].pair;
        ovf.a[j] = pj[j].ovf;
        solve.a[j] = pj[j].solve;
        nb_pair = std::max(nb_pair, pj[j].nb_pair);
        nb_ovf = std::max(nb_ovf, pj[j].nb_ovf);
        nb_solve = std::max(nb_solve, pj[j].nb_solve);
    }
    hipLaunchKernelGGL(k_moments_pair<N>, dim3(nb_pair, (unsigned)nj), dim3(kNbBlock), 0, stream, pair);
    if (nb_ovf > 0)
        hipLaunchKernelGGL((k_moments_overflow<0, N>), dim3(nb_ovf, (unsigned)nj), dim3(kWave * kWavesPerBlock), 0,
                           stream, ovf);
    hipLaunchKernelGGL(k_solve_pair<N>, dim3(nb_solve, (unsigned)nj), dim3(kNbBlock), 0, stream, solve);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

int32_t launch_normals_gradient(gca_ctx ctx, gca_cloud c, float radius, int32_t K, float gradient_radius,
                                bool grid_attrs)
{
    PairJobs pj;
    GCA_TRY(pair_prepare(ctx, c, radius, K, gradient_radius, grid_attrs, &pj));
    {
        StageTimer timer(ctx, GCA_STAGE_NORMALS);
        GCA_TRY(pair_launch<1>(ctx->stream, &pj, 1));
    }
    if (grid_attrs) c->grid.attrs_valid = true;
    return GCA_OK;
}

int32_t launch_normals_gradient_levels(gca_ctx launch, const NormalsLevelJob *jobs, int nj)
{
    if (nj < 1 || nj > kMaxJobs) return GCA_ERR_INVALID_VALUE;
    PairJobs pj[kMaxJobs];
    for (int j = 0; j < nj; ++j)
        GCA_TRY(pair_prepare(jobs[j].ctx, jobs[j].c, jobs[j].radius, jobs[j].max_nn, jobs[j].gradient_radius, true,
                             &pj[j]));
    {
        StageTimer timer(launch, GCA_STAGE_NORMALS);
        GCA_TRY(pair_launch<kMaxJobs>(launch->stream, pj, nj));
    }
    for (int j = 0; j < nj; ++j) jobs[j].c->grid.attrs_valid = true;
    return GCA_OK;
}

}  // namespace gca
