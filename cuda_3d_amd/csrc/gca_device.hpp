// gca_device.hpp -- device-side building blocks shared by the kernels (wave64 reductions and
// scans, fixed-point conversion, grid cell arithmetic, small FP64 solvers).
#pragma once

#include <hip/hip_runtime.h>

#include "gca_internal.hpp"

namespace gca {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Device-scope ordering for the last-block (ticket) pattern.  __threadfence() is a seq_cst
// fence, which on gfx950 both writes back AND invalidates the XCD's L2: issued by every block
// it evicts the data the other blocks are still reading.  Writers only need release
// (write-back); only the block that wins the ticket needs acquire (invalidate).
__device__ __forceinline__ void fence_release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// ---- wave64 scans / reductions ----------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t t = __shfl_up(v, d, kWave);
        if (lane_id() >= d) v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, kWave));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, kWave));
    return v;
}

// Exclusive scan over a block of BLOCK threads (BLOCK a multiple of 64).  `lds` holds
// BLOCK/64 + 1 uint32.  Returns the exclusive prefix; *block_total receives the sum.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds, uint32_t *block_total)
{
    constexpr int NW = BLOCK / kWave;
    const int w = threadIdx.x / kWave;
    uint32_t inc = wave_incl_scan(v);
    if (lane_id() == kWave - 1) lds[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < NW; ++i) {
            uint32_t t = lds[i];
            lds[i] = run;
            run += t;
        }
        lds[NW] = run;
    }
    __syncthreads();
    uint32_t res = lds[w] + inc - v;
    *block_total = lds[NW];
    __syncthreads();
    return res;
}

// ---- FP32 TwoSum-compensated accumulator -------------------------------------------------
struct Comp {
    float s, c;
    __device__ __forceinline__ void add(float x)
    {
        // Knuth TwoSum: exact error of s + x, no branch (valid for any magnitudes).
        float t = s + x;
        float bp = t - s;
        float e = (s - (t - bp)) + (x - bp);
        s = t;
        c += e;
    }
};

// ---- fixed point ------------------------------------------------------------------------
__device__ __forceinline__ long long fx(double v) { return __double2ll_rn(v); }

// fx for |v| < 2^51: adding 1.5 * 2^52 rounds v to an integer (to nearest, ties to even: the
// same rounding as fx), which then sits in the low mantissa bits; two instructions instead of
// the general conversion's six.  Only for values bounded by construction.
__device__ __forceinline__ long long fx_small(double v)
{
    const double t = __dadd_rn(v, 6755399441055744.0);  // 1.5 * 2^52
    return __double_as_longlong(t) - 0x4338000000000000ll;
}

// ---- grid --------------------------------------------------------------------------------
__device__ __forceinline__ int32_t grid_cell(double p, float o, float s, int32_t dim)
{
    double c = floor((p - (double)o) / (double)s);
    c = fmax(c, 0.0);
    c = fmin(c, (double)(dim - 1));
    return (int32_t)c;
}

// Cell range overlapped by the box [q - r, q + r] (r in FP64); false if empty or q not finite.
__device__ __forceinline__ bool grid_range_r(const GridDesc &g, float qx, float qy, float qz, double r,
                                             int32_t lo[3], int32_t hi[3])
{
    const float q[3] = {qx, qy, qz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!isfinite(q[a])) return false;
        double l = floor(((double)q[a] - r - (double)g.origin[a]) / (double)g.cell);
        double h = floor(((double)q[a] + r - (double)g.origin[a]) / (double)g.cell);
        l = fmax(l, 0.0);
        h = fmin(h, (double)(g.dims[a] - 1));
        if (l > h) return false;
        lo[a] = (int32_t)l;
        hi[a] = (int32_t)h;
    }
    return true;
}

// Cell range overlapped by [q - g.radius, q + g.radius].
__device__ __forceinline__ bool grid_range(const GridDesc &g, float qx, float qy, float qz,
                                           int32_t lo[3], int32_t hi[3])
{
    return grid_range_r(g, qx, qy, qz, (double)g.radius, lo, hi);
}

__device__ __forceinline__ double dist2(float px, float py, float pz, float qx, float qy, float qz)
{
    double dx = (double)px - (double)qx, dy = (double)py - (double)qy, dz = (double)pz - (double)qz;
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// lexicographic (d2, idx) order
__device__ __forceinline__ bool key_less(double d2a, int32_t ia, double d2b, int32_t ib)
{
    return d2a < d2b || (d2a == d2b && ia < ib);
}

// ---- rigid transforms --------------------------------------------------------------------
__device__ __forceinline__ void euler_to_T(const double xi[6], double T[16])
{
    double sa, ca, sb, cb, sc, cc;
    sincos(xi[0], &sa, &ca);
    sincos(xi[1], &sb, &cb);
    sincos(xi[2], &sc, &cc);
    T[0] = cc * cb;
    T[1] = cc * sb * sa - sc * ca;
    T[2] = cc * sb * ca + sc * sa;
    T[3] = xi[3];
    T[4] = sc * cb;
    T[5] = sc * sb * sa + cc * ca;
    T[6] = sc * sb * ca - cc * sa;
    T[7] = xi[4];
    T[8] = -sb;
    T[9] = cb * sa;
    T[10] = cb * ca;
    T[11] = xi[5];
    T[12] = T[13] = T[14] = 0.0;
    T[15] = 1.0;
}

__device__ __forceinline__ void matmul4(const double A[16], const double B[16], double C[16])
{
    double t[16];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            t[4 * i + j] = ((A[4 * i] * B[j] + A[4 * i + 1] * B[4 + j]) + A[4 * i + 2] * B[8 + j]) +
                           A[4 * i + 3] * B[12 + j];
#pragma unroll
    for (int k = 0; k < 16; ++k) C[k] = t[k];
}

// LDL^T of the 6x6 normal equations (upper triangle, row-major) and xi = -A^-1 b.
// Returns false (xi = 0) if a pivot is not positive and finite.
__device__ __forceinline__ bool solve6(const double *a21, const double *b6, double xi[6])
{
    double A[6][6], L[6][6], D[6];
    int k = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int b = a; b < 6; ++b) {
            A[a][b] = a21[k];
            A[b][a] = a21[k];
            ++k;
        }
    bool ok = true;
    for (int j = 0; j < 6; ++j) {
        double dj = A[j][j];
        for (int m = 0; m < j; ++m) dj -= L[j][m] * L[j][m] * D[m];
        if (!(dj > 0.0) || !isfinite(dj)) {
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
            double v = -b6[i];
            for (int m = 0; m < i; ++m) v -= L[i][m] * y[m];
            y[i] = v;
        }
        for (int i = 0; i < 6; ++i) y[i] /= D[i];
        for (int i = 5; i >= 0; --i) {
            double v = y[i];
            for (int m = i + 1; m < 6; ++m) v -= L[m][i] * xi[m];
            xi[i] = v;
        }
        for (int i = 0; i < 6; ++i)
            if (!isfinite(xi[i])) ok = false;
    }
    if (!ok)
        for (int i = 0; i < 6; ++i) xi[i] = 0.0;
    return ok;
}

__device__ __forceinline__ float intensity_of(float r, float g, float b)
{
    // color3::to_intensity (type.hpp:45-51): FP64 products and sums, no contraction
    return (float)__dadd_rn(__dadd_rn(__dmul_rn(0.2126, (double)r), __dmul_rn(0.7152, (double)g)),
                            __dmul_rn(0.0722, (double)b));
}

}  // namespace gca
