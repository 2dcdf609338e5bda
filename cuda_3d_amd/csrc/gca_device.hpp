// gca_device.hpp -- device-side building blocks shared by the kernels (wave64 reductions and
// scans, fixed-point conversion, grid cell arithmetic, small FP64 solvers).
#pragma once

#include <hip/hip_runtime.h>

#include "gca_internal.hpp"

namespace gca {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Workgroups are dealt to the device's 8 XCDs round-robin (linear workgroup id mod 8), and each
// XCD has its own L2.  xcd_runs<G>(b, m) permutes [0, m) so that each XCD takes runs of G
// consecutive logical blocks, the runs dealt to the XCDs in turn: neighbouring blocks (which
// read the same neighbourhoods) share an L2 while the work stays spread over every XCD.  Blocks
// beyond the last whole round of 8 runs keep their place.
constexpr int kXcds = 8;
template <int G>
__device__ __forceinline__ int xcd_runs(int b, int m)
{
    const int whole = m / (kXcds * G) * (kXcds * G);
    if (b >= whole) return b;
    const int x = b % kXcds, i = b / kXcds;
    return ((i / G) * kXcds + x) * G + i % G;
}

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

// ---- runs of equal keys in a wavefront ---------------------------------------------------
// Segment id of each lane: the lane index of the first lane of its run of equal keys (runs
// are maximal sequences of consecutive lanes with the same key).
__device__ __forceinline__ int run_head(int64_t key)
{
    const int lane = lane_id();
    const int64_t prev = __shfl_up(key, 1, kWave);
    int head = (lane == 0 || prev != key) ? lane : 0;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(head, d, kWave);
        if (lane >= d) head = max(head, o);
    }
    return head;
}

// In-run suffix reduction: afterwards the run head holds the run's total.
template <typename T, typename Op>
__device__ __forceinline__ T run_reduce(T v, int head, Op op)
{
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T o = __shfl_down(v, d, kWave);
        const int oh = __shfl_down(head, d, kWave);
        if (lane + d < kWave && oh == head) v = op(v, o);
    }
    return v;
}

// ---- decoupled look-back of one-pass scans ----------------------------------------------
// Chunks (workgroups) are handed out in start order by a counter, so a chunk only ever waits on
// chunks whose workgroups are already running.  Each chunk publishes one 64-bit word per call:
// (call sequence << 34) | (state << 32) | value, first with its own total (state 1), then, once the
// totals before it are known, with its inclusive prefix (state 2).  Wavefront 0 looks back over 64
// predecessor words at a time and stops at the nearest inclusive prefix.  Flag, value and sequence
// travel in one word, so no ordering between separate stores is needed, and words left by earlier
// calls carry an older sequence: nothing is cleared between calls (graph-capture safe).  A wait
// that outlasts kScanSpin polls stops and raises the device error word; with chunks taken in start
// order it cannot happen unless the device stops making progress.
constexpr uint32_t kScanSpin = 1u << 22;
constexpr unsigned long long kScanAgg = 1ull << 32, kScanInc = 2ull << 32;
constexpr uint32_t kScanSeqMask = (1u << 30) - 1u;

struct ScanCtl {
    unsigned long long ctr;  // (sequence << 32) | chunks taken
};

// Run by the 64 lanes of wavefront 0 of chunk c (sequence seq, chunk total tot): publishes the
// chunk's words and returns its exclusive prefix (on every lane).
__device__ __forceinline__ uint32_t lookback_exclusive(unsigned long long *flags, int c, uint32_t seq, uint32_t tot,
                                                       uint32_t *dev_err)
{
    const int lane = lane_id();
    const unsigned long long tag = (unsigned long long)(seq & kScanSeqMask) << 34;
    uint32_t excl = 0;
    if (c == 0) {
        if (lane == 0) __hip_atomic_store(&flags[0], tag | kScanInc | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (lane == 0) __hip_atomic_store(&flags[c], tag | kScanAgg | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int end = c;  // window: chunks end - 64 .. end - 1, lane j reads chunk end - 1 - j
    uint32_t spins = 0;
    for (;;) {
        const int k = end - 1 - lane;
        const unsigned long long w =
            k >= 0 ? __hip_atomic_load(&flags[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag | kScanInc);
        const uint32_t st = (uint32_t)(w >> 32) & 3u;
        const bool valid = (w >> 34) == (seq & kScanSeqMask) && st != 0u;
        const unsigned long long inc = __ballot(valid && st == 2u);
        const int first = inc ? __builtin_ctzll(inc) : kWave;  // nearest inclusive prefix
        const unsigned long long need = first >= kWave - 1 ? ~0ull : ((2ull << first) - 1ull);
        if (__ballot(!valid) & need) {
            if (++spins > kScanSpin) {
                if (lane == 0) __hip_atomic_store(dev_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += wave_sum(lane <= first ? (uint32_t)w : 0u);
        if (first < kWave) break;
        end -= kWave;
    }
    if (lane == 0)
        __hip_atomic_store(&flags[c], tag | kScanInc | (uint32_t)(excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
}

// Workgroup-wide variant of the look-back, for kernels with many chunks that all start together
// (every chunk resident at once publishes its total at about the same time, so the look-back
// walks over totals rather than stopping at an early inclusive prefix): all BLOCK threads read
// BLOCK predecessor words per round, so a walk over C chunks takes C / BLOCK rounds instead of
// C / 64.  Same words and protocol as lookback_exclusive; called by every thread of the
// workgroup (the result is on every thread).  lds: 4 * (BLOCK / 64) words of 64 bits.
template <int BLOCK>
__device__ __forceinline__ uint32_t lookback_exclusive_block(unsigned long long *flags, int c, uint32_t seq,
                                                             uint32_t tot, uint32_t *dev_err,
                                                             unsigned long long *lds)
{
    constexpr int NW = BLOCK / kWave;
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const unsigned long long tag = (unsigned long long)(seq & kScanSeqMask) << 34;
    if (c == 0) {
        if (threadIdx.x == 0)
            __hip_atomic_store(&flags[0], tag | kScanInc | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0;
    }
    if (threadIdx.x == 0) __hip_atomic_store(&flags[c], tag | kScanAgg | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0, spins = 0;
    int end = c;  // window: chunks end - BLOCK .. end - 1, thread t reads chunk end - 1 - t
    for (int round = 0;; ++round) {
        unsigned long long *buf = lds + (round & 1) * 2 * NW;  // [0, NW): masks, [NW, 2NW): sums
        const int k = end - 1 - (int)threadIdx.x;
        const unsigned long long w =
            k >= 0 ? __hip_atomic_load(&flags[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : (tag | kScanInc);
        const uint32_t st = (uint32_t)(w >> 32) & 3u;
        const bool valid = (w >> 34) == (seq & kScanSeqMask) && st != 0u;
        const unsigned long long inc = __ballot(valid && st == 2u);
        const unsigned long long bad = __ballot(!valid);
        const int f = inc ? __builtin_ctzll(inc) : kWave;  // this wavefront's nearest inclusive prefix
        const uint32_t s_all = wave_sum(valid ? (uint32_t)w : 0u);
        const uint32_t s_first = wave_sum(valid && lane <= f ? (uint32_t)w : 0u);
        if (lane == 0) {
            const unsigned long long need = f >= kWave - 1 ? ~0ull : ((2ull << f) - 1ull);
            // bit 0: an inclusive prefix in this wavefront; bit 1: a word not yet valid in the part
            // of this wavefront that would be summed
            buf[wv] = (inc ? 1ull : 0ull) | ((bad & need) ? 2ull : 0ull);
            buf[NW + wv] = ((unsigned long long)s_first << 32) | s_all;
        }
        __syncthreads();
        bool wait = false, done = false;
        uint32_t sum = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const unsigned long long m = buf[i], s = buf[NW + i];
            if (!done) {
                wait |= (m & 2ull) != 0ull;
                if (m & 1ull) {
                    sum += (uint32_t)(s >> 32);
                    done = true;
                } else {
                    sum += (uint32_t)s;
                }
            }
        }
        if (wait) {
            if (++spins > kScanSpin) {
                if (threadIdx.x == 0) __hip_atomic_store(dev_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        excl += sum;
        if (done) break;
        end -= BLOCK;
    }
    if (threadIdx.x == 0)
        __hip_atomic_store(&flags[c], tag | kScanInc | (uint32_t)(excl + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return excl;
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
