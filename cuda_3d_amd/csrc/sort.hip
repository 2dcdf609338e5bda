// sort.hip -- LSD radix sort of 64-bit keys with 32-bit values (stable), for the sorted voxel path.
//
// One pass per 8-bit digit: a per-tile histogram (tile = 2048 keys, one 256-thread workgroup),
// an exclusive scan of the digit-major histogram table (one-pass look-back scan), and a stable
// scatter.  In the scatter each wavefront takes a quarter of the tile in key order, 64 keys at a
// time; a key's rank among its wavefront's keys of the same digit comes from a match of the
// digit across lanes (eight ballots), so equal digits keep their input order: within a chunk by
// lane, across chunks by the wavefront's running counters, across wavefronts by their counts,
// across tiles by the scanned table.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

constexpr int kSortBlock = 256;
constexpr int kSortPer = 8;                          // keys per thread
constexpr int kSortTile = kSortBlock * kSortPer;     // 2048 keys per workgroup
constexpr int kSortWaves = kSortBlock / kWave;       // 4
constexpr int kSortChunks = kSortTile / kSortWaves / kWave;  // 8 chunks of 64 per wavefront

__device__ __forceinline__ uint32_t digit_of(unsigned long long k, int shift) { return (uint32_t)(k >> shift) & 255u; }

// Lanes of the wavefront holding the same digit as this lane (among active lanes).
__device__ __forceinline__ unsigned long long match_digit(uint32_t d, bool active)
{
    unsigned long long m = __ballot(active);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const unsigned long long s = __ballot(active && ((d >> b) & 1u));
        m &= ((d >> b) & 1u) ? s : ~s;
    }
    return active ? m : 0ull;
}

__global__ __launch_bounds__(kSortBlock) void k_rs_hist(const unsigned long long *__restrict__ keys, int64_t n, int shift,
                                                       int ntiles, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int k = 0; k < kSortPer; ++k) {
        const int64_t i = base + (int64_t)k * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[digit_of(keys[i], shift)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];  // digit-major
}

__global__ __launch_bounds__(kSortBlock) void k_rs_scatter(const unsigned long long *__restrict__ kin,
                                                          const uint32_t *__restrict__ vin, int64_t n, int shift,
                                                          int ntiles, const uint32_t *__restrict__ offs,
                                                          unsigned long long *__restrict__ kout,
                                                          uint32_t *__restrict__ vout)
{
    __shared__ uint32_t cnt[kSortWaves][256];
    __shared__ uint32_t base[kSortWaves][256];
    const int w = threadIdx.x / kWave, lane = lane_id();
    for (int d = lane; d < 256; d += kWave) cnt[w][d] = 0u;
    __syncthreads();
    const int64_t wbase = (int64_t)blockIdx.x * kSortTile + (int64_t)w * (kSortChunks * kWave);
    const unsigned long long below = (1ull << lane) - 1ull;
    // 1. per-wavefront digit counts
    for (int c = 0; c < kSortChunks; ++c) {
        const int64_t i = wbase + (int64_t)c * kWave + lane;
        const bool on = i < n;
        const uint32_t d = on ? digit_of(kin[i], shift) : 0u;
        const unsigned long long peers = match_digit(d, on);
        if (on && (peers & below) == 0ull) cnt[w][d] += (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();  // the next chunk's leaders read what this one wrote
    }
    __syncthreads();
    // 2. where each wavefront's keys of digit d start
    {
        const int d = threadIdx.x;  // 256 threads, one digit each
        uint32_t run = offs[(int64_t)d * ntiles + blockIdx.x];
#pragma unroll
        for (int v = 0; v < kSortWaves; ++v) {
            base[v][d] = run;
            run += cnt[v][d];
        }
    }
    __syncthreads();
    // 3. scatter, in key order within the wavefront
    for (int c = 0; c < kSortChunks; ++c) {
        const int64_t i = wbase + (int64_t)c * kWave + lane;
        const bool on = i < n;
        const unsigned long long k = on ? kin[i] : 0ull;
        const uint32_t d = digit_of(k, shift);
        const unsigned long long peers = match_digit(d, on);
        if (on) {
            const uint32_t dst = base[w][d] + (uint32_t)__popcll(peers & below);
            kout[dst] = k;
            vout[dst] = vin[i];
        }
        if (on && (peers & below) == 0ull) base[w][d] += (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
}

// Sorts n (key, value) pairs by the low `bits` bits of the key (stable); the result ends in
// (k0, v0) or (k1, v1) as returned in *sorted_in_first.  k1 / v1: scratch of n entries.
int32_t radix_sort_u64(gca_ctx ctx, unsigned long long *k0, uint32_t *v0, unsigned long long *k1, uint32_t *v1,
                       int64_t n, int bits, bool *sorted_in_first)
{
    *sorted_in_first = true;
    if (n <= 1 || bits <= 0) return GCA_OK;
    const int64_t ntiles = (n + kSortTile - 1) / kSortTile;
    if (ntiles > INT32_MAX / 256) return GCA_ERR_INVALID_VALUE;
    uint32_t *hist = (uint32_t *)ctx_scratch(ctx, 38, sizeof(uint32_t) * (size_t)(256 * ntiles), false);
    uint32_t *offs = (uint32_t *)ctx_scratch(ctx, 39, sizeof(uint32_t) * (size_t)(256 * ntiles + 1), false);
    if (!hist || !offs) return GCA_ERR_OUT_OF_MEMORY;
    unsigned long long *ka = k0, *kb = k1;
    uint32_t *va = v0, *vb = v1;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(kSortBlock), 0, ctx->stream, ka, n, shift,
                           (int)ntiles, hist);
        GCA_HIP(hipGetLastError());
        GCA_TRY(exclusive_scan_u32(ctx, hist, offs, 256 * ntiles, nullptr));
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(kSortBlock), 0, ctx->stream, ka, va, n, shift,
                           (int)ntiles, offs, kb, vb);
        GCA_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
        *sorted_in_first = !*sorted_in_first;
    }
    return GCA_OK;
}

}  // namespace gca
