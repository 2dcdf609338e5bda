// voxel.hip -- voxel-grid downsample (K3).
//
// Replaces cuda_voxel_grid_downsample (src/geometry/cuda_voxel_grid_down_sample.cu:123-161)
// and point_cloud::voxel_grid_down_sample (src/geometry/point_cloud.cu:198-229).
//
// The reference sorts int3 keys with a comparison sort and reduces by key.  Here the output
// order (lexicographic (kx, ky, kz) == ascending linear key) comes, for lattices up to 2^34
// cells, from an occupancy bitmap over the voxel lattice instead of a sort (beyond that, and by
// tuning, from a radix sort of packed keys and a segmented reduce: the sorted path below):
//   1. mark: the lattice from the input's device header (every workgroup for itself), then
//      the key in the reference's FP32 formula, set bit L = (kx*ny + ky)*nz + kz;
//   2. popcount prefix over the bitmap words (one-pass scan with decoupled look-back) -> rank
//      of every occupied voxel in lexicographic order, and the voxel count M;
//   3. accumulate: each point adds its coordinates and colour in 2^-30 fixed point, first into
//      a workgroup's LDS table, then once per voxel and workgroup into int64 sums in global
//      memory (integer atomics: exact, so the sums do not depend on the order of arrival);
//   4. finalize: mean = sum / count, intensity from the mean colour; the bitmap words and
//      accumulators are cleared on the way out, so no memset is needed before the next call.
// The output keeps the input's bounding box, as the reference does (point_cloud.cu:224-226).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

constexpr double kVoxelMaxBits = 17179869184.0;  // 2^34 lattice cells

struct VoxelParams {
    float bmin[3];
    float voxel;
    int64_t dims[3];
    int32_t bad;  // lattice refused (extent / voxel beyond INT_MAX, or more bits than planned)
    int32_t pad;
};

// Lattice of a voxel call from the input's device header, so a launch sequence needs no host
// round trip: origin = unpadded min bound, dims = floor(extent / v) + 1 per axis in FP32
// (point_cloud.cu:205-218, cuda_voxel_grid_down_sample.cu:37-47).  A lattice the reference
// would refuse, or one with more than max_bits cells, is flagged bad (and yields an empty
// output).
__device__ __forceinline__ VoxelParams voxel_params_of(const CloudHdr *__restrict__ in, float voxel, double max_bits)
{
    const int n = in->n;
    VoxelParams p{};
    p.voxel = voxel;
    double nbits = 1.0;
    for (int a = 0; a < 3; ++a) {
        const float mn = key_to_float(in->bmin_key[a]), mx = key_to_float(in->bmax_key[a]);
        p.bmin[a] = mn;
        const float ext = __fsub_rn(mx, mn);
        if (n > 0 && (double)voxel * (double)INT_MAX < (double)ext) p.bad = 1;
        p.dims[a] = n > 0 ? (int64_t)floorf(__fdiv_rn(ext, voxel)) + 1 : 1;
        nbits *= (double)p.dims[a];
    }
    if (nbits > max_bits) p.bad = 1;
    if (p.bad) p.dims[0] = p.dims[1] = p.dims[2] = 1;
    return p;
}

// The voxel's integer coordinates: __float2int_rd((p - min) / v) in FP32
// (cuda_voxel_grid_down_sample.cu:37-47), clamped to the lattice.
__device__ __forceinline__ void voxel_key3(const VoxelParams &vp, float4 p, int64_t k[3])
{
    const float c[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float q = floorf(__fdiv_rn(__fsub_rn(c[a], vp.bmin[a]), vp.voxel));
        int64_t v = (int64_t)q;
        v = v < 0 ? 0 : (v >= vp.dims[a] ? vp.dims[a] - 1 : v);
        k[a] = v;
    }
}

__device__ __forceinline__ int64_t voxel_key(const VoxelParams &vp, float4 p)
{
    int64_t k[3];
    voxel_key3(vp, p, k);
    return (k[0] * vp.dims[1] + k[1]) * vp.dims[2] + k[2];
}

// Raster-ordered points hit the same voxel in runs, so each wavefront first merges runs and
// only the first lane of a run issues the atomics.  Every workgroup derives the lattice from the
// header itself; workgroup 0 also stores it for the later kernels of the call and flags a
// refused lattice (*err).
struct MarkArgs {
    const float4 *xyzi;
    const CloudHdr *hdr;
    float voxel;
    double max_bits;
    VoxelParams *vp_out;
    uint32_t *err, *bitmap;
};

template <int N>
__global__ __launch_bounds__(256) void k_voxel_mark(Jobs<MarkArgs, N> J)
{
    const MarkArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    const float4 *__restrict__ xyzi = A.xyzi;
    const CloudHdr *__restrict__ hdr = A.hdr;
    uint32_t *__restrict__ bitmap = A.bitmap;
    __shared__ VoxelParams s_vp;
    if (threadIdx.x == 0) {
        s_vp = voxel_params_of(hdr, A.voxel, A.max_bits);
        if (blockIdx.x == 0) {
            *A.vp_out = s_vp;
            if (s_vp.bad && A.err) atomicOr(A.err, 1u);
        }
    }
    __syncthreads();
    const VoxelParams vp = s_vp;
    const int n = vp.bad ? 0 : hdr->n;
    const int stride = gridDim.x * blockDim.x;
    for (int b = blockIdx.x * blockDim.x; b < n; b += stride) {  // block-uniform trip count
        const int i = b + threadIdx.x;
        int64_t L = -1 - (int64_t)lane_id();  // unique keys for inactive lanes
        if (i < n) L = voxel_key(vp, xyzi[i]);
        const int64_t word = L >= 0 ? (L >> 5) : L;
        const int head = run_head(word);
        uint32_t bits = L >= 0 ? (1u << (uint32_t)(L & 31)) : 0u;
        bits = run_reduce(bits, head, [](uint32_t a, uint32_t c) { return a | c; });
        if (L >= 0 && head == lane_id()) atomicOr(&bitmap[word], bits);
    }
}

// Accumulation with workgroup pre-aggregation.  Each point's voxel key is recomputed from the
// point (the marking pass stores none: the point is read here anyway, and the key costs a few
// FP32 operations against 16 bytes written and read back per point).  A workgroup takes
// kVoxChunk consecutive points (raster order: a few image rows, whose points share few voxels)
// and sums their fixed-point values per voxel in an LDS hash table keyed by the voxel's rank
// (open addressing, linear probing); it then hands each distinct voxel's sums to global memory
// once.  Device-scope 64-bit
// atomics bound this kernel, so the hand-over takes one compare-and-swap on the voxel's claim
// word: the workgroup that wins it stores its sums plainly into the voxel's "A" record, the
// others add theirs atomically into the "B" record (a point whose probe sequence found no slot
// does the same with its own values).  Raster chunks of 2048 points meet 1.3 / 1.9 / 3.0
// workgroups per voxel at 1 / 2 / 4 cm on a 640x480 frame, so most voxels take one CAS instead
// of seven atomics.  Integer sums are exact, so neither the arrival order nor the split into A
// and B changes the result.  Every voxel's A record is written by its claimant, so only B and
// the claim need clearing after use.
// Per-voxel record (kVoxRec words): A sums [0, 7), lattice key [7], B sums [8, 15), claim [15].
constexpr int kVoxRec = 16;
constexpr int kVoxChunk = 2048;
constexpr int kVoxBlock = 1024;  // 2 points per thread: short per-thread chains, many waves per chunk
constexpr int kVoxTableBits = 10;
constexpr int kVoxTable = 1 << kVoxTableBits;
constexpr int kVoxProbe = 16;

struct AccumArgs {
    const float4 *xyzi, *rgb;
    const CloudHdr *hdr;
    const VoxelParams *vpp;
    const uint32_t *bitmap, *word_pre;
    unsigned long long *acc;
};

template <int N>
__global__ __launch_bounds__(kVoxBlock) void k_voxel_accum(Jobs<AccumArgs, N> J)
{
    const AccumArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    const float4 *__restrict__ xyzi = A.xyzi;
    const float4 *__restrict__ rgb = A.rgb;
    const uint32_t *__restrict__ bitmap = A.bitmap;
    const uint32_t *__restrict__ word_pre = A.word_pre;
    unsigned long long *__restrict__ acc = A.acc;
    __shared__ int32_t t_rank[kVoxTable];
    __shared__ long long t_key[kVoxTable];
    __shared__ unsigned long long t_sum[7][kVoxTable];
    __shared__ VoxelParams s_vp;
    if (threadIdx.x == 0) s_vp = *A.vpp;
    for (int h = threadIdx.x; h < kVoxTable; h += blockDim.x) {
        t_rank[h] = -1;
#pragma unroll
        for (int k = 0; k < 7; ++k) t_sum[k][h] = 0ull;
    }
    __syncthreads();
    const VoxelParams vp = s_vp;
    const int n = vp.bad ? 0 : A.hdr->n;
    for (int base = blockIdx.x * kVoxChunk; base < n; base += gridDim.x * kVoxChunk) {  // block-uniform
        __syncthreads();  // table clear (initial, or by the previous chunk's flush)
        const int end = min(base + kVoxChunk, n);
        for (int i = base + (int)threadIdx.x; i < end; i += blockDim.x) {
            const float4 p = xyzi[i], c = rgb[i];
            const int64_t L = voxel_key(vp, p);  // the marking pass's key, recomputed
            const uint32_t w = bitmap[L >> 5];
            const uint32_t below = w & ((1u << (uint32_t)(L & 31)) - 1u);
            const int32_t r = (int32_t)(word_pre[L >> 5] + (uint32_t)__popc(below));
            unsigned long long v[7];
            v[0] = (unsigned long long)fx((double)p.x * kFx30);
            v[1] = (unsigned long long)fx((double)p.y * kFx30);
            v[2] = (unsigned long long)fx((double)p.z * kFx30);
            v[3] = (unsigned long long)fx((double)c.x * kFx30);
            v[4] = (unsigned long long)fx((double)c.y * kFx30);
            v[5] = (unsigned long long)fx((double)c.z * kFx30);
            v[6] = 1ull;
            uint32_t h = ((uint32_t)r * 2654435761u) >> (32 - kVoxTableBits);  // multiplicative hash
            int slot = -1;
            for (int t = 0; t < kVoxProbe; ++t) {
                const int32_t old = atomicCAS(&t_rank[h], -1, r);
                if (old == -1) t_key[h] = L;  // same value from every point of the voxel
                if (old == -1 || old == r) {
                    slot = (int)h;
                    break;
                }
                h = (h + 1) & (kVoxTable - 1);
            }
            if (slot >= 0) {
#pragma unroll
                for (int k = 0; k < 7; ++k) atomicAdd(&t_sum[k][slot], v[k]);
            } else {
                unsigned long long *a = acc + kVoxRec * (int64_t)r;
                const bool own = atomicCAS(&a[15], 0ull, 1ull) == 0ull;
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    if (own) a[k] = v[k];
                    else atomicAdd(&a[8 + k], v[k]);
                }
                a[7] = (unsigned long long)L;
            }
        }
        __syncthreads();
        for (int h = threadIdx.x; h < kVoxTable; h += blockDim.x) {
            const int32_t r = t_rank[h];
            if (r < 0) continue;
            unsigned long long *a = acc + kVoxRec * (int64_t)r;
            const bool own = atomicCAS(&a[15], 0ull, 1ull) == 0ull;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                if (own) a[k] = t_sum[k][h];
                else atomicAdd(&a[8 + k], t_sum[k][h]);
                t_sum[k][h] = 0ull;
            }
            a[7] = (unsigned long long)t_key[h];
            t_rank[h] = -1;
        }
    }
}

// gc_count != nullptr: each output point is also counted into its cell of the grid gc (and its
// cell recorded), which is the first kernel of that grid's build.
struct FinalArgs {
    unsigned long long *acc;
    const CloudHdr *in_hdr;
    CloudHdr *out_hdr;
    uint32_t *bitmap;
    float4 *out_xyzi, *out_rgb;
    GridDesc gd;
    uint32_t *gc_count, *gc_cell_of;
};

template <int N>
__global__ __launch_bounds__(256) void k_voxel_finalize(Jobs<FinalArgs, N> J)
{
    const FinalArgs &A = J.a[N == 1 ? 0 : blockIdx.y];
    unsigned long long *__restrict__ acc = A.acc;
    const CloudHdr *__restrict__ in_hdr = A.in_hdr;
    CloudHdr *__restrict__ out_hdr = A.out_hdr;
    uint32_t *__restrict__ bitmap = A.bitmap;
    float4 *__restrict__ out_xyzi = A.out_xyzi;
    float4 *__restrict__ out_rgb = A.out_rgb;
    const GridDesc &gd = A.gd;
    uint32_t *__restrict__ gc_count = A.gc_count;
    uint32_t *__restrict__ gc_cell_of = A.gc_cell_of;
    const int m = out_hdr->n;
    if (blockIdx.x == 0 && threadIdx.x < 3) {
        out_hdr->bmin_key[threadIdx.x] = in_hdr->bmin_key[threadIdx.x];
        out_hdr->bmax_key[threadIdx.x] = in_hdr->bmax_key[threadIdx.x];
    }
    for (int b = blockIdx.x * blockDim.x; b < m; b += gridDim.x * blockDim.x) {  // block-uniform
        const int r = b + (int)threadIdx.x;
        const bool on = r < m;
        int64_t cell = -1 - (int64_t)lane_id();  // unique keys for inactive lanes
        if (on) {
            unsigned long long *a = acc + kVoxRec * (int64_t)r;
            long long sum[7];
#pragma unroll
            for (int c = 0; c < 7; ++c) sum[c] = (long long)(a[c] + a[8 + c]);  // A + B (two's complement)
            const double cnt = (double)sum[6];
            float o[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) o[c] = (float)(((double)sum[c] / cnt) * (1.0 / kFx30));
            out_xyzi[r] = make_float4(o[0], o[1], o[2], intensity_of(o[3], o[4], o[5]));
            out_rgb[r] = make_float4(o[3], o[4], o[5], 0.0f);
            if (gc_count) {  // the grid's count pass for this point (same cell rule as k_grid_count)
                const int32_t cx = grid_cell(o[0], gd.origin[0], gd.cell, gd.dims[0]);
                const int32_t cy = grid_cell(o[1], gd.origin[1], gd.cell, gd.dims[1]);
                const int32_t cz = grid_cell(o[2], gd.origin[2], gd.cell, gd.dims[2]);
                const uint32_t lin = (uint32_t)(((int64_t)cx * gd.dims[1] + cy) * gd.dims[2] + cz);
                gc_cell_of[r] = lin;
                cell = lin;
            }
            if (bitmap) {
                const int64_t L = (int64_t)a[7];
                bitmap[L >> 5] = 0u;
            }
#pragma unroll
            for (int c = 8; c < kVoxRec; ++c) a[c] = 0ull;  // B and the claim (A is always rewritten)
        }
        if (gc_count) {
            // voxels come out in (kx, ky, kz) order, so neighbours often share a grid cell: one
            // atomic per run of equal cells in the wavefront
            const int head = run_head(cell);
            const uint32_t len = run_reduce(on ? 1u : 0u, head, [](uint32_t x, uint32_t y) { return x + y; });
            if (on && head == lane_id()) atomicAdd(&gc_count[cell], len);
        }
    }
}

// ---- sorted path: radix sort of packed keys, then a segmented reduce -------------------------
// For lattices beyond the bitmap's reach (or by tuning).  Key of a point: (kx << (by + bz)) |
// (ky << bz) | kz, with bx, by, bz the bit widths of dims - 1, so ascending keys are the
// reference's lexicographic (kx, ky, kz) order.  Widths beyond 64 bits in all are sorted in two
// stable passes: by (ky << bz) | kz, then by kx.  Equal keys form segments; a segment's rank
// (its output position) is the count of segment starts before it.  The sums are the same
// fixed-point integers as the bitmap path's, so both paths give the same bits.
__global__ __launch_bounds__(256) void k_vsort_key(const float4 *__restrict__ xyzi, int n, VoxelParams vp,
                                                   int shift_x, int shift_y, int split,
                                                   unsigned long long *__restrict__ keys, uint32_t *__restrict__ idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t k[3];
    voxel_key3(vp, xyzi[i], k);
    const unsigned long long lo = ((unsigned long long)k[1] << shift_y) | (unsigned long long)k[2];
    keys[i] = split ? lo : (((unsigned long long)k[0] << shift_x) | lo);
    idx[i] = (uint32_t)i;
}

// Second key of a split sort: kx of the point at each sorted position.
__global__ __launch_bounds__(256) void k_vsort_rekey(const float4 *__restrict__ xyzi, const uint32_t *__restrict__ idx,
                                                     int n, VoxelParams vp, unsigned long long *__restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t k[3];
    voxel_key3(vp, xyzi[idx[i]], k);
    keys[i] = (unsigned long long)k[0];
}

// flag[i] = 1 where sorted position i starts a voxel.  A split sort leaves only kx in the keys, so
// it compares the full coordinates of the two points instead.
__global__ __launch_bounds__(256) void k_vsort_flag(const unsigned long long *__restrict__ keys,
                                                    const uint32_t *__restrict__ idx, const float4 *__restrict__ xyzi,
                                                    int n, VoxelParams vp, int split, uint32_t *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool head = i == 0 || keys[i] != keys[i - 1];
    if (!head && split) {
        int64_t a[3], b[3];
        voxel_key3(vp, xyzi[idx[i]], a);
        voxel_key3(vp, xyzi[idx[i - 1]], b);
        head = a[1] != b[1] || a[2] != b[2];
    }
    flag[i] = head ? 1u : 0u;
}

// Segmented reduce: each wavefront merges the runs of equal rank among its 64 sorted positions;
// the first lane of a run stores its sums into the voxel's A record when the whole voxel lies in
// this wavefront (one writer); a voxel that crosses wavefronts is handed over by claim, as in
// the bitmap path.
__global__ __launch_bounds__(256) void k_vsort_reduce(const float4 *__restrict__ xyzi, const float4 *__restrict__ rgb,
                                                      const uint32_t *__restrict__ idx,
                                                      const uint32_t *__restrict__ flag,
                                                      const uint32_t *__restrict__ seg, int n,
                                                      unsigned long long *__restrict__ acc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // wavefronts cover 64 consecutive positions
    const int lane = lane_id();
    const bool on = i < n;
    unsigned long long v[7] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    int64_t r = -1 - (int64_t)lane;  // unique keys for inactive lanes
    uint32_t f = 0u;
    if (on) {
        f = flag[i];
        r = (int64_t)seg[i] + (int64_t)f - 1;
        const uint32_t j = idx[i];
        const float4 p = xyzi[j], c = rgb[j];
        v[0] = (unsigned long long)fx((double)p.x * kFx30);
        v[1] = (unsigned long long)fx((double)p.y * kFx30);
        v[2] = (unsigned long long)fx((double)p.z * kFx30);
        v[3] = (unsigned long long)fx((double)c.x * kFx30);
        v[4] = (unsigned long long)fx((double)c.y * kFx30);
        v[5] = (unsigned long long)fx((double)c.z * kFx30);
        v[6] = 1ull;
    }
    const int head = run_head(r);
    const int64_t r_next = __shfl_down(r, 1, kWave);
    // does the voxel end inside this wavefront at this lane?
    uint32_t ends = 0u;
    if (on) {
        if (lane < kWave - 1 && i + 1 < n) ends = r_next != r ? 1u : 0u;
        else ends = (i + 1 >= n || flag[i + 1]) ? 1u : 0u;
    }
    ends = run_reduce(ends, head, [](uint32_t a, uint32_t c) { return a | c; });
#pragma unroll
    for (int k = 0; k < 7; ++k)
        v[k] = run_reduce(v[k], head, [](unsigned long long a, unsigned long long c) { return a + c; });
    if (on && head == lane) {
        unsigned long long *a = acc + kVoxRec * r;
        // a voxel inside this wavefront has one writer; a crossing one is claimed as in the
        // bitmap path (the claimant's A record, the others' B record)
        const bool own = (f && ends) || atomicCAS(&a[15], 0ull, 1ull) == 0ull;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if (own) a[k] = v[k];
            else atomicAdd(&a[8 + k], v[k]);
        }
    }
}

// Sorted-path kernel sequence of one voxel call, on a lattice planned on the host (vp).
static int32_t voxel_sort_enqueue(gca_ctx ctx, gca_cloud in, const VoxelParams &vp, gca_cloud out, int n)
{
    int w[3];
    for (int a = 0; a < 3; ++a)
        w[a] = vp.dims[a] > 1 ? 64 - __builtin_clzll((unsigned long long)(vp.dims[a] - 1)) : 0;
    const bool split = w[0] + w[1] + w[2] > 64;
    const int64_t nc = std::max(n, 1);
    auto *k0 = (unsigned long long *)ctx_scratch(ctx, 40, sizeof(unsigned long long) * nc, false);
    auto *k1 = (unsigned long long *)ctx_scratch(ctx, 41, sizeof(unsigned long long) * nc, false);
    auto *i0 = (uint32_t *)ctx_scratch(ctx, 42, sizeof(uint32_t) * nc, false);
    auto *i1 = (uint32_t *)ctx_scratch(ctx, 43, sizeof(uint32_t) * nc, false);
    auto *flag = (uint32_t *)ctx_scratch(ctx, 44, sizeof(uint32_t) * nc, false);
    auto *seg = (uint32_t *)ctx_scratch(ctx, 45, sizeof(uint32_t) * (nc + 1), false);
    auto *acc = (unsigned long long *)ctx_scratch(ctx, 7, sizeof(unsigned long long) * kVoxRec * nc, true);
    if (!k0 || !k1 || !i0 || !i1 || !flag || !seg || !acc) return GCA_ERR_OUT_OF_MEMORY;
    StageTimer timer(ctx, GCA_STAGE_VOXEL);
    const unsigned nb = (unsigned)((nc + 255) / 256);
    hipLaunchKernelGGL(k_vsort_key, dim3(nb), dim3(256), 0, ctx->stream, in->xyzi, n, vp, w[1] + w[2], w[2],
                       split ? 1 : 0, k0, i0);
    GCA_HIP(hipGetLastError());
    bool first = true;
    GCA_TRY(radix_sort_u64(ctx, k0, i0, k1, i1, n, split ? w[1] + w[2] : w[0] + w[1] + w[2], &first));
    unsigned long long *ks = first ? k0 : k1, *kt = first ? k1 : k0;
    uint32_t *is = first ? i0 : i1, *it = first ? i1 : i0;
    if (split) {
        hipLaunchKernelGGL(k_vsort_rekey, dim3(nb), dim3(256), 0, ctx->stream, in->xyzi, is, n, vp, ks);
        GCA_HIP(hipGetLastError());
        GCA_TRY(radix_sort_u64(ctx, ks, is, kt, it, n, w[0], &first));
        if (!first) {
            std::swap(ks, kt);
            std::swap(is, it);
        }
    }
    hipLaunchKernelGGL(k_vsort_flag, dim3(nb), dim3(256), 0, ctx->stream, ks, is, in->xyzi, n, vp, split ? 1 : 0,
                       flag);
    GCA_HIP(hipGetLastError());
    GCA_TRY(exclusive_scan_u32(ctx, flag, seg, n, (uint32_t *)&out->hdr->n));  // n = 0: M = 0
    hipLaunchKernelGGL(k_vsort_reduce, dim3(nb), dim3(256), 0, ctx->stream, in->xyzi, in->rgb, is, flag, seg, n, acc);
    GCA_HIP(hipGetLastError());
    const Jobs<FinalArgs, 1> fin{{{acc, in->hdr, out->hdr, nullptr, out->xyzi, out->rgb, GridDesc{}, nullptr, nullptr}}};
    hipLaunchKernelGGL(k_voxel_finalize<1>, dim3((unsigned)std::min<int64_t>(4096, nb)), dim3(256), 0, ctx->stream, fin);
    GCA_HIP(hipGetLastError());
    cloud_points_changed(out);
    return GCA_OK;
}

// The kernels of one bitmap-path voxel call, as jobs.
struct VoxelJobs {
    MarkArgs mark;
    ScanJob scan;
    AccumArgs accum;
    FinalArgs fin;
    unsigned nb, nb_acc;
};

// Host side of one voxel call: the ctx's bitmap and scratch for `n_cap` input points and a
// popcount scan over `words` bitmap words (>= the lattice's); enqueues nothing but a first-use
// clear of the bitmap on ctx's stream.
static int32_t voxel_prepare(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out, int64_t n_cap, int64_t words,
                             double max_bits, uint32_t *err, const GridCount *gc, bool wide, VoxelJobs *vj)
{
    if (ctx->bitmap_words < words || ctx->bitmap_dirty) {
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->bitmap_words < words) {
            hipFree(ctx->bitmap);
            ctx->bitmap = nullptr;
            ctx->bitmap_words = 0;
            int64_t want = std::max<int64_t>(words + words / 4, 1 << 16);
            GCA_HIP(hipMalloc(&ctx->bitmap, sizeof(uint32_t) * want));
            ctx->bitmap_words = want;
        }
        GCA_HIP(hipMemsetAsync(ctx->bitmap, 0, sizeof(uint32_t) * ctx->bitmap_words, ctx->stream));
        ctx->bitmap_dirty = false;
    }
    const int64_t nc = std::max<int64_t>(n_cap, 1);
    uint32_t *word_pre = (uint32_t *)ctx_scratch(ctx, 6, sizeof(uint32_t) * (words + 1), false);
    unsigned long long *acc = (unsigned long long *)ctx_scratch(ctx, 7, sizeof(unsigned long long) * kVoxRec * nc, true);
    VoxelParams *vp = (VoxelParams *)ctx_scratch(ctx, 22, sizeof(VoxelParams), false);
    if (!word_pre || !acc || !vp) return GCA_ERR_OUT_OF_MEMORY;
    GCA_TRY(scan_job(ctx, ctx->bitmap, word_pre, words, (uint32_t *)&out->hdr->n, wide, &vj->scan));
    // grid widths in multiples of 8: workgroups go to the 8 XCDs by linear id mod 8, so in an
    // all-levels launch (level = blockIdx.y) the levels' workgroups on the same points then share
    // an XCD and its L2, and the points are fetched from HBM once rather than once per level
    vj->nb = (unsigned)((std::min<int64_t>(4096, (nc + 255) / 256) + 7) / 8 * 8);
    vj->nb_acc = (unsigned)((std::min<int64_t>(4096, (nc + kVoxChunk - 1) / kVoxChunk) + 7) / 8 * 8);
    vj->mark = MarkArgs{in->xyzi, in->hdr, voxel, max_bits, vp, err, ctx->bitmap};
    vj->accum = AccumArgs{in->xyzi, in->rgb, in->hdr, vp, ctx->bitmap, word_pre, acc};
    vj->fin = FinalArgs{acc, in->hdr, out->hdr, ctx->bitmap, out->xyzi, out->rgb, gc ? gc->d : GridDesc{},
                        gc ? gc->count : nullptr, gc ? gc->cell_of : nullptr};
    return GCA_OK;
}

// mark (with the device lattice), popcount scan, accumulate, finalize: each kernel once for all
// nj calls (blockIdx.y = call), on stream.
template <int N>
static int32_t voxel_launch(hipStream_t stream, const VoxelJobs *vj, int nj, bool wide)
{
    Jobs<MarkArgs, N> mark{};
    Jobs<AccumArgs, N> accum{};
    Jobs<FinalArgs, N> fin{};
    ScanJob scans[kMaxJobs];
    unsigned nb = 1, nb_acc = 1;
    for (int j = 0; j < nj; ++j) {
        mark.a[j] = vj[j].mark;
        accum.a[j] = vj[j].accum;
        fin.a[j] = vj[j].fin;
        scans[j] = vj[j].scan;
        nb = std::max(nb, vj[j].nb);
        nb_acc = std::max(nb_acc, vj[j].nb_acc);
    }
    hipLaunchKernelGGL(k_voxel_mark<N>, dim3(nb, (unsigned)nj), dim3(256), 0, stream, mark);
    GCA_HIP(hipGetLastError());
    GCA_TRY(scan_launch(stream, scans, nj, true, wide));
    hipLaunchKernelGGL(k_voxel_accum<N>, dim3(nb_acc, (unsigned)nj), dim3(kVoxBlock), 0, stream, accum);
    hipLaunchKernelGGL(k_voxel_finalize<N>, dim3(nb, (unsigned)nj), dim3(256), 0, stream, fin);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// Kernel sequence of one voxel call: mark (with the device lattice), popcount scan over `words` bitmap
// words (>= the lattice's), accumulate, finalize.  Scratch is sized for `n_cap` input points.
static int32_t voxel_enqueue(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out, int64_t n_cap, int64_t words,
                             double max_bits, uint32_t *err, const GridCount *gc = nullptr)
{
    const bool wide = scan_wide(words);
    VoxelJobs vj;
    GCA_TRY(voxel_prepare(ctx, in, voxel, out, n_cap, words, max_bits, err, gc, wide, &vj));
    ctx->bitmap_dirty = true;  // until the whole sequence is enqueued
    {
        StageTimer timer(ctx, GCA_STAGE_VOXEL);
        GCA_TRY(voxel_launch<1>(ctx->stream, &vj, 1, wide));
    }
    ctx->bitmap_dirty = false;
    cloud_points_changed(out);
    return GCA_OK;
}

int32_t launch_voxel(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out)
{
    GCA_TRY(cloud_sync_header(in));
    const int64_t n = in->host_hdr.n;
    float mn[3], mx[3];
    cloud_host_bounds(in, mn, mx);
    VoxelParams vp{};
    vp.voxel = voxel;
    double nbits = 1.0;
    for (int a = 0; a < 3; ++a) {  // the same FP32 formulas as voxel_params_of
        const float ext = mx[a] - mn[a];
        if (n > 0 && (double)voxel * (double)INT_MAX < (double)ext)  // point_cloud.cu:213-218
            return GCA_ERR_INVALID_VALUE;
        vp.bmin[a] = mn[a];
        vp.dims[a] = n > 0 ? (int64_t)floorf(ext / voxel) + 1 : 1;
        nbits *= (double)vp.dims[a];
    }
    GCA_TRY(cloud_reserve(out, std::max<int64_t>(n, 1)));
    if (ctx->tune[GCA_TUNE_VOXEL] == 2 || nbits > kVoxelMaxBits) return voxel_sort_enqueue(ctx, in, vp, out, (int)n);
    return voxel_enqueue(ctx, in, voxel, out, n, ((int64_t)nbits + 31) / 32, kVoxelMaxBits, nullptr);
}

int32_t launch_voxel_planned(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out, double max_bits, uint32_t *err,
                             const GridCount *gc)
{
    const int64_t words = ((int64_t)max_bits + 31) / 32;
    GCA_TRY(cloud_reserve(out, std::max<int64_t>(in->cap, 1)));
    return voxel_enqueue(ctx, in, voxel, out, in->cap, words, max_bits, err, gc);
}

int32_t launch_voxel_levels(gca_ctx launch, const VoxelLevelJob *jobs, int nj)
{
    if (nj < 1 || nj > kMaxJobs) return GCA_ERR_INVALID_VALUE;
    int64_t words[kMaxJobs];
    bool wide = false;
    for (int j = 0; j < nj; ++j) {
        words[j] = ((int64_t)jobs[j].max_bits + 31) / 32;
        wide = wide || scan_wide(words[j]);
    }
    VoxelJobs vj[kMaxJobs];
    for (int j = 0; j < nj; ++j) {
        const VoxelLevelJob &J = jobs[j];
        GCA_TRY(cloud_reserve(J.out, std::max<int64_t>(J.in->cap, 1)));
        GCA_TRY(voxel_prepare(J.ctx, J.in, J.voxel, J.out, J.in->cap, words[j], J.max_bits, J.err, &J.gc, wide, &vj[j]));
        J.ctx->bitmap_dirty = true;
    }
    {
        StageTimer timer(launch, GCA_STAGE_VOXEL);
        GCA_TRY(voxel_launch<kMaxJobs>(launch->stream, vj, nj, wide));
    }
    for (int j = 0; j < nj; ++j) {
        jobs[j].ctx->bitmap_dirty = false;
        cloud_points_changed(jobs[j].out);
    }
    return GCA_OK;
}

}  // namespace gca
