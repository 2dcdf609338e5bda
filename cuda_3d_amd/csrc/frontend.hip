// frontend.hip -- depth + colour frame -> point cloud (K1).
//
// Replaces __kernel_make_pointcloud_Z16_BGR8 + remove_invalid_points
// (src/geometry/cuda_point_cloud_factory.cu:189-313, src/geometry/geometry_util.cuh:124-130).
//
// Kernel A (one block = 256 consecutive pixels of one row): the depth rows y-R..y+R of the
// segment are staged in LDS with coalesced 16-bit loads, the bilateral filter runs from LDS
// with FP64 weights taken from host-built tables (bit-identical to exp() of the exact tap
// arguments), then back-projection, depth->colour re-projection and BGR sampling in the
// reference's FP32 operation order.  Results go to a per-pixel staging buffer and each block
// records its count of valid pixels.
// Kernel B: stable compaction (wave ballot + in-block prefix) into the SoA cloud at the block's
// offset, which each block sums itself from the counts of the blocks before it in raster order
// (at most a few thousand L2-resident words: no separate scan launch); the last block writes
// the count.  Each block's min/max goes into its own slot.
// Kernel C: one workgroup folds the block bounds into the cloud header (no contended atomics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

constexpr int kRowSeg = 256;
constexpr int kMaxR = 15;
constexpr int kStageBatch = 8;  // staging loads in flight per thread
constexpr int kWrLds = 4096;    // longest range-weight table staged in LDS (32 KB)
static_assert(sizeof(double) * ((2 * kMaxR + 1) * (2 * kMaxR + 1) + kWrLds) +
                      sizeof(uint16_t) * (2 * kMaxR + 1) * (kRowSeg + 2 * kMaxR) <= 60 * 1024,
              "the pixel kernel's dynamic LDS must stay within a workgroup's 64 KB");

struct FrontParams {
    int32_t W, H, nbx;
    float scale, tmin, tmax;
    gca_intrinsics di, ci;
    float R[9];
    float t[3];
    int32_t mode;
    int32_t radius;           // fixed mode
    int32_t ws_off[6];        // per adaptive step (fixed mode uses [0])
    int32_t wr_off[6];
    int32_t wr_len[6];
};

// Back-projection of pixel (x, y) at depth z: x = ((u - 0.5 - cx) z) / fx in the reference's
// FP32 operation order (cuda_point_cloud_factory.cu:189-200).
__device__ __forceinline__ void backproject(int x, int y, float z, const gca_intrinsics &di, float &X, float &Y)
{
    const float u = (float)x - 0.5f, v = (float)y - 0.5f;
    X = __fdiv_rn(__fmul_rn(__fsub_rn(u, di.cx), z), di.fx);
    Y = __fdiv_rn(__fmul_rn(__fsub_rn(v, di.cy), z), di.fy);
}

// A point's xyzi and rgb records from its position and its colour pixel's B, G, R bytes
// (bits 0-7, 8-15, 16-23 of bgr3): colour = byte / 255 (cuda_point_cloud_factory.cu:256-261).
__device__ __forceinline__ void point_records(float X, float Y, float Z, uint32_t bgr3, float4 &oxyz, float4 &orgb)
{
    const float inv255 = 1.0f / 255.0f;
    const float b = __fmul_rn((float)(bgr3 & 0xffu), inv255);
    const float g = __fmul_rn((float)((bgr3 >> 8) & 0xffu), inv255);
    const float rr = __fmul_rn((float)((bgr3 >> 16) & 0xffu), inv255);
    oxyz = make_float4(X, Y, Z, intensity_of(rr, g, b));
    orgb = make_float4(rr, g, b, 0.0f);
}

// The pixel work of a workgroup (the row segment x0 .. x0 + 255 of row y): stage the depth rows
// y-R .. y+R and the weight tables in LDS, then for this thread's pixel the bilateral filter,
// back-projection, depth -> colour re-projection and BGR sampling in the reference's FP32
// operation order.  Returns whether the pixel is a valid point (then oxyz / orgb are its records
// and bgr3 its colour pixel's bytes); contains a barrier (every thread of the workgroup calls it).
__device__ __forceinline__ bool front_pixel(const uint16_t *__restrict__ depth, const uint8_t *__restrict__ bgr,
                                            const double *__restrict__ lut, const FrontParams &fp, int y, int x0,
                                            int n_ws, int n_wr, float4 &oxyz, float4 &orgb, uint32_t &bgr3)
{
    // dynamic LDS, sized by the host for this filter: the spatial and range weight tables (fixed
    // mode: every tap reads one of each) followed by the depth tile of the filter's own radius
    extern __shared__ __align__(16) unsigned char front_lds[];
    double *s_ws = reinterpret_cast<double *>(front_lds);
    double *s_wr = s_ws + n_ws;
    uint16_t *tile = reinterpret_cast<uint16_t *>(s_wr + n_wr);

    const int W = fp.W, H = fp.H;
    const int x = x0 + threadIdx.x;
    int R = 0;
    if (fp.mode == GCA_BILATERAL_FIXED) R = fp.radius;
    else if (fp.mode == GCA_BILATERAL_REFERENCE_ADAPTIVE) R = 5;

    // stage rows y-R .. y+R, columns x0-R .. x0+255+R (out of frame -> 0, which the filter
    // skips exactly like the reference skips out-of-frame and zero taps)
    // (each thread's loads go out kStageBatch at a time before any is written to LDS: one
    // memory round trip per batch, not per element)
    const int rows = 2 * R + 1, cols = kRowSeg + 2 * R;
    for (int b0 = 0; b0 < rows * cols; b0 += kRowSeg * kStageBatch) {
        uint16_t v[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {  // unconditional loads from clamped (in-frame) addresses
            const int k = b0 + u * kRowSeg + (int)threadIdx.x;
            const int ry = k / cols, rx = k - ry * cols;
            const int gy = y - R + ry, gx = x0 - R + rx;
            const bool in = k < rows * cols && gy >= 0 && gy < H && gx >= 0 && gx < W;
            const int cy = min(max(gy, 0), H - 1), cx = min(max(gx, 0), W - 1);
            const uint16_t d = depth[(int64_t)cy * W + cx];
            v[u] = in ? d : (uint16_t)0;
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int k = b0 + u * kRowSeg + (int)threadIdx.x;
            if (k < rows * cols) tile[k] = v[u];
        }
    }
    for (int b0 = 0; b0 < n_ws + n_wr; b0 += kRowSeg * kStageBatch) {
        double w[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {  // unconditional loads from clamped table indices
            const int k = min(b0 + u * kRowSeg + (int)threadIdx.x, n_ws + n_wr - 1);
            w[u] = lut[k < n_ws ? fp.ws_off[0] + k : fp.wr_off[0] + (k - n_ws)];
        }
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) {
            const int k = b0 + u * kRowSeg + (int)threadIdx.x;
            if (k < n_ws + n_wr) s_ws[k] = w[u];  // s_wr follows s_ws
        }
    }
    __syncthreads();

    bool valid = false;
    oxyz = make_float4(0.f, 0.f, 0.f, 0.f);
    orgb = make_float4(0.f, 0.f, 0.f, 0.f);
    bgr3 = 0u;
    if (x < W) {
        const uint16_t d0 = tile[R * cols + threadIdx.x + R];
        const float z0 = (float)d0 * fp.scale;
        if (d0 != 0 && !(z0 < fp.tmin || z0 > fp.tmax)) {
            double f = (double)d0;
            int r = -1, s = 0;
            if (fp.mode == GCA_BILATERAL_FIXED) {
                r = fp.radius;
            } else if (fp.mode == GCA_BILATERAL_REFERENCE_ADAPTIVE) {
                // cuda_point_cloud_factory.cu:119-125
                unsigned step = (unsigned)((z0 - fp.tmin) / 0.5f);
                if (step > 5u) step = 5u;
                s = (int)step;
                r = (int)(1.0f * (float)step);
            }
            if (r >= 0) {
                const int wl = fp.wr_len[s];
                const int side = 2 * r + 1;
                double sw = 0.0, swd = 0.0;
                auto taps = [&](const double *ws, const double *wr) {
                    for (int dy = -r; dy <= r; ++dy) {
                        const uint16_t *row = &tile[(R + dy) * cols + threadIdx.x + R];
                        for (int dx = -r; dx <= r; ++dx) {
                            uint16_t dn = row[dx];
                            if (!dn) continue;
                            int diff = abs((int)d0 - (int)dn);
                            double wrv = diff < wl ? wr[diff] : 0.0;
                            double w = __dmul_rn(ws[(dy + r) * side + (dx + r)], wrv);
                            sw = __dadd_rn(sw, w);
                            swd = __dadd_rn(swd, __dmul_rn(w, (double)dn));
                        }
                    }
                };
                // two copies of the loop, so each reads its tables with the right instructions
                if (n_ws > 0)
                    taps(s_ws, s_wr);
                else
                    taps(lut + fp.ws_off[s], lut + fp.wr_off[s]);
                f = swd / sw;
            }
            const float z = (float)f * fp.scale;
            if (!(z < fp.tmin || z > fp.tmax)) {
                float X, Y;
                backproject(x, y, z, fp.di, X, Y);
                const float Z = z;
                const float c0 = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(fp.R[0], X), __fmul_rn(fp.R[3], Y)),
                                                     __fmul_rn(fp.R[6], Z)), fp.t[0]);
                const float c1 = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(fp.R[1], X), __fmul_rn(fp.R[4], Y)),
                                                     __fmul_rn(fp.R[7], Z)), fp.t[1]);
                const float c2 = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(fp.R[2], X), __fmul_rn(fp.R[5], Y)),
                                                     __fmul_rn(fp.R[8], Z)), fp.t[2]);
                if (c2 > 0.0f) {
                    float fu = __fadd_rn(__fdiv_rn(__fmul_rn(c0, fp.ci.fx), c2), fp.ci.cx);
                    float fv = __fadd_rn(__fdiv_rn(__fmul_rn(c1, fp.ci.fy), c2), fp.ci.cy);
                    fu = __fadd_rn(fu, 0.5f);
                    fv = __fadd_rn(fv, 0.5f);
                    if (fu > -1.0f && fu < (float)W && fv > -1.0f && fv < (float)H) {
                        const int tx = (int)fu, ty = (int)fv;
                        const uint8_t *c = bgr + 3 * ((int64_t)ty * W + tx);
                        bgr3 = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16);
                        point_records(X, Y, Z, bgr3, oxyz, orgb);
                        valid = true;
                    }
                }
            }
        }
    }
    return valid;
}

__global__ __launch_bounds__(kRowSeg) void k_rgbd_pixels(const uint16_t *__restrict__ depth,
                                                         const uint8_t *__restrict__ bgr,
                                                         const double *__restrict__ lut, FrontParams fp,
                                                         uint2 *__restrict__ st_pix,
                                                         uint32_t *__restrict__ block_cnt,
                                                         CloudHdr *__restrict__ hdr, uint32_t *clear_word,
                                                         int n_ws, int n_wr)
{
    __shared__ uint32_t wcount[kRowSeg / kWave];
    const int W = fp.W;
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * kRowSeg;
    const int x = x0 + threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (clear_word) *clear_word = 0u;
        for (int a = 0; a < 3; ++a) {
            hdr->bmin_key[a] = float_to_key(FLT_MAX);
            hdr->bmax_key[a] = float_to_key(FLT_MIN);  // geometry_util.cuh:92
        }
    }
    float4 oxyz, orgb;
    uint32_t bgr3;
    const bool valid = front_pixel(depth, bgr, lut, fp, y, x0, n_ws, n_wr, oxyz, orgb, bgr3);
    if (x < W)  // 8 bytes per pixel: the depth in metres, the colour bytes and the valid bit
        st_pix[(int64_t)y * W + x] = make_uint2(__float_as_uint(oxyz.z), valid ? (bgr3 | (1u << 24)) : 0u);
    uint64_t m = __ballot(valid);
    if (lane_id() == 0) wcount[threadIdx.x / kWave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int i = 0; i < kRowSeg / kWave; ++i) s += wcount[i];
        block_cnt[(int64_t)y * fp.nbx + blockIdx.x] = s;
    }
}

// Block bounds of the valid points of a workgroup into its slot (FLT_MAX / -FLT_MAX if it has
// none): one small kernel folds the slots, instead of every block hitting the same six words
// with atomics.
__device__ __forceinline__ void block_bounds(bool valid, float4 a, float *bp)
{
    __shared__ float smin[3][kRowSeg / kWave], smax[3][kRowSeg / kWave];
    const int w = threadIdx.x / kWave;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (valid) {
        mn[0] = mx[0] = a.x;
        mn[1] = mx[1] = a.y;
        mn[2] = mx[2] = a.z;
    }
    for (int k = 0; k < 3; ++k) {
        float v0 = wave_min(mn[k]), v1 = wave_max(mx[k]);
        if (lane_id() == 0) smin[k][w] = v0, smax[k][w] = v1;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        float v0 = smin[k][0], v1 = smax[k][0];
        for (int q = 1; q < kRowSeg / kWave; ++q) v0 = fminf(v0, smin[k][q]), v1 = fmaxf(v1, smax[k][q]);
        bp[k] = v0;
        bp[3 + k] = v1;
    }
}

// One-pass front end: the pixel work, the workgroup's offset in the compacted cloud by decoupled
// look-back (workgroup-wide, lookback_exclusive_block: every chunk is resident at once), the
// stable compaction and the block bounds, in one kernel.
// Workgroups take row segments in raster order from a counter (so a look-back only waits on
// segments already being processed); the last one writes the point count and re-arms the counter.
// No per-pixel staging: each valid pixel is written once, to its place in the cloud.
__global__ __launch_bounds__(kRowSeg) void k_rgbd_onepass(const uint16_t *__restrict__ depth,
                                                          const uint8_t *__restrict__ bgr,
                                                          const double *__restrict__ lut, FrontParams fp,
                                                          unsigned long long *flags, ScanCtl *ctl, int nchunks,
                                                          float4 *__restrict__ out_xyzi, float4 *__restrict__ out_rgb,
                                                          float *__restrict__ bpart, CloudHdr *__restrict__ hdr,
                                                          uint32_t *clear_word, int n_ws, int n_wr, uint32_t *dev_err)
{
    __shared__ uint32_t lds[kRowSeg / kWave + 1];
    __shared__ unsigned long long lb[4 * (kRowSeg / kWave)];
    __shared__ uint32_t s_chunk, s_seq;
    if (threadIdx.x == 0) {
        const unsigned long long w = __hip_atomic_fetch_add(&ctl->ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_chunk = (uint32_t)w;
        s_seq = (uint32_t)(w >> 32);
    }
    __syncthreads();
    const int c = (int)s_chunk;
    const uint32_t seq = s_seq;
    if (c >= nchunks) {  // only after an interrupted call left the counter behind
        if (threadIdx.x == 0) __hip_atomic_store(dev_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    const int y = c / fp.nbx, x0 = (c - y * fp.nbx) * kRowSeg;
    if (c == 0 && threadIdx.x == 0 && clear_word) *clear_word = 0u;
    float4 oxyz, orgb;
    uint32_t bgr3;
    const bool valid = front_pixel(depth, bgr, lut, fp, y, x0, n_ws, n_wr, oxyz, orgb, bgr3);
    uint32_t tot;
    const uint32_t rank = block_excl_scan<kRowSeg>(valid ? 1u : 0u, lds, &tot);
    const uint32_t excl = lookback_exclusive_block<kRowSeg>(flags, c, seq, tot, dev_err, lb);
    if (threadIdx.x == 0 && c == nchunks - 1) {
        hdr->n = (int32_t)(excl + tot);
        __hip_atomic_store(&ctl->ctr, (unsigned long long)(seq + 1u) << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (valid) {
        const uint32_t dst = excl + rank;
        out_xyzi[dst] = oxyz;
        out_rgb[dst] = orgb;
    }
    block_bounds(valid, oxyz, bpart + (int64_t)c * 6);
}

// Stable compaction of the valid pixels; each point's records are rebuilt from its staged depth
// and colour bytes with the pixel kernel's own operations (the same bits as records staged whole,
// at 8 bytes of staging per pixel instead of 33).
__global__ __launch_bounds__(kRowSeg) void k_rgbd_compact(const uint2 *__restrict__ st_pix,
                                                          const uint32_t *__restrict__ block_cnt,
                                                          gca_intrinsics di, int32_t W, int32_t nbx,
                                                          float4 *__restrict__ out_xyzi,
                                                          float4 *__restrict__ out_rgb,
                                                          float *__restrict__ bpart, CloudHdr *__restrict__ hdr)
{
    __shared__ uint32_t woff[kRowSeg / kWave];
    __shared__ uint32_t wpre[kRowSeg / kWave];
    __shared__ float smin[3][kRowSeg / kWave], smax[3][kRowSeg / kWave];
    const int y = blockIdx.y;
    const int x = blockIdx.x * kRowSeg + threadIdx.x;
    const int64_t p = (int64_t)y * W + x;
    const int64_t b = (int64_t)y * nbx + blockIdx.x;  // raster order of the block
    // this block's offset: the counts of the blocks before it (strided over the threads)
    uint32_t pre = 0;
    for (int64_t k = threadIdx.x; k < b; k += kRowSeg) pre += block_cnt[k];
    pre = wave_sum(pre);
    const uint2 pix = x < W ? st_pix[p] : make_uint2(0u, 0u);
    const bool valid = (pix.y >> 24) != 0u;
    uint64_t m = __ballot(valid);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        woff[w] = (uint32_t)__popcll(m);
        wpre[w] = pre;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int i = 0; i < kRowSeg / kWave; ++i) s += wpre[i];
        for (int i = 0; i < kRowSeg / kWave; ++i) {
            uint32_t t = woff[i];
            woff[i] = s;
            s += t;
        }
        if (b == (int64_t)gridDim.x * gridDim.y - 1) hdr->n = (int32_t)s;  // the last block: the count
    }
    __syncthreads();
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (valid) {
        uint64_t below = m & ((1ull << lane_id()) - 1ull);
        uint32_t dst = woff[w] + (uint32_t)__popcll(below);
        const float z = __uint_as_float(pix.x);
        float X, Y;
        backproject(x, y, z, di, X, Y);
        float4 a, c;
        point_records(X, Y, z, pix.y, a, c);
        out_xyzi[dst] = a;
        out_rgb[dst] = c;
        mn[0] = mx[0] = a.x;
        mn[1] = mx[1] = a.y;
        mn[2] = mx[2] = a.z;
    }
    for (int a = 0; a < 3; ++a) {
        float v0 = wave_min(mn[a]), v1 = wave_max(mx[a]);
        if (lane_id() == 0) smin[a][w] = v0, smax[a][w] = v1;
    }
    __syncthreads();
    // the block's bounds go to its own slot (FLT_MAX / -FLT_MAX if it has no valid pixel); one
    // small kernel folds them, instead of every block hitting the same six words with atomics
    if (threadIdx.x < 3) {
        int a = threadIdx.x;
        float v0 = smin[a][0], v1 = smax[a][0];
        for (int k = 1; k < kRowSeg / kWave; ++k) v0 = fminf(v0, smin[a][k]), v1 = fmaxf(v1, smax[a][k]);
        float *bp = bpart + ((int64_t)y * nbx + blockIdx.x) * 6;
        bp[a] = v0;
        bp[3 + a] = v1;
    }
}

// Folds the per-block bounds into the header keys: min over blocks (initial FLT_MAX), max over
// blocks with the reference's initial FLT_MIN (geometry_util.cuh:92).
__global__ __launch_bounds__(256) void k_bounds_fold(const float *__restrict__ bpart, int64_t nblk,
                                                     CloudHdr *__restrict__ hdr)
{
    __shared__ float smin[3][4], smax[3][4];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {FLT_MIN, FLT_MIN, FLT_MIN};
    for (int64_t b = threadIdx.x; b < nblk; b += 256) {
        const float *bp = bpart + b * 6;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], bp[a]);
            mx[a] = fmaxf(mx[a], bp[3 + a]);
        }
    }
    const int w = threadIdx.x / kWave;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float v0 = wave_min(mn[a]), v1 = wave_max(mx[a]);
        if (lane_id() == 0) smin[a][w] = v0, smax[a][w] = v1;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        float v0 = smin[a][0], v1 = smax[a][0];
        for (int k = 1; k < 4; ++k) v0 = fminf(v0, smin[a][k]), v1 = fmaxf(v1, smax[a][k]);
        hdr->bmin_key[a] = float_to_key(v0);
        hdr->bmax_key[a] = float_to_key(v1);
    }
}

// Host: FP64 weight tables, exactly exp() of the tap arguments the oracle evaluates.
static int32_t build_lut(gca_ctx ctx, const gca_bilateral *f, FrontParams *fp)
{
    bool same = ctx->lut && ctx->lut_key_mode == f->mode;
    if (same && f->mode == GCA_BILATERAL_FIXED)
        same = ctx->lut_key_radius == f->radius && ctx->lut_key_ss == f->sigma_space &&
               ctx->lut_key_sd == f->sigma_depth;
    std::vector<double> tab;
    int nsteps = (f->mode == GCA_BILATERAL_REFERENCE_ADAPTIVE) ? 6 : 1;
    for (int s = 0; s < nsteps; ++s) {
        int R;
        double ss, sd;
        if (f->mode == GCA_BILATERAL_REFERENCE_ADAPTIVE) {
            R = (int)(1.0f * (float)s);
            ss = (double)(8.0f + (float)s * 1.0f);
            sd = (double)(100.0f + (float)s * 10.0f);
        } else {
            R = f->radius;
            ss = (double)f->sigma_space;
            sd = (double)f->sigma_depth;
        }
        fp->ws_off[s] = (int32_t)tab.size();
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) tab.push_back(exp(-((double)(dx * dx + dy * dy)) / (ss * ss)));
        fp->wr_off[s] = (int32_t)tab.size();
        int len = 0;
        for (int k = 0; k < 65536; ++k) {
            double w = exp(-((double)k * (double)k) / (sd * sd));
            tab.push_back(w);
            ++len;
            if (w == 0.0) break;
        }
        fp->wr_len[s] = len;
    }
    if (!same) {
        size_t bytes = tab.size() * sizeof(double);
        if (ctx->lut_bytes < bytes) {
            GCA_HIP(hipStreamSynchronize(ctx->stream));
            hipFree(ctx->lut);
            ctx->lut = nullptr;
            ctx->lut_bytes = 0;
            GCA_HIP(hipMalloc(&ctx->lut, bytes));
            ctx->lut_bytes = bytes;
        }
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        GCA_HIP(hipMemcpy(ctx->lut, tab.data(), bytes, hipMemcpyHostToDevice));
        ctx->lut_key_mode = f->mode;
        ctx->lut_key_radius = f->radius;
        ctx->lut_key_ss = f->sigma_space;
        ctx->lut_key_sd = f->sigma_depth;
    }
    return GCA_OK;
}

struct FrontLaunch {
    bool onepass = false;  // k_rgbd_onepass, else k_rgbd_pixels
    dim3 grid, block;
    unsigned lds = 0;
    // the kernel's arguments (those of k_rgbd_pixels, then those only k_rgbd_onepass takes)
    const uint16_t *depth = nullptr;
    const uint8_t *bgr = nullptr;
    const double *lut = nullptr;
    FrontParams fp{};
    uint2 *st_pix = nullptr;
    uint32_t *bcnt = nullptr;
    CloudHdr *hdr = nullptr;
    uint32_t *clear_word = nullptr;
    int n_ws = 0, n_wr = 0;
    unsigned long long *flags = nullptr;
    ScanCtl *ctl = nullptr;
    int nchunks = 0;
    float4 *out_xyzi = nullptr, *out_rgb = nullptr;
    float *bpart = nullptr;
    uint32_t *dev_err = nullptr;
    bool valid = false;
    void *func() const
    {
        return onepass ? reinterpret_cast<void *>(&k_rgbd_onepass) : reinterpret_cast<void *>(&k_rgbd_pixels);
    }
};

FrontLaunch *front_launch_new() { return new (std::nothrow) FrontLaunch(); }

void front_launch_free(FrontLaunch *f) { delete f; }

hipGraphNode_t front_launch_node(const FrontLaunch *f, hipGraph_t g)
{
    if (!f || !f->valid || !g) return nullptr;
    size_t n = 0;
    if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess || n == 0) return nullptr;
    std::vector<hipGraphNode_t> nodes(n);
    if (hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return nullptr;
    hipGraphNode_t found = nullptr;
    for (size_t k = 0; k < n; ++k) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[k], &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp{};
        if (hipGraphKernelNodeGetParams(nodes[k], &kp) != hipSuccess) continue;
        if (kp.func == f->func()) {
            if (found) return nullptr;  // more than one: not the graph of one pipeline step
            found = nodes[k];
        }
    }
    return found;
}

hipError_t front_launch_update(FrontLaunch *f, hipGraphExec_t exec, hipGraphNode_t node, const uint16_t *depth,
                               const uint8_t *bgr)
{
    if (!f || !f->valid || !exec || !node) return hipErrorInvalidValue;
    f->depth = depth;
    f->bgr = bgr;
    void *pix[] = {&f->depth, &f->bgr, &f->lut, &f->fp, &f->st_pix, &f->bcnt, &f->hdr, &f->clear_word, &f->n_ws,
                   &f->n_wr};
    void *one[] = {&f->depth, &f->bgr, &f->lut, &f->fp, &f->flags, &f->ctl, &f->nchunks, &f->out_xyzi,
                   &f->out_rgb, &f->bpart, &f->hdr, &f->clear_word, &f->n_ws, &f->n_wr, &f->dev_err};
    hipKernelNodeParams kp{};
    kp.func = f->func();
    kp.gridDim = f->grid;
    kp.blockDim = f->block;
    kp.sharedMemBytes = f->lds;
    kp.kernelParams = f->onepass ? one : pix;
    kp.extra = nullptr;
    return hipGraphExecKernelNodeSetParams(exec, node, &kp);
}

int32_t launch_from_rgbd(gca_ctx ctx, const uint16_t *depth, const uint8_t *bgr, const gca_camera *cam,
                         const gca_bilateral *bil, float tmin, float tmax, gca_cloud out, uint32_t *clear_word)
{
    const int W = cam->width, H = cam->height;
    const int64_t P = (int64_t)W * H;
    FrontParams fp{};
    fp.W = W;
    fp.H = H;
    fp.nbx = (W + kRowSeg - 1) / kRowSeg;
    fp.scale = cam->depth_scale;
    fp.tmin = tmin;
    fp.tmax = tmax;
    fp.di = cam->depth_intrin;
    fp.ci = cam->color_intrin;
    for (int k = 0; k < 9; ++k) fp.R[k] = cam->depth_to_color.rotation[k];
    for (int k = 0; k < 3; ++k) fp.t[k] = cam->depth_to_color.translation[k];
    fp.mode = bil->mode;
    fp.radius = bil->radius;
    if (bil->mode != GCA_BILATERAL_NONE) GCA_TRY(build_lut(ctx, bil, &fp));
    GCA_TRY(cloud_reserve(out, P));
    const int64_t nblocks = (int64_t)fp.nbx * H;
    const bool onepass = ctx->tune[GCA_TUNE_FRONTEND] == 1;
    float *bpart = (float *)ctx_scratch(ctx, 20, sizeof(float) * 6 * nblocks, false);
    if (!bpart) return GCA_ERR_OUT_OF_MEMORY;
    // LDS of the pixel work: weight tables (fixed mode; the range table while it is short)
    // and the depth tile of the filter's radius
    const int R = bil->mode == GCA_BILATERAL_FIXED ? bil->radius : (bil->mode == GCA_BILATERAL_REFERENCE_ADAPTIVE ? 5 : 0);
    const bool stage = bil->mode == GCA_BILATERAL_FIXED && fp.wr_len[0] <= kWrLds;  // both tables or neither
    const int n_ws = stage ? (2 * R + 1) * (2 * R + 1) : 0;
    const int n_wr = stage ? fp.wr_len[0] : 0;
    const size_t lds = sizeof(double) * (size_t)(n_ws + n_wr) + sizeof(uint16_t) * (size_t)(2 * R + 1) * (kRowSeg + 2 * R);
    StageTimer timer(ctx, GCA_STAGE_FRONTEND);
    if (onepass) {
        // look-back words and the segment counter: zeroed once when allocated, then carried
        // from call to call by their sequence numbers
        unsigned long long *flags =
            (unsigned long long *)ctx_scratch(ctx, 36, sizeof(unsigned long long) * (size_t)nblocks, true);
        ScanCtl *ctl = (ScanCtl *)ctx_scratch(ctx, 37, sizeof(ScanCtl), true);
        if (!flags || !ctl) return GCA_ERR_OUT_OF_MEMORY;
        if (nblocks > INT_MAX) return GCA_ERR_INVALID_VALUE;
        const dim3 grid((unsigned)nblocks);
        if (FrontLaunch *r = ctx->front_rec) {
            *r = FrontLaunch{};
            r->onepass = true;
            r->grid = grid;
            r->block = dim3(kRowSeg);
            r->lds = (unsigned)lds;
            r->depth = depth;
            r->bgr = bgr;
            r->lut = (const double *)ctx->lut;
            r->fp = fp;
            r->hdr = out->hdr;
            r->clear_word = clear_word;
            r->n_ws = n_ws;
            r->n_wr = n_wr;
            r->flags = flags;
            r->ctl = ctl;
            r->nchunks = (int)nblocks;
            r->out_xyzi = out->xyzi;
            r->out_rgb = out->rgb;
            r->bpart = bpart;
            r->dev_err = ctx->dev_err;
            r->valid = true;
        }
        hipLaunchKernelGGL(k_rgbd_onepass, grid, dim3(kRowSeg), lds, ctx->stream, depth, bgr, (const double *)ctx->lut,
                           fp, flags, ctl, (int)nblocks, out->xyzi, out->rgb, bpart, out->hdr, clear_word, n_ws, n_wr,
                           ctx->dev_err);
        GCA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_bounds_fold, dim3(1), dim3(256), 0, ctx->stream, (const float *)bpart, nblocks, out->hdr);
    } else {
        uint2 *st_pix = (uint2 *)ctx_scratch(ctx, 0, sizeof(uint2) * P, false);
        uint32_t *bcnt = (uint32_t *)ctx_scratch(ctx, 3, sizeof(uint32_t) * (nblocks + 1), false);
        if (!st_pix || !bcnt) return GCA_ERR_OUT_OF_MEMORY;
        dim3 grid(fp.nbx, H);
        if (FrontLaunch *r = ctx->front_rec) {
            *r = FrontLaunch{};
            r->grid = grid;
            r->block = dim3(kRowSeg);
            r->lds = (unsigned)lds;
            r->depth = depth;
            r->bgr = bgr;
            r->lut = (const double *)ctx->lut;
            r->fp = fp;
            r->st_pix = st_pix;
            r->bcnt = bcnt;
            r->hdr = out->hdr;
            r->clear_word = clear_word;
            r->n_ws = n_ws;
            r->n_wr = n_wr;
            r->valid = true;
        }
        hipLaunchKernelGGL(k_rgbd_pixels, grid, dim3(kRowSeg), lds, ctx->stream, depth, bgr,
                           (const double *)ctx->lut, fp, st_pix, bcnt, out->hdr, clear_word, n_ws, n_wr);
        GCA_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_rgbd_compact, grid, dim3(kRowSeg), 0, ctx->stream, (const uint2 *)st_pix,
                           (const uint32_t *)bcnt, fp.di, W, fp.nbx, out->xyzi, out->rgb, bpart, out->hdr);
        hipLaunchKernelGGL(k_bounds_fold, dim3(1), dim3(256), 0, ctx->stream, (const float *)bpart, nblocks, out->hdr);
    }
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

}  // namespace gca
