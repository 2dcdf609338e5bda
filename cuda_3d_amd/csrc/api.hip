// api.hip -- C ABI of libgca.so: contexts, clouds, argument validation and the small utility
// kernels (bounds, transform, upload/copy, scans).  The heavy kernels live in frontend.hip,
// voxel.hip, grid.hip, normals.hip and icp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <new>

#include "gca_device.hpp"
#include "gca_internal.hpp"

namespace gca {

int32_t hip_status(hipError_t e)
{
    if (e == hipErrorOutOfMemory) return GCA_ERR_OUT_OF_MEMORY;
    return GCA_ERR_DEVICE;
}

int32_t ctx_take_device_error(gca_ctx ctx)
{
    if (!ctx->dev_err || __atomic_load_n(ctx->dev_err, __ATOMIC_ACQUIRE) == 0u) return GCA_OK;
    __atomic_store_n(ctx->dev_err, 0u, __ATOMIC_RELEASE);
    return GCA_ERR_DEVICE;
}

void *ctx_scratch(gca_ctx ctx, int slot, size_t bytes, bool zeroed)
{
    if (slot < 0 || slot >= gca_ctx_s::kScratchSlots) return nullptr;
    if (bytes == 0) bytes = 16;
    if (ctx->scratch_bytes[slot] < bytes) {
        if (ctx->scratch[slot]) {
            hipStreamSynchronize(ctx->stream);
            hipFree(ctx->scratch[slot]);
            ctx->scratch[slot] = nullptr;
        }
        size_t want = bytes + bytes / 2;
        want = (want + 255) & ~size_t(255);
        if (hipMalloc(&ctx->scratch[slot], want) != hipSuccess) {
            ctx->scratch[slot] = nullptr;
            ctx->scratch_bytes[slot] = 0;
            return nullptr;
        }
        ctx->scratch_bytes[slot] = want;
        ctx->scratch_zeroed[slot] = false;
    }
    if (zeroed && !ctx->scratch_zeroed[slot]) {
        if (hipMemsetAsync(ctx->scratch[slot], 0, ctx->scratch_bytes[slot], ctx->stream) != hipSuccess)
            return nullptr;
        ctx->scratch_zeroed[slot] = true;
    }
    return ctx->scratch[slot];
}

// ---------------------------------------------------------------------------------------
// utility kernels

__global__ void k_bounds_init(CloudHdr *h)
{
    for (int a = 0; a < 3; ++a) {
        h->bmin_key[a] = float_to_key(FLT_MAX);
        h->bmax_key[a] = float_to_key(FLT_MIN);  // geometry_util.cuh:92 initialises max with FLT_MIN
    }
}

// Block-reduced AABB, one atomic per block and component.
__global__ __launch_bounds__(256) void k_bounds_reduce(const float4 *__restrict__ xyzi,
                                                       CloudHdr *__restrict__ h)
{
    __shared__ float smin[3][4], smax[3][4];
    const int n = h->n;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = xyzi[i];
        mn[0] = fminf(mn[0], p.x), mn[1] = fminf(mn[1], p.y), mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x), mx[1] = fmaxf(mx[1], p.y), mx[2] = fmaxf(mx[2], p.z);
    }
    const int w = threadIdx.x / kWave;
    for (int a = 0; a < 3; ++a) {
        float a0 = wave_min(mn[a]), a1 = wave_max(mx[a]);
        if (lane_id() == 0) smin[a][w] = a0, smax[a][w] = a1;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int a = threadIdx.x;
        float m0 = smin[a][0], m1 = smax[a][0];
        for (int k = 1; k < 4; ++k) m0 = fminf(m0, smin[a][k]), m1 = fmaxf(m1, smax[a][k]);
        if (m0 <= m1 || n > 0) {
            atomicMin(&h->bmin_key[a], float_to_key(m0));
            atomicMax(&h->bmax_key[a], float_to_key(m1));
        }
    }
}

__global__ __launch_bounds__(256) void k_transform(float4 *__restrict__ xyzi, float4 *__restrict__ nrm,
                                                   float4 *__restrict__ grad, const CloudHdr *__restrict__ h,
                                                   double t00, double t01, double t02, double t03,
                                                   double t10, double t11, double t12, double t13,
                                                   double t20, double t21, double t22, double t23,
                                                   int do_n, int do_g)
{
    const int n = h->n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 p = xyzi[i];
        double x = p.x, y = p.y, z = p.z;
        p.x = (float)(((t00 * x + t01 * y) + t02 * z) + t03);
        p.y = (float)(((t10 * x + t11 * y) + t12 * z) + t13);
        p.z = (float)(((t20 * x + t21 * y) + t22 * z) + t23);
        xyzi[i] = p;
        if (do_n) {
            float4 v = nrm[i];
            double a = v.x, b = v.y, c = v.z;
            v.x = (float)((t00 * a + t01 * b) + t02 * c);
            v.y = (float)((t10 * a + t11 * b) + t12 * c);
            v.z = (float)((t20 * a + t21 * b) + t22 * c);
            nrm[i] = v;
        }
        if (do_g) {
            float4 v = grad[i];
            double a = v.x, b = v.y, c = v.z;
            v.x = (float)((t00 * a + t01 * b) + t02 * c);
            v.y = (float)((t10 * a + t11 * b) + t12 * c);
            v.z = (float)((t20 * a + t21 * b) + t22 * c);
            grad[i] = v;
        }
    }
}

// ---- exclusive scan of uint32: one pass with decoupled look-back ---------------------------
// (lookback_exclusive, gca_device.hpp).  The workgroup that takes the last chunk re-arms the
// counter with the next sequence, so nothing is cleared between calls; a look-back that gives up
// raises the ctx's device error (reported by gca_ctx_synchronize).
// Chunks of BLOCK x kScanPer values: 256-thread workgroups for short scans (the shortest
// per-chunk latency), 1024-thread ones beyond kScanWideFrom 256-thread chunks (a look-back covers
// 64 chunks per memory round trip, so fewer, wider chunks shorten a long scan's prefix chain).
constexpr int kScanPer = 16;
constexpr int64_t kScanWideFrom = 128;

template <bool POPC>
__device__ __forceinline__ uint32_t scan_val(uint32_t v) { return POPC ? (uint32_t)__popc(v) : v; }

template <bool POPC, int BLOCK, int N>
__global__ __launch_bounds__(BLOCK) void k_scan_onepass(Jobs<ScanJob, N> J)
{
    const ScanJob &A = J.a[N == 1 ? 0 : blockIdx.y];
    if ((int)blockIdx.x >= A.nchunks) return;  // a smaller job's surplus workgroups
    const uint32_t *__restrict__ in = A.in;
    uint32_t *__restrict__ out = A.out;
    const int64_t n = A.n;
    const int nchunks = A.nchunks;
    __shared__ uint32_t lds[BLOCK / kWave + 1];
    __shared__ uint32_t s_chunk, s_seq, s_excl;
    if (threadIdx.x == 0) {
        const unsigned long long w = __hip_atomic_fetch_add(&A.ctl->ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_chunk = (uint32_t)w;
        s_seq = (uint32_t)(w >> 32);
    }
    __syncthreads();
    const int c = (int)s_chunk;
    const uint32_t seq = s_seq;
    if (c >= nchunks) {  // only after an interrupted call left the counter behind
        if (threadIdx.x == 0) __hip_atomic_store(A.dev_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        return;
    }
    const int64_t base = (int64_t)c * (BLOCK * kScanPer) + (int64_t)threadIdx.x * kScanPer;
    const bool full = base + kScanPer <= n && ((reinterpret_cast<uintptr_t>(in) & 15u) == 0);
    uint32_t v[kScanPer];
    if (full) {
        const uint4 *p4 = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
        for (int q = 0; q < kScanPer / 4; ++q) {
            const uint4 u = p4[q];
            v[4 * q] = u.x;
            v[4 * q + 1] = u.y;
            v[4 * q + 2] = u.z;
            v[4 * q + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = scan_val<POPC>(v[k]);
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<BLOCK>(s, lds, &tot);
    if (threadIdx.x < kWave) {
        const int lane = lane_id();
        const uint32_t excl = lookback_exclusive(A.flags, c, seq, tot, A.dev_err);
        if (lane == 0) {
            s_excl = excl;
            if (c == nchunks - 1) {
                if (A.total_dev) *A.total_dev = excl + tot;
                if (A.out_end) *A.out_end = excl + tot;
                __hip_atomic_store(&A.ctl->ctr, (unsigned long long)(seq + 1u) << 32, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    run += s_excl;
    uint32_t o[kScanPer];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (base + kScanPer <= n && ((reinterpret_cast<uintptr_t>(out) & 15u) == 0)) {
        uint4 *p4 = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
        for (int q = 0; q < kScanPer / 4; ++q) p4[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
}

bool scan_wide(int64_t n) { return (n + 256 * kScanPer - 1) / (256 * kScanPer) > kScanWideFrom; }

int32_t scan_job(gca_ctx ctx, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total_dev, bool wide,
                 ScanJob *job)
{
    const int64_t chunk = (wide ? 1024 : 256) * kScanPer;
    int64_t nchunks = (n + chunk - 1) / chunk;
    if (nchunks < 1) nchunks = 1;
    if (nchunks > 1024 * 64) return GCA_ERR_INVALID_VALUE;
    unsigned long long *flags = (unsigned long long *)ctx_scratch(ctx, 15, sizeof(unsigned long long) * nchunks, true);
    ScanCtl *ctl = (ScanCtl *)ctx_scratch(ctx, 23, sizeof(ScanCtl), true);
    if (!flags || !ctl) return GCA_ERR_OUT_OF_MEMORY;
    *job = ScanJob{in, n, (int32_t)nchunks, 0, flags, ctl, out, total_dev, out + n, ctx->dev_err};
    return GCA_OK;
}

template <bool POPC, int BLOCK, int N>
static void scan_launch_n(hipStream_t stream, const ScanJob *jobs, int nj)
{
    Jobs<ScanJob, N> J{};
    int gx = 1;
    for (int j = 0; j < nj; ++j) {
        J.a[j] = jobs[j];
        gx = std::max(gx, (int)jobs[j].nchunks);
    }
    hipLaunchKernelGGL((k_scan_onepass<POPC, BLOCK, N>), dim3((unsigned)gx, (unsigned)nj), dim3(BLOCK), 0, stream, J);
}

template <bool POPC, int BLOCK>
static void scan_launch_b(hipStream_t stream, const ScanJob *jobs, int nj)
{
    if (nj == 1) scan_launch_n<POPC, BLOCK, 1>(stream, jobs, nj);
    else scan_launch_n<POPC, BLOCK, kMaxJobs>(stream, jobs, nj);
}

int32_t scan_launch(hipStream_t stream, const ScanJob *jobs, int nj, bool popcount, bool wide)
{
    if (nj < 1 || nj > kMaxJobs) return GCA_ERR_INVALID_VALUE;
    if (popcount && wide) scan_launch_b<true, 1024>(stream, jobs, nj);
    else if (popcount) scan_launch_b<true, 256>(stream, jobs, nj);
    else if (wide) scan_launch_b<false, 1024>(stream, jobs, nj);
    else scan_launch_b<false, 256>(stream, jobs, nj);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

int32_t exclusive_scan_u32(gca_ctx ctx, const uint32_t *in, uint32_t *out, int64_t n,
                           uint32_t *total_dev, bool popcount)
{
    const bool wide = scan_wide(n);
    ScanJob job;
    GCA_TRY(scan_job(ctx, in, out, n, total_dev, wide, &job));
    return scan_launch(ctx->stream, &job, 1, popcount, wide);
}

// ---------------------------------------------------------------------------------------

int32_t cloud_reserve(gca_cloud c, int64_t cap)
{
    if (cap <= c->cap) return GCA_OK;
    if (cap > (int64_t)INT_MAX / 2) return GCA_ERR_INVALID_VALUE;
    GCA_HIP(hipStreamSynchronize(c->ctx->stream));
    hipFree(c->xyzi);
    hipFree(c->rgb);
    hipFree(c->nrm);
    hipFree(c->grad);
    c->xyzi = c->rgb = c->nrm = c->grad = nullptr;
    c->cap = 0;
    size_t b = sizeof(float4) * (size_t)cap;
    GCA_HIP(hipMalloc(&c->xyzi, b));
    GCA_HIP(hipMalloc(&c->rgb, b));
    GCA_HIP(hipMalloc(&c->nrm, b));
    GCA_HIP(hipMalloc(&c->grad, b));
    c->cap = cap;
    return GCA_OK;
}

int32_t cloud_sync_header(gca_cloud c)
{
    if (c->host_valid) return GCA_OK;
    GCA_HIP(hipMemcpyAsync(&c->host_hdr, c->hdr, sizeof(CloudHdr), hipMemcpyDeviceToHost,
                           c->ctx->stream));
    GCA_HIP(hipStreamSynchronize(c->ctx->stream));
    GCA_TRY(ctx_take_device_error(c->ctx));  // a header behind a failed scan is not valid
    c->host_valid = true;
    return GCA_OK;
}

void cloud_host_bounds(gca_cloud c, float mn[3], float mx[3])
{
    for (int a = 0; a < 3; ++a) {
        mn[a] = key_to_float(c->host_hdr.bmin_key[a]);
        mx[a] = key_to_float(c->host_hdr.bmax_key[a]);
    }
}

void cloud_points_changed(gca_cloud c)
{
    c->host_valid = false;
    c->grid.valid = false;
    c->grid.attrs_valid = false;
}

int32_t launch_bounds(gca_ctx ctx, gca_cloud c)
{
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, ctx->stream, c->hdr);
    // few blocks: each issues six atomics on the same header words
    int64_t nb = std::min<int64_t>(64, std::max<int64_t>(1, (c->cap + 255) / 256));
    hipLaunchKernelGGL(k_bounds_reduce, dim3((unsigned)nb), dim3(256), 0, ctx->stream, c->xyzi, c->hdr);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

int32_t launch_transform(gca_ctx ctx, gca_cloud c, const double *T)
{
    int64_t nb = std::min<int64_t>(2048, std::max<int64_t>(1, (c->cap + 255) / 256));
    hipLaunchKernelGGL(k_transform, dim3((unsigned)nb), dim3(256), 0, ctx->stream, c->xyzi, c->nrm,
                       c->grad, c->hdr, T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9],
                       T[10], T[11], c->has_normals ? 1 : 0, c->has_gradient ? 1 : 0);
    GCA_HIP(hipGetLastError());
    return GCA_OK;
}

// Host-side canonical grid geometry; identical FP32 formulas to the oracle.
void grid_params(const float bmin[3], const float bmax[3], float r, GridDesc *d, int32_t cell_quarters)
{
    float pad = 1.5f * r;
    float top[3];
    for (int a = 0; a < 3; ++a) {
        d->origin[a] = bmin[a] - pad;
        top[a] = bmax[a] + pad;
    }
    float s = cell_quarters > 4 ? r * (0.25f * (float)cell_quarters) : r;
    for (;;) {
        double total = 1.0;
        for (int a = 0; a < 3; ++a) {
            float q = (top[a] - d->origin[a]) / s;
            d->dims[a] = (q < 2.0e9f) ? (int32_t)q + 1 : 2000000000;
            total *= (double)d->dims[a];
        }
        if (total <= (double)kGridCellCap) break;
        s = s * 2.0f;
    }
    d->cell = s;
    d->radius = r;
    d->ncells = (int64_t)d->dims[0] * d->dims[1] * d->dims[2];
    d->cell_q = cell_quarters > 4 ? cell_quarters : 0;
    d->pad = 0;
}

int32_t grid_build(gca_ctx ctx, gca_cloud c, float radius, bool exact)
{
    // the tuned cell edge (GCA_TUNE_GRID_CELL) applies to grids that only serve searches; an exact
    // build keeps the canonical cell = radius (cell binning as the reference's)
    const int32_t q = exact || ctx->tune[GCA_TUNE_GRID_CELL] <= 4 ? 0 : ctx->tune[GCA_TUNE_GRID_CELL];
    if (c->grid.valid && c->grid.desc.cell_q == q) {
        if (c->grid.desc.radius == radius) return GCA_OK;
        // a canonical grid built for a nearby radius serves searches too (its cell is >= radius)
        if (!exact && q == 0 && c->grid.desc.cell >= radius && c->grid.desc.cell <= 1.5f * radius) return GCA_OK;
    }
    GCA_TRY(cloud_sync_header(c));
    float mn[3], mx[3];
    cloud_host_bounds(c, mn, mx);
    GridDesc d;
    grid_params(mn, mx, radius, &d, q);
    GCA_TRY(launch_grid_build(ctx, c, d));
    return GCA_OK;
}

// Events are recorded on the launching ctx's stream; the record is kept by the owner's profiler
// (a frame pipeline's level ctxs report to the pipeline's ctx).
StageTimer::StageTimer(gca_ctx c, int stage) : ctx(c), owner(c->prof_owner ? c->prof_owner : c)
{
    if (!owner->prof_on || ctx->capturing) return;
    while (owner->prof_pool.size() < owner->prof_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        owner->prof_pool.push_back(e);
    }
    gca_ctx_s::ProfRec r{stage, owner->prof_pool[owner->prof_used], owner->prof_pool[owner->prof_used + 1]};
    owner->prof_used += 2;
    if (hipEventRecord(r.a, ctx->stream) != hipSuccess) return;
    idx = (int)owner->prof.size();
    owner->prof.push_back(r);
}

StageTimer::~StageTimer()
{
    if (idx >= 0) (void)hipEventRecord(owner->prof[idx].b, ctx->stream);
}

}  // namespace gca

using namespace gca;

// =========================================================================================
// C ABI

extern "C" {

const char *gca_status_string(int32_t s)
{
    switch (s) {
    case GCA_OK: return "ok";
    case GCA_ERR_INVALID_VALUE: return "invalid value";
    case GCA_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GCA_ERR_DEVICE: return "device (HIP) error";
    case GCA_ERR_NOT_SPD: return "normal equations not positive definite (identity step)";
    case GCA_ERR_NOT_READY: return "missing prerequisite (normals / gradient / correspondences)";
    default: return "unknown status";
    }
}

int32_t gca_abi_version(void) { return GCA_ABI_VERSION; }

int32_t gca_ctx_create(int32_t device, gca_ctx *out)
{
    if (!out) return GCA_ERR_INVALID_VALUE;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GCA_ERR_DEVICE;
    if (device < 0 || device >= ndev) return GCA_ERR_INVALID_VALUE;
    GCA_HIP(hipSetDevice(device));
    gca_ctx c = new (std::nothrow) gca_ctx_s();
    if (!c) return GCA_ERR_OUT_OF_MEMORY;
    c->device = device;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->n_cu < 1)
        c->n_cu = 1;
    if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) c->wall_khz = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_a) != hipSuccess || hipEventCreate(&c->ev_b) != hipSuccess ||
        hipHostMalloc((void **)&c->dev_err, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        if (c->stream) hipStreamDestroy(c->stream);
        if (c->ev_a) hipEventDestroy(c->ev_a);
        if (c->ev_b) hipEventDestroy(c->ev_b);
        delete c;
        return GCA_ERR_DEVICE;
    }
    *c->dev_err = 0u;
    c->noalloc_scratch = noalloc_setup(device);  // this device's static scratch (no allocation)
    *out = c;
    return GCA_OK;
}

int32_t gca_ctx_destroy(gca_ctx c)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (int i = 0; i < 16; ++i)
        if (c->level_clouds[i]) gca_cloud_destroy(c->level_clouds[i]);
    for (int i = 0; i < gca_ctx_s::kScratchSlots; ++i) hipFree(c->scratch[i]);
    for (hipEvent_t e : c->prof_pool) hipEventDestroy(e);
    hipFree(c->bitmap);
    hipFree(c->lut);
    hipFree(c->icp_state);
    hipFree(c->partials);
    hipEventDestroy(c->ev_a);
    hipEventDestroy(c->ev_b);
    hipStreamDestroy(c->stream);
    hipHostFree(c->dev_err);
    delete c;
    return GCA_OK;
}

int32_t gca_ctx_synchronize(gca_ctx c)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    GCA_HIP(hipStreamSynchronize(c->stream));
    for (gca_ctx o : c->sync_also)
        if (o) {
            GCA_HIP(hipStreamSynchronize(o->stream));
            GCA_TRY(ctx_take_device_error(o));
        }
    return ctx_take_device_error(c);
}

int32_t gca_ctx_get_stream(gca_ctx c, void **stream)
{
    if (!c || !stream) return GCA_ERR_INVALID_VALUE;
    *stream = (void *)c->stream;
    return GCA_OK;
}

int32_t gca_ctx_set_tuning(gca_ctx c, int32_t key, int32_t value)
{
    if (!c || key < 0 || key >= GCA_TUNE_COUNT || value < 0) return GCA_ERR_INVALID_VALUE;
    auto pow2 = [](int32_t v, int32_t hi) { return v == 0 || (v <= hi && (v & (v - 1)) == 0); };
    if ((key <= GCA_TUNE_ICP_LANES_LARGE || key == GCA_TUNE_ICP_LANES_BATCH) && !pow2(value, 64))
        return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_ICP_MAX_UNITS && value > 1024) return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_ICP_PERSIST_BLOCKS && value > 65536) return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_ICP_LAUNCH && value > 3) return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_FRONTEND && value > 2) return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_VOXEL && value > 2) return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_GRID_CELL && value != 0 && (value < 4 || value > 12)) return GCA_ERR_INVALID_VALUE;
    if (key == GCA_TUNE_NORMALS_LANES && !pow2(value, 64)) return GCA_ERR_INVALID_VALUE;
    if (c->tune[key] != value) ++c->tune_version;
    c->tune[key] = value;
    return GCA_OK;
}

int32_t gca_ctx_get_tuning(gca_ctx c, int32_t key, int32_t *value)
{
    if (!c || !value || key < 0 || key >= GCA_TUNE_COUNT) return GCA_ERR_INVALID_VALUE;
    *value = c->tune[key];
    return GCA_OK;
}

int32_t gca_ctx_profile(gca_ctx c, int32_t enable)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    GCA_HIP(hipStreamSynchronize(c->stream));
    c->prof.clear();
    c->prof_used = 0;
    c->prof_on = enable != 0;
    return GCA_OK;
}

int32_t gca_ctx_profile_read(gca_ctx c, int64_t *launches, double *total_ms)
{
    if (!c || !launches || !total_ms) return GCA_ERR_INVALID_VALUE;
    GCA_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < GCA_STAGE_COUNT; ++s) {
        launches[s] = 0;
        total_ms[s] = 0.0;
    }
    for (const auto &r : c->prof) {
        float ms = 0.f;
        GCA_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        launches[r.stage] += 1;
        total_ms[r.stage] += ms;
    }
    return GCA_OK;
}

int32_t gca_cloud_create(gca_ctx ctx, int64_t capacity, gca_cloud *out)
{
    if (!ctx || !out || capacity < 0) return GCA_ERR_INVALID_VALUE;
    *out = nullptr;
    gca_cloud c = new (std::nothrow) gca_cloud_s();
    if (!c) return GCA_ERR_OUT_OF_MEMORY;
    c->ctx = ctx;
    if (hipMalloc(&c->hdr, sizeof(CloudHdr)) != hipSuccess) {
        delete c;
        return GCA_ERR_OUT_OF_MEMORY;
    }
    CloudHdr h{};
    h.n = 0;
    for (int a = 0; a < 3; ++a) {
        h.bmin_key[a] = float_to_key(FLT_MAX);
        h.bmax_key[a] = float_to_key(FLT_MIN);
    }
    if (hipMemcpy(c->hdr, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(c->hdr);
        delete c;
        return GCA_ERR_DEVICE;
    }
    c->host_hdr = h;
    c->host_valid = true;
    int32_t s = cloud_reserve(c, std::max<int64_t>(capacity, 1));
    if (s != GCA_OK) {
        gca_cloud_destroy(c);
        return s;
    }
    *out = c;
    return GCA_OK;
}

static void grid_free(Grid &g)
{
    hipFree(g.count);
    hipFree(g.start);
    hipFree(g.order);
    hipFree(g.sxyzi);
    hipFree(g.cell_of);
    hipFree(g.snrm);
    hipFree(g.sgrad);
    g = Grid();
}

int32_t gca_cloud_destroy(gca_cloud c)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    hipStreamSynchronize(c->ctx->stream);
    hipFree(c->xyzi);
    hipFree(c->rgb);
    hipFree(c->nrm);
    hipFree(c->grad);
    hipFree(c->hdr);
    grid_free(c->grid);
    delete c;
    return GCA_OK;
}

int32_t gca_cloud_size(gca_cloud c, int64_t *n)
{
    if (!c || !n) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_sync_header(c));
    *n = c->host_hdr.n;
    return GCA_OK;
}

int32_t gca_cloud_capacity(gca_cloud c, int64_t *cap)
{
    if (!c || !cap) return GCA_ERR_INVALID_VALUE;
    *cap = c->cap;
    return GCA_OK;
}

int32_t gca_cloud_has_normals(gca_cloud c, int32_t *flag)
{
    if (!c || !flag) return GCA_ERR_INVALID_VALUE;
    *flag = c->has_normals ? 1 : 0;
    return GCA_OK;
}

int32_t gca_cloud_upload(gca_cloud c, const float *xyz, const float *rgb, int64_t n)
{
    if (!c || n < 0 || (n > 0 && (!xyz || !rgb))) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_reserve(c, std::max<int64_t>(n, 1)));
    std::vector<float4> a((size_t)n), b((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        float r = rgb[3 * i], g = rgb[3 * i + 1], bb = rgb[3 * i + 2];
        // color3::to_intensity (type.hpp:45-51), evaluated exactly as on the device
        float I = (float)(0.2126 * (double)r + 0.7152 * (double)g + 0.0722 * (double)bb);
        a[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], I);
        b[i] = make_float4(r, g, bb, 0.0f);
    }
    GCA_HIP(hipStreamSynchronize(c->ctx->stream));
    if (n > 0) {
        GCA_HIP(hipMemcpy(c->xyzi, a.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        GCA_HIP(hipMemcpy(c->rgb, b.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    }
    int32_t n32 = (int32_t)n;
    GCA_HIP(hipMemcpy(&c->hdr->n, &n32, sizeof(int32_t), hipMemcpyHostToDevice));
    c->has_normals = c->has_gradient = false;
    cloud_points_changed(c);
    GCA_TRY(launch_bounds(c->ctx, c));
    return GCA_OK;
}

__global__ void k_set_count(CloudHdr *h, int32_t n) { h->n = n; }

int32_t gca_cloud_upload_device(gca_ctx ctx, gca_cloud c, const void *xyzi, const void *rgb, int64_t n)
{
    if (!ctx || !c || n < 0 || (n > 0 && !xyzi) || n > (int64_t)INT_MAX / 4) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_reserve(c, std::max<int64_t>(n, 1)));
    if (n > 0) {
        GCA_HIP(hipMemcpyAsync(c->xyzi, xyzi, sizeof(float4) * n, hipMemcpyDeviceToDevice, ctx->stream));
        if (rgb)
            GCA_HIP(hipMemcpyAsync(c->rgb, rgb, sizeof(float4) * n, hipMemcpyDeviceToDevice, ctx->stream));
        else
            GCA_HIP(hipMemsetAsync(c->rgb, 0, sizeof(float4) * n, ctx->stream));
    }
    hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, ctx->stream, c->hdr, (int32_t)n);
    GCA_HIP(hipGetLastError());
    c->has_normals = c->has_gradient = false;
    cloud_points_changed(c);
    GCA_TRY(launch_bounds(ctx, c));
    return GCA_OK;
}

static int32_t upload_vec3(gca_cloud c, float4 *dst, const float *v, int64_t n)
{
    int64_t have;
    GCA_TRY(gca_cloud_size(c, &have));
    if (n != have || (n > 0 && !v)) return GCA_ERR_INVALID_VALUE;
    GCA_HIP(hipStreamSynchronize(c->ctx->stream));
    std::vector<float4> a((size_t)n);
    for (int64_t i = 0; i < n; ++i) a[i] = make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0f);
    if (n > 0) GCA_HIP(hipMemcpy(dst, a.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    return GCA_OK;
}

int32_t gca_cloud_upload_normals(gca_cloud c, const float *normals, int64_t n)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(upload_vec3(c, c->nrm, normals, n));
    c->has_normals = true;
    c->has_gradient = false;
    c->grid.attrs_valid = false;
    return GCA_OK;
}

int32_t gca_cloud_upload_gradient(gca_cloud c, const float *gradient, int64_t n)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(upload_vec3(c, c->grad, gradient, n));
    c->has_gradient = true;
    c->grid.attrs_valid = false;
    return GCA_OK;
}

int32_t gca_cloud_download(gca_cloud c, float *xyz, float *rgb, float *intensity, float *normals,
                           float *gradient)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    int64_t n;
    GCA_TRY(gca_cloud_size(c, &n));
    // the ctx stream is non-blocking: wait for queued producers before reading on the host
    GCA_HIP(hipStreamSynchronize(c->ctx->stream));
    std::vector<float4> a((size_t)n);
    auto get = [&](const float4 *src, float *dst, bool with_w) -> int32_t {
        if (!dst || n == 0) return GCA_OK;
        GCA_HIP(hipMemcpy(a.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            dst[3 * i] = a[i].x;
            dst[3 * i + 1] = a[i].y;
            dst[3 * i + 2] = a[i].z;
        }
        (void)with_w;
        return GCA_OK;
    };
    GCA_TRY(get(c->xyzi, xyz, false));
    if (intensity && n > 0) {
        GCA_HIP(hipMemcpy(a.data(), c->xyzi, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) intensity[i] = a[i].w;
    }
    GCA_TRY(get(c->rgb, rgb, false));
    if (normals && !c->has_normals) return GCA_ERR_NOT_READY;
    GCA_TRY(get(c->nrm, normals, false));
    if (gradient && !c->has_gradient) return GCA_ERR_NOT_READY;
    GCA_TRY(get(c->grad, gradient, false));
    return GCA_OK;
}

int32_t gca_cloud_device_arrays(gca_cloud c, void **xyzi, void **rgb, void **normals,
                                void **gradient, void **count)
{
    if (!c) return GCA_ERR_INVALID_VALUE;
    if (xyzi) *xyzi = c->xyzi;
    if (rgb) *rgb = c->rgb;
    if (normals) *normals = c->nrm;
    if (gradient) *gradient = c->grad;
    if (count) *count = &c->hdr->n;
    return GCA_OK;
}

int32_t gca_cloud_bounds(gca_cloud c, float mn[3], float mx[3])
{
    if (!c || !mn || !mx) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_sync_header(c));
    cloud_host_bounds(c, mn, mx);
    return GCA_OK;
}

int32_t gca_cloud_copy(gca_ctx ctx, gca_cloud src, gca_cloud dst)
{
    if (!ctx || !src || !dst || src == dst) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_reserve(dst, src->cap));
    size_t b = sizeof(float4) * (size_t)src->cap;
    GCA_HIP(hipMemcpyAsync(dst->xyzi, src->xyzi, b, hipMemcpyDeviceToDevice, ctx->stream));
    GCA_HIP(hipMemcpyAsync(dst->rgb, src->rgb, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (src->has_normals)
        GCA_HIP(hipMemcpyAsync(dst->nrm, src->nrm, b, hipMemcpyDeviceToDevice, ctx->stream));
    if (src->has_gradient)
        GCA_HIP(hipMemcpyAsync(dst->grad, src->grad, b, hipMemcpyDeviceToDevice, ctx->stream));
    GCA_HIP(hipMemcpyAsync(dst->hdr, src->hdr, sizeof(CloudHdr), hipMemcpyDeviceToDevice, ctx->stream));
    dst->has_normals = src->has_normals;
    dst->has_gradient = src->has_gradient;
    cloud_points_changed(dst);
    return GCA_OK;
}

int32_t gca_cloud_append(gca_ctx ctx, gca_cloud dst, gca_cloud src)
{
    // point_cloud::operator+= (point_cloud.cu:55-73)
    if (!ctx || !dst || !src) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_sync_header(dst));
    GCA_TRY(cloud_sync_header(src));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t nd = dst->host_hdr.n, ns = src->host_hdr.n, total = nd + ns;
    if (total > (int64_t)INT_MAX / 2) return GCA_ERR_INVALID_VALUE;
    const bool normals = dst->has_normals && src->has_normals;
    if (total > dst->cap) {
        // grow, keeping dst's points (and normals when they survive)
        const int64_t cap = std::max<int64_t>(total, dst->cap + dst->cap / 2);
        const size_t b = sizeof(float4) * (size_t)cap;
        float4 *a[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int k = 0; k < 4; ++k)
            if (hipMalloc(&a[k], b) != hipSuccess) {
                for (int j = 0; j < k; ++j) hipFree(a[j]);
                return GCA_ERR_OUT_OF_MEMORY;
            }
        const size_t keep = sizeof(float4) * (size_t)nd;
        if (keep) {
            GCA_HIP(hipMemcpyAsync(a[0], dst->xyzi, keep, hipMemcpyDeviceToDevice, ctx->stream));
            GCA_HIP(hipMemcpyAsync(a[1], dst->rgb, keep, hipMemcpyDeviceToDevice, ctx->stream));
            if (normals) GCA_HIP(hipMemcpyAsync(a[2], dst->nrm, keep, hipMemcpyDeviceToDevice, ctx->stream));
        }
        GCA_HIP(hipStreamSynchronize(ctx->stream));
        hipFree(dst->xyzi);
        hipFree(dst->rgb);
        hipFree(dst->nrm);
        hipFree(dst->grad);
        dst->xyzi = a[0];
        dst->rgb = a[1];
        dst->nrm = a[2];
        dst->grad = a[3];
        dst->cap = cap;  // (when src is dst, its arrays are these new ones: the first nd entries)
    }
    if (ns > 0) {
        const size_t b = sizeof(float4) * (size_t)ns;
        GCA_HIP(hipMemcpyAsync(dst->xyzi + nd, src->xyzi, b, hipMemcpyDeviceToDevice, ctx->stream));
        GCA_HIP(hipMemcpyAsync(dst->rgb + nd, src->rgb, b, hipMemcpyDeviceToDevice, ctx->stream));
        if (normals) GCA_HIP(hipMemcpyAsync(dst->nrm + nd, src->nrm, b, hipMemcpyDeviceToDevice, ctx->stream));
    }
    dst->host_hdr.n = (int32_t)total;
    GCA_HIP(hipMemcpyAsync(&dst->hdr->n, &dst->host_hdr.n, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    GCA_TRY(launch_bounds(ctx, dst));  // get_min_bound / get_max_bound recompute them (point_cloud.cu:107-132)
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    dst->has_normals = normals;
    dst->has_gradient = false;
    cloud_points_changed(dst);
    return GCA_OK;
}

int32_t gca_cloud_from_rgbd(gca_ctx ctx, const uint16_t *depth_dev, const uint8_t *bgr_dev,
                            const gca_camera *cam, const gca_bilateral *filter, float tmin,
                            float tmax, gca_cloud out)
{
    // cuda_point_cloud_factory.cu:286-290
    if (!ctx || !out || !cam || !depth_dev || !bgr_dev) return GCA_ERR_INVALID_VALUE;
    if (cam->width <= 0 || cam->height <= 0 || !(cam->depth_scale > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if ((int64_t)cam->width * cam->height > (int64_t)INT_MAX / 4) return GCA_ERR_INVALID_VALUE;
    gca_bilateral none{GCA_BILATERAL_NONE, 0, 1.0f, 1.0f};
    const gca_bilateral *f = filter ? filter : &none;
    if (f->mode < 0 || f->mode > 2) return GCA_ERR_INVALID_VALUE;
    if (f->mode == GCA_BILATERAL_FIXED &&
        (f->radius < 0 || f->radius > 15 || !(f->sigma_space > 0.0f) || !(f->sigma_depth > 0.0f)))
        return GCA_ERR_INVALID_VALUE;
    GCA_TRY(launch_from_rgbd(ctx, depth_dev, bgr_dev, cam, f, tmin, tmax, out));
    out->has_normals = out->has_gradient = false;
    cloud_points_changed(out);
    return GCA_OK;
}

int32_t gca_voxel_downsample(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out)
{
    if (!ctx || !in || !out || in == out) return GCA_ERR_INVALID_VALUE;
    if (!(voxel > 0.0f)) return GCA_ERR_INVALID_VALUE;  // point_cloud.cu:200-204
    GCA_TRY(launch_voxel(ctx, in, voxel, out));
    out->has_normals = out->has_gradient = false;
    return GCA_OK;
}

int32_t gca_radius_outlier_removal(gca_ctx ctx, gca_cloud in, float radius, int32_t min_nn,
                                   gca_cloud out)
{
    if (!ctx || !in || !out || in == out) return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;  // point_cloud.cu:243-247
    GCA_TRY(grid_build(ctx, in, radius));
    GCA_TRY(launch_outlier(ctx, in, radius, min_nn, out));
    out->has_normals = out->has_gradient = false;
    cloud_points_changed(out);
    return GCA_OK;
}

int32_t gca_euclidean_clustering(gca_ctx ctx, gca_cloud c, float tolerance, int32_t min_cluster_size,
                                 int32_t max_cluster_size, int32_t *labels_dev, int32_t *n_clusters,
                                 int32_t *n_cut)
{
    if (!ctx || !c || !labels_dev || !n_clusters) return GCA_ERR_INVALID_VALUE;
    // cuda_clustering.cu:38-42 and point_cloud.cu:297-302
    if (min_cluster_size < 0 || max_cluster_size <= 0 || max_cluster_size < min_cluster_size)
        return GCA_ERR_INVALID_VALUE;
    if (!(tolerance > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!c->has_normals) return GCA_ERR_NOT_READY;  // point_cloud.cu:290-295
    GCA_TRY(cloud_sync_header(c));
    *n_clusters = 0;
    if (n_cut) *n_cut = 0;
    if (c->host_hdr.n == 0) return GCA_OK;
    GCA_TRY(grid_build(ctx, c, tolerance));
    uint32_t *counts = (uint32_t *)ctx_scratch(ctx, 26, 2 * sizeof(uint32_t), false);
    if (!counts) return GCA_ERR_OUT_OF_MEMORY;
    GCA_TRY(launch_euclidean_clustering(ctx, c, tolerance, (uint32_t)min_cluster_size,
                                        (uint32_t)max_cluster_size, labels_dev, counts));
    uint32_t h[2] = {0, 0};
    GCA_HIP(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    *n_clusters = (int32_t)h[0];
    if (n_cut) *n_cut = (int32_t)h[1];
    return GCA_OK;
}

int32_t gca_residual_segmentation(gca_ctx ctx, gca_cloud c, float tolerance, const float *abs_residual_dev,
                                  int32_t min_cluster_size, int32_t max_cluster_size, int32_t *labels_dev,
                                  int32_t *n_clusters, int32_t *n_cut)
{
    if (!ctx || !c || !abs_residual_dev || !labels_dev || !n_clusters) return GCA_ERR_INVALID_VALUE;
    // cuda_clustering.cu:260-264 and point_cloud.cu:397-402; no normals needed (point_cloud.cu:391-435)
    if (min_cluster_size < 0 || max_cluster_size <= 0 || max_cluster_size < min_cluster_size)
        return GCA_ERR_INVALID_VALUE;
    if (!(tolerance > 0.0f)) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_sync_header(c));
    *n_clusters = 0;
    if (n_cut) *n_cut = 0;
    if (c->host_hdr.n == 0) return GCA_OK;
    GCA_TRY(grid_build(ctx, c, tolerance));
    uint32_t *counts = (uint32_t *)ctx_scratch(ctx, 26, 2 * sizeof(uint32_t), false);
    if (!counts) return GCA_ERR_OUT_OF_MEMORY;
    GCA_TRY(launch_euclidean_clustering(ctx, c, tolerance, (uint32_t)min_cluster_size, (uint32_t)max_cluster_size,
                                        labels_dev, counts, abs_residual_dev));
    uint32_t h[2] = {0, 0};
    GCA_HIP(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    *n_clusters = (int32_t)h[0];
    if (n_cut) *n_cut = (int32_t)h[1];
    return GCA_OK;
}

int32_t gca_residual_segmentation_host(gca_ctx ctx, gca_cloud c, float tolerance, const float *abs_residual_host,
                                       int32_t min_cluster_size, int32_t max_cluster_size, int32_t *labels_host,
                                       int64_t labels_cap, int32_t *n_clusters, int32_t *n_cut)
{
    if (!ctx || !c || !n_clusters) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_sync_header(c));
    const int64_t n = c->host_hdr.n;
    if (n > 0 && (!labels_host || labels_cap < n || !abs_residual_host)) return GCA_ERR_INVALID_VALUE;
    int32_t *lab = (int32_t *)ctx_scratch(ctx, 27, sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1), false);
    float *res = (float *)ctx_scratch(ctx, 33, sizeof(float) * (size_t)std::max<int64_t>(n, 1), false);
    if (!lab || !res) return GCA_ERR_OUT_OF_MEMORY;
    if (n > 0) GCA_HIP(hipMemcpyAsync(res, abs_residual_host, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    GCA_TRY(gca_residual_segmentation(ctx, c, tolerance, res, min_cluster_size, max_cluster_size, lab, n_clusters,
                                      n_cut));
    if (n > 0) {
        GCA_HIP(hipMemcpyAsync(labels_host, lab, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
        GCA_HIP(hipStreamSynchronize(ctx->stream));
    }
    return GCA_OK;
}

int32_t gca_euclidean_clustering_host(gca_ctx ctx, gca_cloud c, float tolerance, int32_t min_cluster_size,
                                      int32_t max_cluster_size, int32_t *labels_host, int64_t labels_cap,
                                      int32_t *n_clusters, int32_t *n_cut)
{
    if (!ctx || !c || !n_clusters) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(cloud_sync_header(c));
    const int64_t n = c->host_hdr.n;
    if (n > 0 && (!labels_host || labels_cap < n)) return GCA_ERR_INVALID_VALUE;
    int32_t *lab = (int32_t *)ctx_scratch(ctx, 27, sizeof(int32_t) * (size_t)std::max<int64_t>(n, 1), false);
    if (!lab) return GCA_ERR_OUT_OF_MEMORY;
    GCA_TRY(gca_euclidean_clustering(ctx, c, tolerance, min_cluster_size, max_cluster_size, lab, n_clusters,
                                     n_cut));
    if (n > 0) {
        GCA_HIP(hipMemcpyAsync(labels_host, lab, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
        GCA_HIP(hipStreamSynchronize(ctx->stream));
    }
    return GCA_OK;
}

int32_t gca_nn_search(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t *out_dev)
{
    if (!ctx || !q || !ref || !out_dev) return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;  // point_cloud.cu:512-516
    GCA_TRY(grid_build(ctx, ref, radius));
    return launch_nn(ctx, q, ref, radius, out_dev);
}

int32_t gca_nn_search_host(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t *out_host,
                           int64_t capacity, int64_t *n_out)
{
    if (!ctx || !q || !ref || !n_out || capacity < 0) return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;  // point_cloud.cu:512-516
    GCA_TRY(cloud_sync_header(q));
    const int64_t n = q->host_hdr.n;
    *n_out = n;
    if (n > capacity || (n > 0 && !out_host)) return GCA_ERR_INVALID_VALUE;
    if (n == 0) return GCA_OK;
    int32_t *dev = (int32_t *)ctx_scratch(ctx, 21, sizeof(int32_t) * n, false);
    if (!dev) return GCA_ERR_OUT_OF_MEMORY;
    GCA_TRY(grid_build(ctx, ref, radius));
    GCA_TRY(launch_nn(ctx, q, ref, radius, dev));
    GCA_HIP(hipMemcpyAsync(out_host, dev, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    GCA_HIP(hipStreamSynchronize(ctx->stream));
    return GCA_OK;
}

int32_t gca_radius_search(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t max_nn,
                          int32_t *offsets_dev, int32_t *idx_dev, int64_t idx_capacity,
                          int64_t *total)
{
    if (!ctx || !q || !ref || !offsets_dev || !total || max_nn < 0 || idx_capacity < 0)
        return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(grid_build(ctx, ref, radius));
    return launch_radius(ctx, q, ref, radius, max_nn, offsets_dev, idx_dev, idx_capacity, total);
}

int32_t gca_grid_cells(gca_ctx ctx, gca_cloud c, float radius, int32_t *cells_dev, float origin[3],
                       float *cell_size, int32_t dims[3])
{
    if (!ctx || !c || !cells_dev) return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(grid_build(ctx, c, radius, true));
    GCA_TRY(launch_grid_cells(ctx, c, cells_dev));
    for (int a = 0; a < 3; ++a) {
        if (origin) origin[a] = c->grid.desc.origin[a];
        if (dims) dims[a] = c->grid.desc.dims[a];
    }
    if (cell_size) *cell_size = c->grid.desc.cell;
    return GCA_OK;
}

int32_t gca_estimate_normals(gca_ctx ctx, gca_cloud c, float radius, int32_t max_nn)
{
    if (!ctx || !c || max_nn < 0) return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(grid_build(ctx, c, radius));
    GCA_TRY(launch_normals(ctx, c, radius, max_nn));
    c->has_normals = true;
    c->has_gradient = false;
    c->grid.attrs_valid = false;
    return GCA_OK;
}

int32_t gca_compute_color_gradient(gca_ctx ctx, gca_cloud c, float radius, int32_t max_nn)
{
    if (!ctx || !c || max_nn < 0) return GCA_ERR_INVALID_VALUE;
    if (!(radius > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!c->has_normals) return GCA_ERR_NOT_READY;  // color_icp.cu:168-174
    GCA_TRY(grid_build(ctx, c, radius));
    GCA_TRY(launch_gradient(ctx, c, radius, max_nn));
    c->has_gradient = true;
    c->grid.attrs_valid = false;
    return GCA_OK;
}

int32_t gca_transform(gca_ctx ctx, gca_cloud c, const double T[16])
{
    if (!ctx || !c || !T) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(launch_transform(ctx, c, T));
    cloud_points_changed(c);
    GCA_TRY(launch_bounds(ctx, c));  // point_cloud.cu:96-100 recomputes the bounds
    return GCA_OK;
}

}  // extern "C"
