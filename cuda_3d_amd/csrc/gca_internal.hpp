// gca_internal.hpp -- shared internals of libgca.so (HIP, gfx950 only).
//
// Data layout in HBM (structure of arrays, 16-byte aligned float4 records so that every hot
// read is one dwordx4 load):
//   xyzi[n]  float4 (x, y, z, intensity)      rgb[n]  float4 (r, g, b, 0)
//   nrm[n]   float4 (nx, ny, nz, 0)           grad[n] float4 (dx, dy, dz, 0)
// The point count and the bounding box live in a small device header (CloudHdr) so kernels
// read sizes from device memory and launches are sized by capacity: no host round trip is
// needed between pipeline stages.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/gca/gca.h"

namespace gca {

constexpr int kWave = 64;
constexpr double kFx30 = 1073741824.0;       // 2^30: voxel sums fixed point
constexpr double kFx40 = 1099511627776.0;    // 2^40: neighbourhood sums fixed point
constexpr int64_t kGridCellCap = 1LL << 25;  // max cells of a neighbour grid

// ---------------------------------------------------------------------------------------
// Device header of a cloud.  Bounds are stored as order-preserving uint32 keys so that
// atomicMin/atomicMax on unsigned ints give float min/max.
struct CloudHdr {
    int32_t n;
    uint32_t bmin_key[3];
    uint32_t bmax_key[3];
    int32_t pad;
};

__host__ __device__ inline uint32_t float_to_key(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__host__ __device__ inline float key_to_float(uint32_t k)
{
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------------------
// Neighbour grid: dense cells of edge `cell` >= radius over the padded bounds, points
// sorted by cell (counting sort).  Cell of a point = floor(((double)p - origin) / cell) in
// FP64; a query at q scans the cells overlapped by [q - r, q + r], which makes the
// neighbour set exactly { j : d2_fp64 < r^2 } whatever the cell edge.
struct GridDesc {
    float origin[3];
    float cell;
    int32_t dims[3];
    float radius;
    int64_t ncells;
    int32_t cell_q;  // 0: canonical cell (the radius, doubled while too many cells); else the tuned
                     // edge in quarters of the radius
    int32_t pad;
};

struct Grid {
    GridDesc desc{};
    bool valid = false;
    int64_t cap_cells = 0, cap_pts = 0;
    uint32_t *count = nullptr;   // ncells, kept at zero between builds
    uint32_t *start = nullptr;   // ncells + 1 (exclusive prefix)
    int32_t *order = nullptr;    // original index of each sorted slot
    float4 *sxyzi = nullptr;     // points in cell order
    uint32_t *cell_of = nullptr; // linear cell of each original point (cap_pts), 32-bit
    float4 *snrm = nullptr;      // normals / colour gradients in cell order (registration
    float4 *sgrad = nullptr;     // targets), filled by grid_gather_attributes
    bool attrs_valid = false;
};

// Grid counting done by a voxel call's finalize for its output points (grid_prepare'd buffers).
struct GridCount {
    GridDesc d;
    uint32_t *count;
    uint32_t *cell_of;
};

struct FrontLaunch;  // a recorded front-end launch (pipeline graph updates)
struct ScanCtl;

// One launch for several pyramid levels ("jobs"): job j's arguments are a[j], blockIdx.y = j,
// and grid.x is the largest job's; a job's surplus workgroups find no work (grid-stride loops
// over device sizes) or return at once.  N = 1 for single-level calls.
constexpr int kMaxJobs = 4;
template <class A, int N>
struct Jobs {
    A a[N];
};

}  // namespace gca

struct gca_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // scratch buffers owned by the ctx, grown on demand
    static constexpr int kScratchSlots = 48;
    void *scratch[kScratchSlots] = {nullptr};
    size_t scratch_bytes[kScratchSlots] = {0};
    bool scratch_zeroed[kScratchSlots] = {false};
    // device-raised error word in mapped host memory (a bounded wait that gave up); read and
    // cleared at synchronisation points
    uint32_t *dev_err = nullptr;
    int n_cu = 0;         // compute units of the device (persistent launch sizing)
    int wall_khz = 0;     // rate of the device wall clock (wall_clock64)
    // persistent bitmap of the voxel downsampler (kept all-zero between calls)
    uint32_t *bitmap = nullptr;
    int64_t bitmap_words = 0;
    bool bitmap_dirty = false;
    // bilateral weight tables (FP64, host-computed), keyed by their parameters
    double *lut = nullptr;
    size_t lut_bytes = 0;
    int32_t lut_key_mode = -1, lut_key_radius = -1;
    float lut_key_ss = -1.f, lut_key_sd = -1.f;
    int32_t lut_len = 0;
    // ICP device state
    void *icp_state = nullptr;
    size_t icp_state_bytes = 0;
    double *partials = nullptr;
    size_t partials_bytes = 0;
    // per-level working clouds of the multi-scale driver (created on first use)
    gca_cloud level_clouds[16] = {nullptr};
    // launch-shape tuning (GCA_TUNE_*), 0 = heuristic; the version counts changes (objects that
    // bake shapes in, e.g. a frame pipeline's plans and graphs, rebuild when it moves)
    int32_t tune[GCA_TUNE_COUNT] = {};
    uint64_t tune_version = 0;
    // stage profiler: event pairs recorded on `stream` while enabled (not while the stream is
    // being captured into a graph)
    bool prof_on = false;
    bool capturing = false;
    // set by a pipeline while it captures a step: launch_from_rgbd records its pixel kernel's
    // launch there (see FrontLaunch)
    gca::FrontLaunch *front_rec = nullptr;
    // the device address of this device's static scratch of the allocation-free entry points
    // (gca::NoAllocScratch), set by gca_ctx_create
    void *noalloc_scratch = nullptr;
    gca_ctx_s *prof_owner = nullptr;  // stage records go to this ctx's profiler when set
    // work this ctx's users wait for as well: synchronising the ctx also synchronises these (a
    // frame pipeline's other streams in two-lane steps)
    gca_ctx_s *sync_also[3] = {nullptr, nullptr, nullptr};
    struct ProfRec {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> prof_pool;
    size_t prof_used = 0;
};

namespace gca {
// Brackets a stage's launches with hipEvents when the ctx profiler is on (RAII).
struct StageTimer {
    gca_ctx ctx;
    gca_ctx owner;  // the ctx whose profiler keeps the record
    int idx = -1;
    StageTimer(gca_ctx c, int stage);
    ~StageTimer();
};
}  // namespace gca

struct gca_cloud_s {
    gca_ctx ctx = nullptr;
    int64_t cap = 0;
    float4 *xyzi = nullptr, *rgb = nullptr, *nrm = nullptr, *grad = nullptr;
    gca::CloudHdr *hdr = nullptr;
    bool has_normals = false, has_gradient = false;
    // host copy of the header (valid when host_valid)
    bool host_valid = false;
    gca::CloudHdr host_hdr{};
    // cached neighbour grid (built for grid.desc.radius); invalidated when points change
    gca::Grid grid;
};

namespace gca {

// Internal helpers implemented in api.hip -------------------------------------------------
int32_t hip_status(hipError_t e);
void *ctx_scratch(gca_ctx ctx, int slot, size_t bytes, bool zeroed);
// GCA_ERR_DEVICE (and clears it) when a kernel raised the ctx's device error word.
int32_t ctx_take_device_error(gca_ctx ctx);
int32_t cloud_reserve(gca_cloud c, int64_t cap);
int32_t cloud_sync_header(gca_cloud c);  // copies hdr to host (synchronises the stream)
void cloud_host_bounds(gca_cloud c, float mn[3], float mx[3]);
void cloud_points_changed(gca_cloud c);
// (Re)builds c->grid so that it serves queries of `radius`.  A cached grid whose cell edge lies
// in [radius, 1.5 radius] is reused (results do not depend on the cell edge); exact = build
// for exactly this radius (canonical cell binning).
int32_t grid_build(gca_ctx ctx, gca_cloud c, float radius, bool exact = false);
// Grid descriptor of c->grid for queries of `radius`.
inline GridDesc query_desc(const Grid &G, float radius)
{
    GridDesc d = G.desc;
    d.radius = radius;
    return d;
}

// The launch of the front end's pixel kernel (the one kernel that reads the frame buffers),
// recorded while a pipeline step is captured, so that the instantiated graph can be pointed at
// other frame buffers by updating that kernel node alone.
struct FrontLaunch;
FrontLaunch *front_launch_new();
void front_launch_free(FrontLaunch *f);
// The recorded launch's kernel node in g, or nullptr.
hipGraphNode_t front_launch_node(const FrontLaunch *f, hipGraph_t g);
// exec's instance of that node reads depth / bgr from now on (launches already made keep theirs).
hipError_t front_launch_update(FrontLaunch *f, hipGraphExec_t exec, hipGraphNode_t node, const uint16_t *depth,
                               const uint8_t *bgr);

// Static device scratch of the allocation-free entry points: one per device, at the address
// noalloc_setup resolves once per device (noalloc.hip); the library never allocates it.
constexpr int kNoAllocMaxBlocks = 1024;  // partials of up to this many workgroups
struct NoAllocScratch {
    alignas(256) uint32_t ticket;                // last-block ticket, zero between calls
    alignas(256) double partials[kNoAllocMaxBlocks * 32];
};
// Device address of `device`'s scratch (resolved on the first call for that device, with the
// device current); nullptr if the runtime refuses.
void *noalloc_setup(int device);

// Kernel launchers (one translation unit each) -------------------------------------------
// clear_word (optional): a device word the first kernel sets to zero (a caller's per-step flag).
int32_t launch_from_rgbd(gca_ctx ctx, const uint16_t *depth, const uint8_t *bgr,
                         const gca_camera *cam, const gca_bilateral *bil, float tmin, float tmax,
                         gca_cloud out, uint32_t *clear_word = nullptr);
int32_t launch_voxel(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out);
// No host round trip: the lattice is computed on the device; the bitmap scan covers max_bits
// cells (a lattice beyond that flags *err and gives an empty output); scratch by capacity.
int32_t launch_voxel_planned(gca_ctx ctx, gca_cloud in, float voxel, gca_cloud out, double max_bits,
                             uint32_t *err, const GridCount *gc = nullptr);
// Grid buffers of c for geometry d (allocation only; no launch).
int32_t grid_prepare(gca_ctx ctx, gca_cloud c, const GridDesc &d);
// counted: the points' cells and counts were already produced (GridCount of a voxel call).
int32_t launch_grid_build(gca_ctx ctx, gca_cloud c, const GridDesc &d, bool counted = false);
int32_t launch_nn(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t *out);
int32_t launch_radius(gca_ctx ctx, gca_cloud q, gca_cloud ref, float radius, int32_t max_nn,
                      int32_t *offsets, int32_t *idx, int64_t idx_cap, int64_t *total);
int32_t launch_grid_cells(gca_ctx ctx, gca_cloud c, int32_t *cells);
int32_t launch_outlier(gca_ctx ctx, gca_cloud in, float radius, int32_t min_nn, gca_cloud out);
// labels of the cloud's own grid (built at `tolerance`); counts_dev[0] = clusters, [1] =
// components cut at max_size.  res (device, one float per point, may be null): residual gate of
// the residual-based segmentation.  Synchronises (the cut is planned on the host).  The header's
// host copy must be current.
int32_t launch_euclidean_clustering(gca_ctx ctx, gca_cloud c, float tolerance, uint32_t min_size,
                                    uint32_t max_size, int32_t *label, uint32_t *counts_dev,
                                    const float *res = nullptr);
// Copies normals and colour gradients into the grid's cell order (c->grid.snrm / sgrad).
int32_t grid_gather_attributes(gca_ctx ctx, gca_cloud c);
int32_t launch_normals(gca_ctx ctx, gca_cloud c, float radius, int32_t max_nn);
int32_t launch_gradient(gca_ctx ctx, gca_cloud c, float radius, int32_t max_nn);
// Normals (radius, max_nn) and colour gradient (gradient_radius <= radius, all neighbours) from
// one neighbourhood walk; results equal launch_normals followed by launch_gradient.  grid_attrs:
// the solve also writes the grid-ordered copies (as grid_gather_attributes would).
int32_t launch_normals_gradient(gca_ctx ctx, gca_cloud c, float radius, int32_t max_nn, float gradient_radius,
                                bool grid_attrs = false);
int32_t launch_transform(gca_ctx ctx, gca_cloud c, const double T[16]);
int32_t launch_bounds(gca_ctx ctx, gca_cloud c);
int32_t launch_copy_cloud(gca_ctx ctx, gca_cloud src, gca_cloud dst);
int32_t launch_upload_finish(gca_ctx ctx, gca_cloud c, int64_t n);

// Registration plans (icp.hip): n_pairs chains of n_levels states set up once (descriptors,
// buffers, pinned result copies); each run enqueues set-up + the registration launches + the
// result copy without a host round trip, so it can be captured in a graph.  Results are read
// after the ctx stream has been synchronised.
struct IcpPlan;
int32_t icp_plan_create(gca_ctx ctx, int n_pairs, int n_levels, const gca_cloud *srcs, const gca_cloud *tgts,
                        const gca_icp_params *params, const float *grad_r, const double *T0, IcpPlan **out);
int32_t icp_plan_run(gca_ctx ctx, IcpPlan *plan);
// Whether every cloud still has the buffers (src: pointers and capacity src_cap[s]; tgt: arrays,
// grid and gathered attributes) the plan's descriptors were made from.
bool icp_plan_clouds_current(const IcpPlan *plan, const gca_cloud *srcs, const int64_t *src_cap,
                             const gca_cloud *tgts);
// The same launch in parts (state set-up; one state's persistent launch, which takes the
// state's sizes from its clouds' headers, so it may be enqueued before they are final; results),
// for overlapping state s's registration with the building of state s + 1's source.
int32_t icp_plan_init(gca_ctx ctx, IcpPlan *plan);
// Whether state s's target cloud still has the grid (geometry and arrays) the plan was made
// with: a plan reads the target's grid through pointers and a geometry fixed at its creation.
bool icp_plan_target_current(const IcpPlan *plan, int s, gca_cloud tgt);
bool grid_same_geometry(const GridDesc &a, const GridDesc &b);
int32_t icp_plan_persist(gca_ctx ctx, IcpPlan *plan, int s);
// Results into the plan's pinned host memory (one kernel); extra: an optional caller word copied
// to host_extra (pinned) by the same kernel.
int32_t icp_plan_copy_results(gca_ctx ctx, IcpPlan *plan, const uint32_t *extra = nullptr,
                              uint32_t *host_extra = nullptr);
int32_t icp_plan_results(gca_ctx ctx, const IcpPlan *plan, double *T_out, gca_icp_report *reports);
void icp_plan_destroy(IcpPlan *plan);

// Host-side canonical grid geometry (same FP32 formulas as the oracle)
// cell_quarters > 4: cell edge = r * cell_quarters / 4 (a tuned cell; results do not change).
void grid_params(const float bmin[3], const float bmax[3], float r, GridDesc *d, int32_t cell_quarters = 0);

// Exclusive scan of uint32 arrays in one pass (decoupled look-back)
// out has n + 1 entries (out[n] = total); popcount = scan the popcounts of the words.
int32_t exclusive_scan_u32(gca_ctx ctx, const uint32_t *in, uint32_t *out, int64_t n,
                           uint32_t *total_dev, bool popcount = false);
// The same scan as one job of a multi-level launch: its look-back words are ctx's scratch;
// wide = 1024-thread chunks (scan_wide(n) is the single-call choice).
struct ScanJob {
    const uint32_t *in;
    int64_t n;
    int32_t nchunks, pad;
    unsigned long long *flags;
    ScanCtl *ctl;
    uint32_t *out, *total_dev, *out_end, *dev_err;
};
bool scan_wide(int64_t n);
int32_t scan_job(gca_ctx ctx, const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total_dev, bool wide,
                 ScanJob *job);
// jobs[0, nj) in one launch on stream (nj <= kMaxJobs, all of one kind).
int32_t scan_launch(hipStream_t stream, const ScanJob *jobs, int nj, bool popcount, bool wide);

// Lockstep build of several levels (each kernel of the sequence launched once for all levels, on
// the launch ctx's stream; scratch from each level's own ctx, whose stream must be the launch
// stream while these run).  Results equal the single-level calls'.
struct VoxelLevelJob {
    gca_ctx ctx;  // the level's scratch
    gca_cloud in, out;
    float voxel;
    double max_bits;
    uint32_t *err;
    GridCount gc;  // the grid count of the output points (grid_prepare'd)
};
int32_t launch_voxel_levels(gca_ctx launch, const VoxelLevelJob *jobs, int nj);
// Counted grids (the voxel finalize did the count pass).
int32_t launch_grid_levels(gca_ctx launch, const gca_ctx *ctxs, const gca_cloud *cs, const GridDesc *ds, int nj);
struct NormalsLevelJob {
    gca_ctx ctx;
    gca_cloud c;
    float radius;
    int32_t max_nn;
    float gradient_radius;  // <= radius
};
// launch_normals_gradient(..., grid_attrs = true) per level.
int32_t launch_normals_gradient_levels(gca_ctx launch, const NormalsLevelJob *jobs, int nj);
// Stable LSD radix sort of n (64-bit key, 32-bit value) pairs by the low `bits` key bits, 8 per
// pass; the result ends in (k0, v0) if *sorted_in_first, else in (k1, v1).
int32_t radix_sort_u64(gca_ctx ctx, unsigned long long *k0, uint32_t *v0, unsigned long long *k1, uint32_t *v1,
                       int64_t n, int bits, bool *sorted_in_first);

}  // namespace gca

#define GCA_HIP(expr)                                   \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return gca::hip_status(_e); \
    } while (0)

#define GCA_TRY(expr)                  \
    do {                               \
        int32_t _s = (expr);           \
        if (_s != GCA_OK) return _s;   \
    } while (0)
