// pipeline.hip -- the per-frame RGB-D odometry step as one native object, replayable as a graph.
//
// Replaces the loop of the reference's demo (apps/app.cpp:111-372): create_from_rgbd ->
// voxel_grid_down_sample -> estimate_normals -> color_icp.align against the previous frame,
// with the coarse-to-fine schedule of apps/test_color_icp.cpp:378-414.  A step enqueues,
// without any host round trip:
//   front end (bilateral, back-projection, compaction, bounds), on the pipeline's stream;
//   per level, on the level's own stream (the levels only read the front end's cloud, so they
//              run side by side: the small coarse levels fill the device next to the fine one):
//              voxel lattice from the device header + downsample, neighbour grid of fixed
//              geometry, normals and colour gradient (one neighbourhood walk when the gradient's
//              radius is the smaller and it has no K cap; its solve then also writes the
//              grid-ordered attributes, else one gather kernel does);
//   registration of the new frame onto the previous one, level by level (coarse to fine): a
//              registration plan built once per frame parity is set up before the fork; the
//              persistent launches (each takes its level's sizes from the built clouds' headers)
//              run one after another on the coarsest level's stream, each once its level is
//              built (so the coarse levels register while the fine ones are still being built,
//              and a launch follows the one handing it the pose with no cross-stream wait);
//              one small kernel right after the last level's launch, on the same stream, stores
//              the results and the step's error words into pinned host memory.
// Nothing in a step depends on host-side sizes: the voxel lattice is computed on the device
// (its scan covers an upper bound from the camera frustum), and each level's grid uses one
// geometry over the frustum box (neighbour results do not depend on the grid geometry: points
// and query ranges are clamped alike).  So after one eager step per frame parity, each parity's
// step is captured into a hipGraph and replayed.
//
// Overlapped steps register the previous frame beside the new frame's build.  When the levels
// can be built in lockstep (each build kernel launched once for all levels), consecutive steps
// alternate between two lanes, each with its own pipeline stream, registration stream and level
// scratch, over four frame slots; each step is then two single-stream graphs (the build, the
// registration), ordered across lanes by events the host waits for and records around them, so
// two frames' steps also run side by side.
//
// The pipeline owns private ctxs (one for the front end and registration, one per level, each
// with its stream and scratch): the scratch buffers a captured graph points into are used by
// nothing else, so no other call can move them.  Every cloud belongs to the main ctx; a level's
// kernels run on its level ctx, ordered by the fork / join events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "gca_device.hpp"
#include "gca_internal.hpp"

using namespace gca;

constexpr int kPipeMaxLevels = 8;
constexpr int kPipeSlots = 4;  // frame slots: 2 sequential, 3 overlapped in one lane, 4 in two lanes
constexpr int kErrStride = 16;  // uint32 words between the slots' error words (own 64-byte lines)

struct gca_pipeline_s {
    gca_ctx ctx = nullptr;  // private: front end, registration, every cloud
    gca_ctx lctx[kPipeMaxLevels] = {};  // private, one per level (stream + scratch)
    hipEvent_t ev_fork = nullptr, ev_join[kPipeMaxLevels] = {};
    hipEvent_t ev_built[kPipeMaxLevels] = {};  // level l's source built
    gca_camera cam{};
    gca_bilateral bil{};
    float tmin = 0.f, tmax = 0.f;
    int nl = 0;
    gca_level_params lv[kPipeMaxLevels]{};
    gca_ctx rctx = nullptr;  // private: the registration of the previous frame in overlapped steps
    hipEvent_t ev_rfork = nullptr, ev_rjoin = nullptr;
    gca_cloud full[kPipeSlots] = {};
    gca_cloud lvc[kPipeSlots][kPipeMaxLevels] = {};
    GridDesc gdesc[kPipeMaxLevels]{};
    double max_bits[kPipeMaxLevels] = {};
    IcpPlan *plan[kPipeSlots] = {};  // plan[s]: the frame in slot s onto the frame in slot prev(s)
    uint32_t *err_dev = nullptr;     // per slot (kErrStride words apart): voxel lattice refusals of its build
    uint32_t *err_host = nullptr;    // pinned copies
    int64_t frames = 0;
    int nslots = 2;                  // frame slots in use: 2 sequential, 3 overlapped
    bool overlap = false;
    int last = -1;                   // slot of the frame whose registration result() reports
    int pending = -1;                // overlapped: slot of the stepped frame not yet registered
    bool use_graph = false;
    int eager_since_reset = 0;  // eager steps since plans were (re)built
    uint64_t tune_version = 0;
    hipGraphExec_t exec[kPipeSlots] = {};
    hipGraph_t graph[kPipeSlots] = {};          // kept for its front-end node
    gca::FrontLaunch *front[kPipeSlots] = {};  // the front end's recorded launch per slot
    hipGraphNode_t front_node[kPipeSlots] = {};
    const void *cap_depth[kPipeSlots] = {}, *cap_bgr[kPipeSlots] = {};
    int64_t captures = 0, updates = 0;
    // overlapped steps in two lanes (lane 0 is ctx / rctx / lctx): lane 1's pipeline stream,
    // registration stream and level scratch (created on first use), and the cross-lane events
    int lanes = 1;
    bool lane1_ready = false;
    gca_ctx ctx1 = nullptr, rctx1 = nullptr;
    gca_ctx lctx1[kPipeMaxLevels] = {};
    hipEvent_t ev_slot_built[kPipeSlots] = {};  // slot s's build done (on its lane)
    hipEvent_t ev_slot_reg[kPipeSlots] = {};    // the registration of slot s's frame done
    bool slot_built_recorded[kPipeSlots] = {}, slot_reg_recorded[kPipeSlots] = {};
    hipGraphExec_t exec_reg[kPipeSlots] = {};  // two lanes: the registration of slot s's frame
    hipGraph_t graph_reg[kPipeSlots] = {};
};

namespace {

int prev_slot(const gca_pipeline_s *p, int s) { return (s + p->nslots - 1) % p->nslots; }

// The streams and scratch one lane of steps runs on.
struct Lane {
    gca_ctx main, reg;
    gca_ctx *levels;
};
Lane lane(gca_pipeline p, int j) { return j ? Lane{p->ctx1, p->rctx1, p->lctx1} : Lane{p->ctx, p->rctx, p->lctx}; }
// Two lanes: slots 0, 2 on lane 0 and 1, 3 on lane 1 (step k on lane k mod 2).
int lane_of(const gca_pipeline_s *p, int slot) { return p->lanes == 2 ? (slot & 1) : 0; }

// Every ctx of the pipeline (both lanes), for f(ctx).
template <class F>
void each_ctx(gca_pipeline p, F f)
{
    gca_ctx all[2 + 2 * kPipeMaxLevels + 2] = {p->ctx, p->rctx, p->ctx1, p->rctx1};
    int n = 4;
    for (int l = 0; l < kPipeMaxLevels; ++l) {
        all[n++] = p->lctx[l];
        all[n++] = p->lctx1[l];
    }
    for (int i = 0; i < n; ++i)
        if (all[i]) f(all[i]);
}

int32_t sync_all(gca_pipeline p)
{
    hipError_t e = hipSuccess;
    each_ctx(p, [&](gca_ctx c) {
        const hipError_t e1 = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = e1;
    });
    return e == hipSuccess ? GCA_OK : hip_status(e);
}

uint32_t *err_dev_of(gca_pipeline p, int s) { return p->err_dev + kErrStride * s; }
uint32_t *err_host_of(gca_pipeline p, int s) { return p->err_host + kErrStride * s; }

void drop_graphs(gca_pipeline p)
{
    for (int s = 0; s < kPipeSlots; ++s) {
        if (p->exec[s]) hipGraphExecDestroy(p->exec[s]);
        if (p->graph[s]) hipGraphDestroy(p->graph[s]);
        p->exec[s] = nullptr;
        p->graph[s] = nullptr;
        p->front_node[s] = nullptr;
        p->cap_depth[s] = p->cap_bgr[s] = nullptr;
        if (p->exec_reg[s]) hipGraphExecDestroy(p->exec_reg[s]);
        if (p->graph_reg[s]) hipGraphDestroy(p->graph_reg[s]);
        p->exec_reg[s] = nullptr;
        p->graph_reg[s] = nullptr;
    }
}

void drop_plans(gca_pipeline p)
{
    drop_graphs(p);
    for (int s = 0; s < kPipeSlots; ++s) {
        icp_plan_destroy(p->plan[s]);
        p->plan[s] = nullptr;
    }
    p->eager_since_reset = 0;
}

int32_t make_plan(gca_pipeline p, int slot)
{
    gca_cloud srcs[kPipeMaxLevels], tgts[kPipeMaxLevels];
    gca_icp_params prm[kPipeMaxLevels];
    float gr[kPipeMaxLevels];
    for (int l = 0; l < p->nl; ++l) {
        srcs[l] = p->lvc[slot][l];
        tgts[l] = p->lvc[prev_slot(p, slot)][l];
        // the plan keeps its target's grid geometry for every later step of this parity, while
        // the target is rebuilt each step with the pipeline's geometry: a target built before a
        // geometry change (a tuned cell edge) is re-binned into the current geometry first
        gca_cloud t = tgts[l];
        if (t->grid.valid && !grid_same_geometry(t->grid.desc, p->gdesc[l])) {
            GCA_TRY(launch_grid_build(p->ctx, t, p->gdesc[l], false));
            GCA_TRY(grid_gather_attributes(p->ctx, t));
        }
        prm[l] = p->lv[l].icp;
        gr[l] = p->lv[l].gradient_radius;
    }
    static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    return icp_plan_create(p->ctx, 1, p->nl, srcs, tgts, prm, gr, I, &p->plan[slot]);
}

// The registration plan's launches for the frame in slot reg onto the frame before it, on rc.
int32_t registration_launches(gca_pipeline p, int reg, gca_ctx rc)
{
    IcpPlan *plan = p->plan[reg];
    GCA_TRY(icp_plan_init(rc, plan));
    for (int l = 0; l < p->nl; ++l) {
        // the plan reads the target's grid through what it saw at its creation
        if (!icp_plan_target_current(plan, l, p->lvc[prev_slot(p, reg)][l])) return GCA_ERR_DEVICE;
        GCA_TRY(icp_plan_persist(rc, plan, l));
    }
    // results and the registered frame's lattice error word (its build ended a step earlier)
    return icp_plan_copy_results(rc, plan, err_dev_of(p, reg), err_host_of(p, reg));
}

// One lane: the registration of the frame in slot reg onto the frame before it, on the
// registration stream, forked from the pipeline stream (so after everything enqueued there
// before); the caller makes the pipeline stream wait for ev_rjoin.  Overlapped steps only.
int32_t enqueue_registration(gca_pipeline p, int reg)
{
    GCA_HIP(hipEventRecord(p->ev_rfork, p->ctx->stream));
    GCA_HIP(hipStreamWaitEvent(p->rctx->stream, p->ev_rfork, 0));
    GCA_TRY(registration_launches(p, reg, p->rctx));
    GCA_HIP(hipEventRecord(p->ev_rjoin, p->rctx->stream));
    return GCA_OK;
}

// Two lanes: the cross-lane order of a step's two parts (the registration, the build), as
// events the host waits for and records around each part (the parts themselves are plain kernel
// sequences on one stream each, replayed as graphs).  A wait for an event never recorded yet (the
// first steps) is skipped.
int32_t wait_slot(gca_ctx c, hipEvent_t e, bool recorded)
{
    if (recorded) GCA_HIP(hipStreamWaitEvent(c->stream, e, 0));
    return GCA_OK;
}

// Before the registration of the frame in slot reg (built on the other lane) onto the frame before
// it (built on this lane), on lane j's registration stream: both builds done.
int32_t before_registration(gca_pipeline p, int reg, int j)
{
    gca_ctx rc = lane(p, j).reg;
    const int tgt = prev_slot(p, reg);
    GCA_TRY(wait_slot(rc, p->ev_slot_built[reg], p->slot_built_recorded[reg]));
    return wait_slot(rc, p->ev_slot_built[tgt], p->slot_built_recorded[tgt]);
}

int32_t after_registration(gca_pipeline p, int reg, int j)
{
    GCA_HIP(hipEventRecord(p->ev_slot_reg[reg], lane(p, j).reg->stream));
    p->slot_reg_recorded[reg] = true;
    return GCA_OK;
}

// Before the build of slot `slot` (frame k) on lane j's pipeline stream: the two registrations
// that last read the slot's clouds are done (frame k - 4's own, run on the other lane, and frame
// k - 3's, which read it as its target, run on this lane's registration stream).
int32_t before_build(gca_pipeline p, int slot, int j)
{
    gca_ctx m = lane(p, j).main;
    const int next = (slot + 1) % p->nslots;
    GCA_TRY(wait_slot(m, p->ev_slot_reg[slot], p->slot_reg_recorded[slot]));
    return wait_slot(m, p->ev_slot_reg[next], p->slot_reg_recorded[next]);
}

int32_t after_build(gca_pipeline p, int slot, int j)
{
    GCA_HIP(hipEventRecord(p->ev_slot_built[slot], lane(p, j).main->stream));
    p->slot_built_recorded[slot] = true;
    return GCA_OK;
}

// Lockstep build: each kernel of the levels' build (voxel, grid, joint normals + gradient)
// launched once for all levels, on one stream.  Possible when every level fits one multi-level
// launch; overlapped steps then run in two lanes (each a pipeline stream and a registration
// stream: per-level streams in two lanes would need more hardware queues than the device gives
// a process).
bool lockstep_ok(const gca_pipeline_s *p)
{
    if (p->nl > kMaxJobs) return false;
    for (int l = 0; l < p->nl; ++l)
        if (p->lv[l].gradient_max_nn != 0 || !(p->lv[l].gradient_radius <= p->lv[l].normal_radius)) return false;
    return true;
}

// While it lives, the level ctxs enqueue on stream s (the lockstep build's scratch set-up:
// first-use clears and re-allocations are then ordered with its launches).
struct LevelStreamLease {
    gca_ctx *lc;
    int n;
    hipStream_t saved[kPipeMaxLevels] = {};
    LevelStreamLease(gca_ctx *levels, int nl, hipStream_t s) : lc(levels), n(nl)
    {
        for (int l = 0; l < n; ++l) {
            saved[l] = lc[l]->stream;
            lc[l]->stream = s;
        }
    }
    ~LevelStreamLease()
    {
        for (int l = 0; l < n; ++l) lc[l]->stream = saved[l];
    }
};

// Slot `slot`'s levels built in lockstep on lane j's pipeline stream, with its level scratch.
int32_t build_lockstep(gca_pipeline p, int slot, int j)
{
    const Lane L = lane(p, j);
    gca_ctx ctx = L.main;
    LevelStreamLease lease(L.levels, p->nl, ctx->stream);
    VoxelLevelJob vj[kMaxJobs];
    NormalsLevelJob nj[kMaxJobs];
    gca_cloud cs[kMaxJobs];
    for (int l = 0; l < p->nl; ++l) {
        gca_ctx lc = L.levels[l];
        gca_cloud c = p->lvc[slot][l];
        const gca_level_params &V = p->lv[l];
        GCA_TRY(cloud_reserve(c, std::max<int64_t>(p->full[slot]->cap, 1)));
        GCA_TRY(grid_prepare(lc, c, p->gdesc[l]));
        vj[l] = VoxelLevelJob{lc, p->full[slot], c, V.voxel_size, p->max_bits[l], err_dev_of(p, slot),
                              GridCount{p->gdesc[l], c->grid.count, c->grid.cell_of}};
        nj[l] = NormalsLevelJob{lc, c, V.normal_radius, V.normal_max_nn, V.gradient_radius};
        cs[l] = c;
    }
    GCA_TRY(launch_voxel_levels(ctx, vj, p->nl));
    GCA_TRY(launch_grid_levels(ctx, L.levels, cs, p->gdesc, p->nl));
    GCA_TRY(launch_normals_gradient_levels(ctx, nj, p->nl));
    for (int l = 0; l < p->nl; ++l) {
        cs[l]->has_normals = cs[l]->has_gradient = true;
        GCA_TRY(grid_gather_attributes(ctx, cs[l]));  // no launch: the solve wrote them
    }
    return GCA_OK;
}

// Two lanes (step k on lane k mod 2): frame k's front end and lockstep build into slot `slot` on
// lane j's pipeline stream.  It runs beside the registration of frame k - 1 (this lane's
// registration stream) and beside the other lane's step; before_build / after_build order it
// across lanes.
int32_t lane_build(gca_pipeline p, int slot, int j, const uint16_t *depth, const uint8_t *bgr)
{
    const Lane L = lane(p, j);
    GCA_TRY(launch_from_rgbd(L.main, depth, bgr, &p->cam, &p->bil, p->tmin, p->tmax, p->full[slot],
                             err_dev_of(p, slot)));
    return build_lockstep(p, slot, j);
}

// Everything a step enqueues (no host round trip, no allocation once warm): the new frame's
// front end and every level's build into slot `slot`, and the registration of the frame in slot
// `reg` (-1: none).  Sequential steps (reg == slot): each level registers onto the previous frame
// as soon as it is built.  Overlapped steps (reg: the previous frame, built by the previous
// step): its registration runs on the registration stream beside this frame's front end and
// build, which touch neither of the two frames it reads.
int32_t enqueue(gca_pipeline p, int slot, int reg, const uint16_t *depth, const uint8_t *bgr)
{
    if (p->lanes == 2) {  // eager two-lane step: the registration, then the build (see gca_pipeline_step)
        const int j = lane_of(p, slot);
        if (reg >= 0) {
            GCA_TRY(before_registration(p, reg, j));
            GCA_TRY(registration_launches(p, reg, lane(p, j).reg));
            GCA_TRY(after_registration(p, reg, j));
        }
        GCA_TRY(before_build(p, slot, j));
        GCA_TRY(lane_build(p, slot, j, depth, bgr));
        return after_build(p, slot, j);
    }
    gca_ctx ctx = p->ctx;
    const bool overlapped = p->overlap && reg >= 0;
    IcpPlan *plan = reg >= 0 && !p->overlap ? p->plan[reg] : nullptr;  // the inline chain
    if (overlapped) GCA_TRY(enqueue_registration(p, reg));
    // the front end's first kernel also clears this slot's voxel-lattice error word
    GCA_TRY(launch_from_rgbd(ctx, depth, bgr, &p->cam, &p->bil, p->tmin, p->tmax, p->full[slot], err_dev_of(p, slot)));
    if (plan) GCA_TRY(icp_plan_init(ctx, plan));
    GCA_HIP(hipEventRecord(p->ev_fork, ctx->stream));
    auto build = [&](int l) -> int32_t {
        gca_ctx lc = p->lctx[l];
        gca_cloud c = p->lvc[slot][l];
        const gca_level_params &L = p->lv[l];
        GCA_HIP(hipStreamWaitEvent(lc->stream, p->ev_fork, 0));
        // the voxel call's finalize also does the grid's count pass for its output points
        GCA_TRY(cloud_reserve(c, std::max<int64_t>(p->full[slot]->cap, 1)));
        GCA_TRY(grid_prepare(lc, c, p->gdesc[l]));
        const GridCount gc{p->gdesc[l], c->grid.count, c->grid.cell_of};
        GCA_TRY(launch_voxel_planned(lc, p->full[slot], L.voxel_size, c, p->max_bits[l], err_dev_of(p, slot), &gc));
        GCA_TRY(launch_grid_build(lc, c, p->gdesc[l], true));
        if (L.gradient_max_nn == 0 && L.gradient_radius <= L.normal_radius) {
            // the solve writes the grid-ordered copies registration reads
            GCA_TRY(launch_normals_gradient(lc, c, L.normal_radius, L.normal_max_nn, L.gradient_radius, true));
        } else {
            GCA_TRY(launch_normals(lc, c, L.normal_radius, L.normal_max_nn));
            GCA_TRY(launch_gradient(lc, c, L.gradient_radius, L.gradient_max_nn));
            c->grid.attrs_valid = false;
        }
        c->has_normals = c->has_gradient = true;
        GCA_TRY(grid_gather_attributes(lc, c));  // no launch when the solve wrote them
        if (l > 0) {
            if (plan) GCA_HIP(hipEventRecord(p->ev_built[l], lc->stream));
            GCA_HIP(hipEventRecord(p->ev_join[l], lc->stream));
            GCA_HIP(hipStreamWaitEvent(ctx->stream, p->ev_join[l], 0));
        } else if (!plan) {
            GCA_HIP(hipEventRecord(p->ev_join[0], lc->stream));
            GCA_HIP(hipStreamWaitEvent(ctx->stream, p->ev_join[0], 0));
        }
        return GCA_OK;
    };
    // the sequential registration chain runs on level 0's stream (after level 0's build): level
    // l once it is built, right after level l - 1's launch, which hands it the pose
    auto reg_level = [&](int l) -> int32_t {
        gca_ctx ic = p->lctx[0];
        // the plan reads the target's grid through what it saw at its creation
        if (!icp_plan_target_current(plan, l, p->lvc[prev_slot(p, slot)][l])) return GCA_ERR_DEVICE;
        if (l > 0) GCA_HIP(hipStreamWaitEvent(ic->stream, p->ev_built[l], 0));
        return icp_plan_persist(ic, plan, l);
    };
    // coarsest first, each level's registration enqueued right after its build: a graph launch
    // submits its nodes in capture order, and a late submission is a late start (the finest
    // level's build first, the chain after the coarsest: 0.328 -> 0.355 ms per 640x480 frame)
    for (int l = 0; l < p->nl; ++l) {
        GCA_TRY(build(l));
        if (plan) GCA_TRY(reg_level(l));
    }
    if (plan) {
        // the registration results and the step's lattice error word, by one kernel right after
        // the last registration launch on its stream (every level's build is ordered before that
        // launch, so the error word is final)
        GCA_TRY(icp_plan_copy_results(p->lctx[0], plan, err_dev_of(p, slot), err_host_of(p, slot)));
        GCA_HIP(hipEventRecord(p->ev_join[0], p->lctx[0]->stream));
        GCA_HIP(hipStreamWaitEvent(ctx->stream, p->ev_join[0], 0));
    } else if (!p->overlap) {
        GCA_HIP(hipMemcpyAsync(err_host_of(p, slot), err_dev_of(p, slot), sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    if (overlapped) GCA_HIP(hipStreamWaitEvent(ctx->stream, p->ev_rjoin, 0));
    return GCA_OK;
}

// What work() enqueues on origin's stream (and on streams it forks from there), captured and
// instantiated.
template <class F>
int32_t capture_graph(gca_pipeline p, gca_ctx origin, F work, hipGraph_t *g_out, hipGraphExec_t *ex_out)
{
    hipGraph_t g = nullptr;
    each_ctx(p, [](gca_ctx c) { c->capturing = true; });
    hipError_t e = hipStreamBeginCapture(origin->stream, hipStreamCaptureModeThreadLocal);
    int32_t st = e == hipSuccess ? work() : hip_status(e);
    if (e == hipSuccess) {
        const hipError_t e2 = hipStreamEndCapture(origin->stream, &g);
        if (st == GCA_OK && e2 != hipSuccess) st = hip_status(e2);
    }
    each_ctx(p, [](gca_ctx c) { c->capturing = false; });
    if (st == GCA_OK && !g) st = GCA_ERR_DEVICE;
    hipGraphExec_t ex = nullptr;
    if (st == GCA_OK) {
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (e != hipSuccess) st = hip_status(e);
    }
    if (st != GCA_OK) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    *g_out = g;
    *ex_out = ex;
    return GCA_OK;
}

// The step of slot `slot` as a graph (two lanes: its build part; the registration part is
// capture_registration's).
int32_t capture(gca_pipeline p, int slot, int reg, const uint16_t *depth, const uint8_t *bgr)
{
    const int j = lane_of(p, slot);
    gca_ctx ctx = lane(p, j).main;
    if (!p->front[slot] && !(p->front[slot] = front_launch_new())) return GCA_ERR_OUT_OF_MEMORY;
    ctx->front_rec = p->front[slot];
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    const int32_t st = capture_graph(
        p, ctx,
        [&]() { return p->lanes == 2 ? lane_build(p, slot, j, depth, bgr) : enqueue(p, slot, reg, depth, bgr); }, &g,
        &ex);
    ctx->front_rec = nullptr;
    GCA_TRY(st);
    if (p->exec[slot]) {  // a replacement: an earlier launch of the old graph may still run
        const int32_t sst = sync_all(p);
        if (sst != GCA_OK) {
            hipGraphExecDestroy(ex);
            hipGraphDestroy(g);
            return sst;
        }
        hipGraphExecDestroy(p->exec[slot]);
    }
    if (p->graph[slot]) hipGraphDestroy(p->graph[slot]);
    p->exec[slot] = ex;
    p->graph[slot] = g;
    // the pixel kernel's node: a later step with other frame buffers updates this node instead
    // of capturing again (nullptr: such a step captures)
    p->front_node[slot] = front_launch_node(p->front[slot], g);
    p->cap_depth[slot] = depth;
    p->cap_bgr[slot] = bgr;
    ++p->captures;
    return GCA_OK;
}

// Two lanes: the registration of slot reg's frame (run by lane j) as a graph.
int32_t capture_registration(gca_pipeline p, int reg, int j)
{
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    gca_ctx rc = lane(p, j).reg;
    GCA_TRY(capture_graph(p, rc, [&]() { return registration_launches(p, reg, rc); }, &g, &ex));
    if (p->exec_reg[reg]) {  // a replacement: an earlier launch of the old graph may still run
        const int32_t sst = sync_all(p);
        if (sst != GCA_OK) {
            hipGraphExecDestroy(ex);
            hipGraphDestroy(g);
            return sst;
        }
        hipGraphExecDestroy(p->exec_reg[reg]);
    }
    if (p->graph_reg[reg]) hipGraphDestroy(p->graph_reg[reg]);
    p->exec_reg[reg] = ex;
    p->graph_reg[reg] = g;
    ++p->captures;
    return GCA_OK;
}

// Axis-aligned box holding every back-projected point: x = ((u - 0.5 - cx) z) / fx over the
// image corners and z in [tmin, tmax] (cuda_point_cloud_factory.cu:232), padded slightly.
void frustum_box(const gca_camera &c, float tmin, float tmax, float mn[3], float mx[3])
{
    const gca_intrinsics &k = c.depth_intrin;
    double lo[3] = {DBL_MAX, DBL_MAX, (double)tmin}, hi[3] = {-DBL_MAX, -DBL_MAX, (double)tmax};
    const double us[2] = {0.0, (double)c.width - 1.0}, vs[2] = {0.0, (double)c.height - 1.0};
    const double zs[2] = {(double)tmin, (double)tmax};
    for (double z : zs) {
        for (double u : us) {
            const double x = ((u - 0.5 - k.cx) * z) / k.fx;
            lo[0] = std::min(lo[0], x);
            hi[0] = std::max(hi[0], x);
        }
        for (double v : vs) {
            const double y = ((v - 0.5 - k.cy) * z) / k.fy;
            lo[1] = std::min(lo[1], y);
            hi[1] = std::max(hi[1], y);
        }
    }
    for (int a = 0; a < 3; ++a) {
        const double pad = 1e-3 * (hi[a] - lo[a]) + 1e-3;
        mn[a] = (float)(lo[a] - pad);
        mx[a] = (float)(hi[a] + pad);
    }
}

// Per level: one grid geometry over the frustum box (cell = the normals / correspondence radius,
// the largest of the level, or the tuned multiple of it), and the voxel lattice bound of the box.
void set_geometry(gca_pipeline p)
{
    float bmn[3], bmx[3];
    frustum_box(p->cam, p->tmin, p->tmax, bmn, bmx);
    for (int l = 0; l < p->nl; ++l) {
        const gca_level_params &L = p->lv[l];
        const float r = std::max(std::max(L.normal_radius, L.icp.max_correspondence_distance), L.gradient_radius);
        grid_params(bmn, bmx, r, &p->gdesc[l], p->ctx->tune[GCA_TUNE_GRID_CELL]);
        double bits = 1.0;
        for (int a = 0; a < 3; ++a) bits *= std::floor((double)(bmx[a] - bmn[a]) / (double)L.voxel_size) + 2.0;
        p->max_bits[l] = bits;
    }
}

// Moves c's work to a stream on a hardware queue of its own.  The runtime spreads ordinary
// streams over a few shared hardware queues (4 per process here) in an order of its own; the
// four streams of two-lane steps (two pipeline streams, two registration streams) must sit on
// four distinct queues, or parts meant to run side by side are serialised by the queue they
// share (measured: 0.215 ms per 640x480 frame for the first pipeline of a process, about 0.26 for
// later ones, by the queues they happened to get).  A stream made with a CU mask gets a queue of
// its own; the mask here is every CU of the device.  A runtime that refuses such a stream leaves
// c on its shared one (slower, not wrong).
int32_t own_queue(gca_ctx c)
{
    hipDeviceProp_t prop;
    GCA_HIP(hipGetDeviceProperties(&prop, c->device));
    const int n = std::max(prop.multiProcessorCount, 1);
    std::vector<uint32_t> mask((size_t)(n + 31) / 32, 0xffffffffu);
    if (n % 32) mask.back() = (1u << (n % 32)) - 1u;
    hipStream_t s = nullptr;
    if (hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()) != hipSuccess || !s) {
        (void)hipGetLastError();
        return GCA_OK;
    }
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        hipStreamDestroy(s);
        return hip_status(e);
    }
    hipStreamDestroy(c->stream);
    c->stream = s;
    return GCA_OK;
}

// Lane 1's ctxs and events (two-lane overlapped steps), made on first use, and the four streams
// of two-lane steps on queues of their own.
int32_t make_lane1(gca_pipeline p)
{
    const int dev = p->ctx->device;
    auto make = [&](gca_ctx *c) -> int32_t {
        if (!*c) {
            GCA_TRY(gca_ctx_create(dev, c));
            (*c)->prof_owner = p->ctx;
        }
        return GCA_OK;
    };
    if (p->lane1_ready) return GCA_OK;
    GCA_TRY(make(&p->ctx1));
    GCA_TRY(make(&p->rctx1));
    for (int l = 0; l < p->nl; ++l) GCA_TRY(make(&p->lctx1[l]));
    for (gca_ctx c : {p->ctx, p->rctx, p->ctx1, p->rctx1}) GCA_TRY(own_queue(c));
    p->lane1_ready = true;
    return GCA_OK;
}

// Plans, graphs and the geometry follow a tuning change; every ctx launches with the pipeline
// ctx's shapes.
int32_t follow_tuning(gca_pipeline p)
{
    if (p->ctx->tune_version != p->tune_version) {  // plans and graphs bake the launch shapes in
        GCA_TRY(sync_all(p));
        drop_plans(p);
        set_geometry(p);  // the grid cell may be a tuned one
        p->tune_version = p->ctx->tune_version;
    }
    each_ctx(p, [&](gca_ctx c) { memcpy(c->tune, p->ctx->tune, sizeof(p->ctx->tune)); });
    return GCA_OK;
}

}  // namespace

extern "C" {

int32_t gca_pipeline_destroy(gca_pipeline p)
{
    if (!p) return GCA_ERR_INVALID_VALUE;
    sync_all(p);
    drop_plans(p);
    for (int s = 0; s < kPipeSlots; ++s) {
        front_launch_free(p->front[s]);
        if (p->full[s]) gca_cloud_destroy(p->full[s]);
        for (int l = 0; l < p->nl; ++l)
            if (p->lvc[s][l]) gca_cloud_destroy(p->lvc[s][l]);
    }
    hipFree(p->err_dev);
    hipHostFree(p->err_host);
    if (p->ev_fork) hipEventDestroy(p->ev_fork);
    if (p->ev_rfork) hipEventDestroy(p->ev_rfork);
    if (p->ev_rjoin) hipEventDestroy(p->ev_rjoin);
    for (int s = 0; s < kPipeSlots; ++s) {
        if (p->ev_slot_built[s]) hipEventDestroy(p->ev_slot_built[s]);
        if (p->ev_slot_reg[s]) hipEventDestroy(p->ev_slot_reg[s]);
    }
    if (p->rctx) gca_ctx_destroy(p->rctx);
    if (p->rctx1) gca_ctx_destroy(p->rctx1);
    if (p->ctx1) gca_ctx_destroy(p->ctx1);
    for (int l = 0; l < kPipeMaxLevels; ++l) {
        if (p->ev_join[l]) hipEventDestroy(p->ev_join[l]);
        if (p->ev_built[l]) hipEventDestroy(p->ev_built[l]);
        if (p->lctx[l]) gca_ctx_destroy(p->lctx[l]);
        if (p->lctx1[l]) gca_ctx_destroy(p->lctx1[l]);
    }
    if (p->ctx) gca_ctx_destroy(p->ctx);
    delete p;
    return GCA_OK;
}

int32_t gca_pipeline_create(gca_ctx ctx, const gca_camera *cam, const gca_bilateral *filter, float tmin, float tmax,
                            int32_t n_levels, const gca_level_params *levels, gca_pipeline *out)
{
    if (!out) return GCA_ERR_INVALID_VALUE;
    *out = nullptr;
    if (!ctx || !cam || !levels || n_levels < 1 || n_levels > kPipeMaxLevels) return GCA_ERR_INVALID_VALUE;
    // cuda_point_cloud_factory.cu:286-290
    if (cam->width <= 0 || cam->height <= 0 || !(cam->depth_scale > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if ((int64_t)cam->width * cam->height > (int64_t)INT_MAX / 4) return GCA_ERR_INVALID_VALUE;
    if (!(cam->depth_intrin.fx > 0.0f) || !(cam->depth_intrin.fy > 0.0f)) return GCA_ERR_INVALID_VALUE;
    if (!(tmax > tmin) || !(tmin >= 0.0f)) return GCA_ERR_INVALID_VALUE;
    const gca_bilateral none{GCA_BILATERAL_NONE, 0, 1.0f, 1.0f};
    const gca_bilateral &f = filter ? *filter : none;
    if (f.mode < 0 || f.mode > 2) return GCA_ERR_INVALID_VALUE;
    if (f.mode == GCA_BILATERAL_FIXED &&
        (f.radius < 0 || f.radius > 15 || !(f.sigma_space > 0.0f) || !(f.sigma_depth > 0.0f)))
        return GCA_ERR_INVALID_VALUE;
    for (int l = 0; l < n_levels; ++l) {
        const gca_level_params &L = levels[l];
        if (!(L.voxel_size > 0.0f) || !(L.normal_radius > 0.0f) || !(L.gradient_radius > 0.0f) ||
            L.normal_max_nn < 0 || L.gradient_max_nn < 0 || L.icp.max_iterations < 0 ||
            !(L.icp.max_correspondence_distance > 0.0f) || !(L.icp.lambda_geometric >= 0.0f) ||
            L.icp.lambda_geometric > 1.0f || L.icp.accumulate_mode != levels[0].icp.accumulate_mode)
            return GCA_ERR_INVALID_VALUE;
    }
    gca_pipeline p = new (std::nothrow) gca_pipeline_s();
    if (!p) return GCA_ERR_OUT_OF_MEMORY;
    p->cam = *cam;
    p->bil = f;
    p->tmin = tmin;
    p->tmax = tmax;
    p->nl = n_levels;
    memcpy(p->lv, levels, sizeof(gca_level_params) * n_levels);
    int32_t st = gca_ctx_create(ctx->device, &p->ctx);
    for (int l = 0; l < n_levels && st == GCA_OK; ++l) {
        st = gca_ctx_create(ctx->device, &p->lctx[l]);
        if (st == GCA_OK) p->lctx[l]->prof_owner = p->ctx;
        if (st == GCA_OK && (hipEventCreateWithFlags(&p->ev_join[l], hipEventDisableTiming) != hipSuccess ||
                             hipEventCreateWithFlags(&p->ev_built[l], hipEventDisableTiming) != hipSuccess))
            st = GCA_ERR_DEVICE;
    }
    if (st == GCA_OK) st = gca_ctx_create(ctx->device, &p->rctx);
    if (st == GCA_OK) p->rctx->prof_owner = p->ctx;
    if (st == GCA_OK && (hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
                         hipEventCreateWithFlags(&p->ev_rfork, hipEventDisableTiming) != hipSuccess ||
                         hipEventCreateWithFlags(&p->ev_rjoin, hipEventDisableTiming) != hipSuccess))
        st = GCA_ERR_DEVICE;
    for (int s = 0; s < kPipeSlots && st == GCA_OK; ++s)
        if (hipEventCreateWithFlags(&p->ev_slot_built[s], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&p->ev_slot_reg[s], hipEventDisableTiming) != hipSuccess)
            st = GCA_ERR_DEVICE;
    const int64_t cap = (int64_t)cam->width * cam->height;
    for (int s = 0; s < kPipeSlots && st == GCA_OK; ++s) {
        st = gca_cloud_create(p->ctx, cap, &p->full[s]);
        for (int l = 0; l < n_levels && st == GCA_OK; ++l) st = gca_cloud_create(p->ctx, cap, &p->lvc[s][l]);
    }
    if (st == GCA_OK && hipMalloc(&p->err_dev, 256) != hipSuccess) st = GCA_ERR_OUT_OF_MEMORY;
    if (st == GCA_OK && hipHostMalloc(&p->err_host, 256) != hipSuccess) st = GCA_ERR_OUT_OF_MEMORY;
    if (st == GCA_OK) memset(p->err_host, 0, 256);
    if (st != GCA_OK) {
        gca_pipeline_destroy(p);
        return st;
    }
    set_geometry(p);
    p->tune_version = p->ctx->tune_version;
    *out = p;
    return GCA_OK;
}

int32_t gca_pipeline_set_graph(gca_pipeline p, int32_t enable)
{
    if (!p) return GCA_ERR_INVALID_VALUE;
    p->use_graph = enable != 0;
    if (!p->use_graph) {
        GCA_TRY(sync_all(p));  // no executable graph is destroyed while a launch of it may run
        drop_graphs(p);
    }
    return GCA_OK;
}

int32_t gca_pipeline_ctx(gca_pipeline p, gca_ctx *ctx)
{
    if (!p || !ctx) return GCA_ERR_INVALID_VALUE;
    *ctx = p->ctx;
    return GCA_OK;
}

int32_t gca_pipeline_step(gca_pipeline p, const uint16_t *depth_dev, const uint8_t *bgr_dev)
{
    if (!p || !depth_dev || !bgr_dev) return GCA_ERR_INVALID_VALUE;
    const int slot = (int)(p->frames % p->nslots);
    GCA_TRY(follow_tuning(p));
    // the registration this step runs: sequential, the new frame's (from its second frame on);
    // overlapped, the previous frame's (unless a flush ran it)
    const int reg = p->overlap ? p->pending : (p->frames >= 1 ? slot : -1);
    // a graph is captured only once every buffer it will point into exists: after an eager
    // step of each slot with its plan in place
    bool plans = true;
    for (int s = 0; s < p->nslots; ++s) plans = plans && (p->plan[s] || s == reg);
    const bool graph = p->use_graph && plans && p->eager_since_reset >= p->nslots &&
                       (p->overlap ? reg == prev_slot(p, slot) : reg == slot);
    // Two lanes: buffers are (re)allocated on the host only by eager steps and captures, and
    // one lane's allocation must not free what the other lane's queued work still reads, so
    // those steps start from an idle pipeline (a replayed step allocates nothing).
    const bool replay = graph && p->exec[slot] && (p->lanes == 1 || p->exec_reg[reg]);
    if (p->lanes == 2 && !replay) GCA_TRY(sync_all(p));
    if (reg >= 0 && !p->plan[reg]) GCA_TRY(make_plan(p, reg));
    const int j = lane_of(p, slot);
    if (graph) {
        if (p->lanes == 2 && !p->exec_reg[reg]) GCA_TRY(capture_registration(p, reg, j));
        if (!p->exec[slot]) {
            GCA_TRY(capture(p, slot, reg, depth_dev, bgr_dev));
        } else if (p->cap_depth[slot] != depth_dev || p->cap_bgr[slot] != bgr_dev) {
            // new frame buffers: point the pixel kernel's node at them (capture again only if
            // the update is refused)
            if (p->front_node[slot] &&
                front_launch_update(p->front[slot], p->exec[slot], p->front_node[slot], depth_dev, bgr_dev) ==
                    hipSuccess) {
                p->cap_depth[slot] = depth_dev;
                p->cap_bgr[slot] = bgr_dev;
                ++p->updates;
            } else {
                (void)hipGetLastError();
                GCA_TRY(capture(p, slot, reg, depth_dev, bgr_dev));
            }
        }
        if (p->lanes == 2) {
            GCA_TRY(before_registration(p, reg, j));
            GCA_HIP(hipGraphLaunch(p->exec_reg[reg], lane(p, j).reg->stream));
            GCA_TRY(after_registration(p, reg, j));
            GCA_TRY(before_build(p, slot, j));
            GCA_HIP(hipGraphLaunch(p->exec[slot], lane(p, j).main->stream));
            GCA_TRY(after_build(p, slot, j));
        } else {
            GCA_HIP(hipGraphLaunch(p->exec[slot], p->ctx->stream));
        }
    } else {
        GCA_TRY(enqueue(p, slot, reg, depth_dev, bgr_dev));
        if (reg >= 0) ++p->eager_since_reset;
    }
    // host bookkeeping of a replayed step: the slot's clouds changed on the device
    p->full[slot]->host_valid = false;
    for (int l = 0; l < p->nl; ++l) p->lvc[slot][l]->host_valid = false;
    if (p->overlap) {
        if (reg >= 0) p->last = reg;
        p->pending = p->frames >= 1 ? slot : -1;  // the first frame has nothing to register onto
    } else {
        p->last = slot;
    }
    ++p->frames;
    return GCA_OK;
}

int32_t gca_pipeline_flush(gca_pipeline p)
{
    if (!p) return GCA_ERR_INVALID_VALUE;
    if (!p->overlap || p->pending < 0) return GCA_OK;
    const int reg = p->pending;
    GCA_TRY(follow_tuning(p));
    if (!p->plan[reg]) {
        if (p->lanes == 2) GCA_TRY(sync_all(p));  // see gca_pipeline_step
        GCA_TRY(make_plan(p, reg));
    }
    if (p->lanes == 2) {
        // on the lane the next step would take (the one that would register this frame)
        const int j = (reg + 1) & 1;
        GCA_TRY(before_registration(p, reg, j));
        GCA_TRY(registration_launches(p, reg, lane(p, j).reg));
        GCA_TRY(after_registration(p, reg, j));
    } else {
        GCA_TRY(enqueue_registration(p, reg));
        GCA_HIP(hipStreamWaitEvent(p->ctx->stream, p->ev_rjoin, 0));
    }
    p->last = reg;
    p->pending = -1;
    return GCA_OK;
}

int32_t gca_pipeline_set_overlap(gca_pipeline p, int32_t enable)
{
    if (!p) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(sync_all(p));
    drop_plans(p);
    p->overlap = enable != 0;
    // overlapped steps run in two lanes when the levels can be built in lockstep
    p->lanes = p->overlap && lockstep_ok(p) ? 2 : 1;
    if (p->lanes == 2) {
        const int32_t st = make_lane1(p);
        if (st != GCA_OK) {  // stay a sequential pipeline
            p->overlap = false;
            p->lanes = 1;
            p->nslots = 2;
            p->frames = 0;
            p->pending = p->last = -1;
            return st;
        }
    }
    // synchronising the pipeline ctx waits for both lanes (registrations are joined by events)
    p->ctx->sync_also[0] = p->lanes == 2 ? p->rctx : nullptr;
    p->ctx->sync_also[1] = p->lanes == 2 ? p->ctx1 : nullptr;
    p->ctx->sync_also[2] = p->lanes == 2 ? p->rctx1 : nullptr;
    p->nslots = p->lanes == 2 ? 4 : (p->overlap ? 3 : 2);
    for (int s = 0; s < kPipeSlots; ++s) p->slot_built_recorded[s] = p->slot_reg_recorded[s] = false;
    p->frames = 0;
    p->pending = -1;
    p->last = -1;
    return GCA_OK;
}

int32_t gca_pipeline_result(gca_pipeline p, double T[16], gca_icp_report *reports, int32_t *valid)
{
    if (!p || !T || !valid) return GCA_ERR_INVALID_VALUE;
    *valid = 0;
    GCA_TRY(sync_all(p));
    int32_t dev_st = GCA_OK;
    each_ctx(p, [&](gca_ctx c) {
        const int32_t e = ctx_take_device_error(c);
        if (dev_st == GCA_OK) dev_st = e;
    });
    GCA_TRY(dev_st);
    if (p->last >= 0 && *err_host_of(p, p->last)) return GCA_ERR_INVALID_VALUE;  // a voxel lattice was refused
    const bool registered = p->overlap ? p->last >= 0 : p->frames >= 2;
    if (!registered) {
        static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        memcpy(T, I, sizeof(I));
        return GCA_OK;
    }
    GCA_TRY(icp_plan_results(p->ctx, p->plan[p->last], T, reports));
    *valid = 1;
    return GCA_OK;
}

int32_t gca_pipeline_slots(gca_pipeline p, int32_t *slots)
{
    if (!p || !slots) return GCA_ERR_INVALID_VALUE;
    *slots = p->nslots;
    return GCA_OK;
}

int32_t gca_pipeline_graph_stats(gca_pipeline p, int64_t *captures, int64_t *node_updates)
{
    if (!p || !captures || !node_updates) return GCA_ERR_INVALID_VALUE;
    *captures = p->captures;
    *node_updates = p->updates;
    return GCA_OK;
}

int32_t gca_pipeline_sizes(gca_pipeline p, int32_t slot, int64_t *full_n, int64_t *level_n)
{
    if (!p || slot < 0 || slot >= p->nslots || !full_n) return GCA_ERR_INVALID_VALUE;
    GCA_TRY(sync_all(p));  // the slot may be another lane's
    GCA_TRY(cloud_sync_header(p->full[slot]));
    *full_n = p->full[slot]->host_hdr.n;
    if (level_n)
        for (int l = 0; l < p->nl; ++l) {
            GCA_TRY(cloud_sync_header(p->lvc[slot][l]));
            level_n[l] = p->lvc[slot][l]->host_hdr.n;
        }
    return GCA_OK;
}

}  // extern "C"
