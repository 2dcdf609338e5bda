/*
 * gca.h -- C ABI of the MI355X (gfx950) RGB-D front end + colored-ICP library (libgca.so).
 *
 * Every entry point replaces one function of the reference's drop-in layer: the
 * `::cudaError_t gca::cuda_*(thrust::device_vector&...)` free functions and the
 * `gca::point_cloud` / `gca::color_icp` methods that call them.  The reference citation
 * (path:line inside XssZuiNB/CUDA_3D) is given above each declaration.
 *
 * Conventions
 *  - Plain C types only: handles, raw pointers, sizes.  No exceptions cross the ABI.
 *  - Return value: a gca_status code (0 = ok).  GCA_ERR_INVALID_VALUE plays the role of the
 *    reference's cudaErrorInvalidValue.  gca_status_string() names each code.
 *  - A gca_ctx owns one HIP stream (default priority) on one device.  Every call enqueues on
 *    that stream.  Calls that return host scalars (sizes, reports, sums) synchronise it.
 *  - A gca_cloud is a library-owned structure-of-arrays on the device:
 *      xyzi     float4 per point: x, y, z (metres), intensity
 *      rgb      float4 per point: r, g, b in [0,1], w = 0
 *      normals  float4 per point: nx, ny, nz, 0            (after gca_estimate_normals)
 *      gradient float4 per point: dx, dy, dz, 0            (after gca_compute_color_gradient)
 *    The point count lives in device memory so that pipelines never wait on the host;
 *    gca_cloud_size() copies it back.
 *  - Device pointers passed in (depth/colour frames, index buffers) are borrowed: the caller
 *    keeps ownership and must keep them alive until the ctx stream has passed the call.
 *  - Poses are double[16], row-major 4x4 homogeneous matrices.
 *  - One ctx per host thread.  Calls on one ctx are not thread-safe.
 */
#ifndef GCA_GCA_H
#define GCA_GCA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCA_ABI_VERSION 1

typedef enum gca_status {
    GCA_OK = 0,
    GCA_ERR_INVALID_VALUE = 1, /* bad argument or precondition (reference: cudaErrorInvalidValue) */
    GCA_ERR_OUT_OF_MEMORY = 2,
    GCA_ERR_DEVICE = 3,        /* HIP runtime or kernel error */
    GCA_ERR_NOT_SPD = 4,       /* 6x6 system not positive definite: identity step taken */
    GCA_ERR_NOT_READY = 5      /* missing prerequisite (e.g. target without normals) */
} gca_status;

typedef struct gca_ctx_s *gca_ctx;
typedef struct gca_cloud_s *gca_cloud;

/* Pinhole intrinsics.  Reference: src/camera/camera_param.hpp:5-11 (gca::intrinsics). */
typedef struct gca_intrinsics {
    float fx, fy, cx, cy;
} gca_intrinsics;

/* Rigid depth->colour transform; rotation is COLUMN-major as librealsense delivers it
 * (read as R[0]x + R[3]y + R[6]z).  Reference: src/camera/camera_param.hpp:13-18,
 * src/geometry/cuda_point_cloud_factory.cu:160-170. */
typedef struct gca_extrinsics {
    float rotation[9];
    float translation[3];
} gca_extrinsics;

/* Camera description.  Reference: gca::cuda_camera_param
 * (src/cuda_container/cuda_container.hpp:65-113; ctor cuda_camera_param_container.cu:35-63).
 * The colour frame has the same width/height as the depth frame (reference uses `width`
 * for both, cuda_point_cloud_factory.cu:244,256). */
typedef struct gca_camera {
    int32_t width;
    int32_t height;
    float depth_scale; /* metres per raw depth unit (0.001 for Z16 millimetres) */
    gca_intrinsics depth_intrin;
    gca_intrinsics color_intrin;
    gca_extrinsics depth_to_color;
} gca_camera;

enum {
    GCA_BILATERAL_NONE = 0,              /* raw depth, reference template arg <false> */
    GCA_BILATERAL_FIXED = 1,             /* fixed (radius, sigma_space px, sigma_depth raw units) */
    GCA_BILATERAL_REFERENCE_ADAPTIVE = 2 /* reference's depth-adaptive schedule,
                                            cuda_point_cloud_factory.cu:100-129 */
};

/* Bilateral depth filter.  Weight of a neighbour = exp(-(dx^2+dy^2)/sigma_space^2) *
 * exp(-(d0-dn)^2/sigma_depth^2) (no 1/2 factor, as cuda_point_cloud_factory.cu:19-27);
 * zero-depth neighbours are skipped (:86-87).  sigma_depth is in raw depth units. */
typedef struct gca_bilateral {
    int32_t mode;
    int32_t radius;
    float sigma_space;
    float sigma_depth;
} gca_bilateral;

enum {
    GCA_ACCUM_COMPENSATED = 0, /* FP32 per-point terms, TwoSum-compensated FP32 partials, FP64 merge */
    GCA_ACCUM_NAIVE_FP32 = 1,  /* ablation: every sum FP32 (within and across work units), like a
                                  plain FP32 transform_reduce */
    GCA_ACCUM_FP64 = 2         /* FP64 partials throughout */
};

/* Colored-ICP parameters.  Reference: gca::color_icp ctor and constants
 * (src/registration/color_icp.hpp:20-21,46-48). */
typedef struct gca_icp_params {
    int32_t max_iterations;            /* reference: maximum_iterations */
    float max_correspondence_distance; /* reference: max_correspondence_distance */
    float lambda_geometric;            /* reference: m_color_icp_lambda = 0.968 */
    int32_t accumulate_mode;           /* GCA_ACCUM_* */
    double conv_rotation;              /* stop when cos(theta_step) >= 1 - conv_rotation ... */
    double conv_translation_sq;        /* ... and |t_step|^2 <= conv_translation_sq (reference 1e-6, 1e-6);
                                          0, 0 runs all max_iterations unless the step is exactly 0 */
} gca_icp_params;

typedef struct gca_icp_report {
    int32_t iterations;      /* Gauss-Newton iterations run */
    int32_t converged;       /* 1 if the stopping rule fired */
    int32_t solver_failures; /* iterations whose 6x6 LDL^T failed (identity step taken) */
    int32_t n_correspondences; /* correspondences in the last iteration */
    double rmse;             /* reference "RMSE": sum r^2 / N_src of the last build (color_icp.cu:231) */
    double fitness;          /* n_correspondences / N_src */
    double latency_us;       /* device clock from the call's set-up until this registration
                                finished (per-pair latency inside a batched launch) */
    double active_us;        /* device clock from the start of this registration's first
                                iteration until it finished (active_us / iterations: the time
                                of one Gauss-Newton iteration) */
} gca_icp_report;

/* Per-level schedule of the multi-scale driver (coarse to fine). */
typedef struct gca_level_params {
    float voxel_size;
    float normal_radius;
    int32_t normal_max_nn;   /* 0 = no cap (reference) */
    float gradient_radius;
    int32_t gradient_max_nn; /* 0 = no cap (reference) */
    gca_icp_params icp;
} gca_level_params;

/* ------------------------------------------------------------------------------------------ */
/* Context, status                                                                           */

const char *gca_status_string(int32_t status);
int32_t gca_abi_version(void);

/* Reference: cuda_warm_up_gpu (src/util/gpu_check.cu:14-27) + the implicit CUDA context. */
int32_t gca_ctx_create(int32_t device, gca_ctx *out);
int32_t gca_ctx_destroy(gca_ctx ctx);
int32_t gca_ctx_synchronize(gca_ctx ctx);
/* Borrow the ctx's hipStream_t (e.g. to order caller work or time it). */
int32_t gca_ctx_get_stream(gca_ctx ctx, void **stream);

/* Stage timing with hipEvents on the ctx stream (no reference counterpart: the reference times
 * only with std::chrono in its demos, apps/app.cpp:117,354-357).  While enabled, every launch of
 * a stage is bracketed by an event pair.  Stages: */
enum {
    GCA_STAGE_FRONTEND = 0,   /* bilateral + back-projection + compaction (3 kernels) */
    GCA_STAGE_VOXEL = 1,      /* mark + popcount scan + accumulate + finalize */
    GCA_STAGE_GRID = 2,       /* count + scan + scatter */
    GCA_STAGE_NORMALS = 3,    /* moments + capped overflow + per-point eigen solve */
    GCA_STAGE_GRADIENT = 4,   /* moments + per-point 3x3 solve */
    GCA_STAGE_ICP = 5,        /* ICP solve: one persistent launch per call (or one per iteration) */
    GCA_STAGE_COUNT = 6
};
int32_t gca_ctx_profile(gca_ctx ctx, int32_t enable);  /* enable != 0 also clears the records */

/* Launch-shape tuning of a ctx (no reference counterpart: the reference takes whatever shapes
 * Thrust picks).  Values are chosen by the autotuner (cuda_3d_amd/autotune.py); 0 = built-in
 * heuristic.  Results do not depend on the tuning beyond the last bits of FP64 reductions. */
enum {
    GCA_TUNE_ICP_LANES_SMALL = 0,  /* lanes per correspondence query in one-pair calls, clouds < 16k
                                      points: 1,2,4,8,16,32,64 */
    GCA_TUNE_ICP_LANES_MEDIUM = 1, /* ... clouds < 48k points */
    GCA_TUNE_ICP_LANES_LARGE = 2,  /* ... larger clouds */
    GCA_TUNE_ICP_MAX_UNITS = 3,    /* cap on work units (FP64 partials) per pair and iteration, 1..1024 */
    GCA_TUNE_NORMALS_LANES = 4,    /* lanes per normal/gradient query: 1,2,4,8,16,32,64 */
    GCA_TUNE_ICP_PERSIST_BLOCKS = 5, /* workgroups of the persistent ICP launch (0 = 2 per CU), 1..65536 */
    GCA_TUNE_ICP_LAUNCH = 6,       /* 0 = resident launches for one-pair calls, a persistent launch for
                                      several pairs; 1 = one persistent launch per call; 2 = one launch
                                      per iteration; 3 = one resident launch per pair and level.  Every
                                      strategy gives the same bits for the same unit partition */
    GCA_TUNE_ICP_LANES_BATCH = 7,  /* lanes per correspondence query in calls with several pairs (0 = 1) */
    GCA_TUNE_FRONTEND = 8,         /* 0/2 = pixels, then compaction (two kernels); 1 = one pass (pixels,
                                      look-back offset and compaction in one kernel); same bits */
    GCA_TUNE_GRID_CELL = 9,        /* neighbour-grid cell edge in quarters of the search radius, 4..12
                                      (0 = 4: cell = radius); any edge >= the radius gives the same
                                      neighbours.  Exact grids (gca_grid_cells) keep cell = radius */
    GCA_TUNE_VOXEL = 10,           /* gca_voxel_downsample binning: 0 = occupancy bitmap over the voxel
                                      lattice up to 2^34 cells, radix sort of the packed keys beyond;
                                      1 = always the bitmap (same fallback); 2 = always the sort.  Same
                                      bits either way */
    GCA_TUNE_COUNT = 11
};
int32_t gca_ctx_set_tuning(gca_ctx ctx, int32_t key, int32_t value);
int32_t gca_ctx_get_tuning(gca_ctx ctx, int32_t key, int32_t *value);
/* Synchronises; launches[s] and total_ms[s] (arrays of GCA_STAGE_COUNT) since the last enable. */
int32_t gca_ctx_profile_read(gca_ctx ctx, int64_t *launches, double *total_ms);

/* ------------------------------------------------------------------------------------------ */
/* Point clouds                                                                              */

/* Reference: gca::point_cloud(size_t n_points) (src/geometry/point_cloud.cu:23-27). */
int32_t gca_cloud_create(gca_ctx ctx, int64_t capacity, gca_cloud *out);
int32_t gca_cloud_destroy(gca_cloud cloud);
/* Reference: point_cloud::points_number (point_cloud.hpp:47-50).  Synchronises. */
int32_t gca_cloud_size(gca_cloud cloud, int64_t *n);
int32_t gca_cloud_capacity(gca_cloud cloud, int64_t *cap);
int32_t gca_cloud_has_normals(gca_cloud cloud, int32_t *flag);
/* Host xyz (n x 3) and rgb (n x 3, [0,1]) -> device cloud; intensity is derived as in
 * color3::to_intensity.  Reference: point_cloud::create_from_pcl (point_cloud.cu:481-504). */
int32_t gca_cloud_upload(gca_cloud cloud, const float *xyz, const float *rgb, int64_t n);
/* Device-resident points: xyzi_dev = n float4 (x, y, z, intensity), rgb_dev = n float4 or NULL
 * (colour zeroed), copied on the ctx stream.  The caller orders its own writes to the sources
 * before the call.  Same semantics as gca_cloud_upload (create_from_pcl, point_cloud.cu:481-504)
 * for points that already live in device memory, e.g. torch tensors. */
int32_t gca_cloud_upload_device(gca_ctx ctx, gca_cloud cloud, const void *xyzi_dev,
                                const void *rgb_dev, int64_t n);
/* Optional host normals (n x 3) for a cloud whose normals come from elsewhere. */
int32_t gca_cloud_upload_normals(gca_cloud cloud, const float *normals, int64_t n);
int32_t gca_cloud_upload_gradient(gca_cloud cloud, const float *gradient, int64_t n);
/* Reference: point_cloud::download / download_normals (point_cloud.cu:75-91).  Any output
 * pointer may be NULL.  xyz/rgb/normals/gradient are n x 3 floats, intensity n floats. */
int32_t gca_cloud_download(gca_cloud cloud, float *xyz, float *rgb, float *intensity,
                           float *normals, float *gradient);
/* Borrowed device pointers to the SoA arrays (float4 each) and the device point count. */
int32_t gca_cloud_device_arrays(gca_cloud cloud, void **xyzi, void **rgb, void **normals,
                                void **gradient, void **count);
/* Reference: point_cloud::get_min_bound/get_max_bound (point_cloud.cu:107-132,
 * cuda_compute_min_max_bound geometry_util.cuh:88-105, including its max-init FLT_MIN). */
int32_t gca_cloud_bounds(gca_cloud cloud, float min_bound[3], float max_bound[3]);
/* Reference: point_cloud copy ctor / operator= (point_cloud.cu:29-49). */
int32_t gca_cloud_copy(gca_ctx ctx, gca_cloud src, gca_cloud dst);
/* dst += src: src's points (xyz, intensity, colour) appended after dst's; normals kept only
 * when both clouds have them (else dst has none); the colour gradient is dropped; the bounds are
 * recomputed over all points.  dst == src appends a copy of the cloud to itself.  Synchronises
 * (sizes are read on the host).  Reference: point_cloud::operator+= (point_cloud.cu:55-73). */
int32_t gca_cloud_append(gca_ctx ctx, gca_cloud dst, gca_cloud src);

/* ------------------------------------------------------------------------------------------ */
/* Front end                                                                                 */

/* Depth (uint16, H*W, device) + colour (BGR8, H*W*3, device) -> point cloud (valid points
 * in raster order).  Reference: cuda_make_point_cloud
 * (src/geometry/cuda_point_cloud_factory.cuh:16-20, .cu:269-313) called by
 * point_cloud::create_from_rgbd (point_cloud.cu:455-479). */
int32_t gca_cloud_from_rgbd(gca_ctx ctx, const uint16_t *depth_dev, const uint8_t *bgr_dev,
                            const gca_camera *camera, const gca_bilateral *filter,
                            float threshold_min_m, float threshold_max_m, gca_cloud out);

/* Voxel-grid downsample (mean xyz and rgb per occupied voxel, lexicographic voxel order).
 * Reference: cuda_voxel_grid_downsample (src/geometry/cuda_voxel_grid_down_sample.cuh:11-13)
 * via point_cloud::voxel_grid_down_sample (point_cloud.cu:198-229). */
int32_t gca_voxel_downsample(gca_ctx ctx, gca_cloud in, float voxel_size, gca_cloud out);

/* Euclidean clustering: the reference's seeded breadth-first searches over the graph joining
 * points closer than tolerance.  Seeds run in index order; a search stops expanding once its queue
 * holds max_cluster_size points, and the rest of that component is left to later seeds; a queue
 * of at least min_cluster_size points is a cluster.  labels_dev (device, one int32 per point)
 * receives each point's cluster number (seed order) or -1.  Neighbours are visited in the order
 * (squared distance, index); only the pieces of a component above max_cluster_size depend on it.
 * *n_cut (may be NULL) counts the components that were cut.  GCA_ERR_NOT_READY without normals,
 * like the reference's empty result.  Synchronises.  Reference: cuda_euclidean_clustering
 * (src/geometry/cuda_clustering.cuh, .cu:31-110) via point_cloud::euclidean_clustering
 * (point_cloud.hpp:71-73, point_cloud.cu:284-335). */
int32_t gca_euclidean_clustering(gca_ctx ctx, gca_cloud cloud, float tolerance,
                                 int32_t min_cluster_size, int32_t max_cluster_size,
                                 int32_t *labels_dev, int32_t *n_clusters, int32_t *n_cut);

/* gca_euclidean_clustering with the labels copied to host memory (labels_cap >= the cloud's
 * point count); for callers without device buffers (the C++ facade). */
int32_t gca_euclidean_clustering_host(gca_ctx ctx, gca_cloud cloud, float tolerance,
                                      int32_t min_cluster_size, int32_t max_cluster_size,
                                      int32_t *labels_host, int64_t labels_cap, int32_t *n_clusters,
                                      int32_t *n_cut);

/* Residual-based segmentation: gca_euclidean_clustering's searches with an edge followed only when
 * the two points' absolute ICP residuals differ by less than 0.003 (|r_a - r_b| in FP32).
 * abs_residual_dev: one float per point (device), e.g. gca_compute_residuals(GCA_RESIDUAL_ABS)
 * of this cloud as the ICP source (color_icp::get_abs_residual, color_icp.cu:90-137).  No normals
 * needed.  Reference: cuda_residual_based_segmentation (src/geometry/cuda_clustering.cuh,
 * .cu:251-341) via point_cloud::icp_residual_segmentation (point_cloud.hpp:79-81,
 * point_cloud.cu:391-435). */
int32_t gca_residual_segmentation(gca_ctx ctx, gca_cloud cloud, float tolerance,
                                  const float *abs_residual_dev, int32_t min_cluster_size,
                                  int32_t max_cluster_size, int32_t *labels_dev, int32_t *n_clusters,
                                  int32_t *n_cut);

/* gca_residual_segmentation with host residuals in and host labels out (labels_cap >= the cloud's
 * point count), for callers without device buffers (the C++ facade). */
int32_t gca_residual_segmentation_host(gca_ctx ctx, gca_cloud cloud, float tolerance,
                                       const float *abs_residual_host, int32_t min_cluster_size,
                                       int32_t max_cluster_size, int32_t *labels_host,
                                       int64_t labels_cap, int32_t *n_clusters, int32_t *n_cut);

/* Radius outlier removal: keep a point when more than min_neighbors points (itself
 * included) lie within radius.  Reference: cuda_grid_radius_outliers_removal
 * (src/geometry/cuda_nn_search.cuh:37-41, .cu:725-800) via
 * point_cloud::radius_outlier_removal (point_cloud.cu:238-282). */
int32_t gca_radius_outlier_removal(gca_ctx ctx, gca_cloud in, float radius,
                                   int32_t min_neighbors, gca_cloud out);

/* ------------------------------------------------------------------------------------------ */
/* Neighbour search                                                                          */

/* 1-NN of each query point among `reference` within `radius` (index or -1), ties broken by
 * lowest index.  out_dev: int32[n_query] device.  Reference: cuda_nn_search
 * (src/geometry/cuda_nn_search.cuh:11-15, .cu:545-619) via point_cloud::nn_search
 * (point_cloud.cu:506-557). */
int32_t gca_nn_search(gca_ctx ctx, gca_cloud query, gca_cloud reference, float radius,
                      int32_t *out_dev);

/* The same, into host memory (capacity int32 entries); *n_query receives the query count
 * (GCA_ERR_INVALID_VALUE if it exceeds capacity).  Synchronises.  Reference: the host-vector
 * overload point_cloud::nn_search(std::vector<index_t>&, ...) (point_cloud.hpp:112-114,
 * point_cloud.cu:540-557). */
int32_t gca_nn_search_host(gca_ctx ctx, gca_cloud query, gca_cloud reference, float radius,
                           int32_t *out_host, int64_t capacity, int64_t *n_query);

/* All neighbours within radius (max_nn = 0) or the max_nn nearest, each query's list sorted
 * by (squared distance, index).  offsets_dev: int32[n_query+1] device (CSR), idx_dev:
 * int32[idx_capacity] device.  *total receives the number of indices; if it exceeds
 * idx_capacity nothing is written to idx_dev and GCA_ERR_INVALID_VALUE is returned.
 * Reference: cuda_search_radius_neighbors (src/geometry/cuda_nn_search.cuh:17-24,
 * .cu:622-723). */
int32_t gca_radius_search(gca_ctx ctx, gca_cloud query, gca_cloud reference, float radius,
                          int32_t max_nn, int32_t *offsets_dev, int32_t *idx_dev,
                          int64_t idx_capacity, int64_t *total);

/* Grid cell of every point for a grid of cell size >= radius built over `cloud`
 * (int32[n*3] device, cell coordinates).  Exposes the cell binning for parity checks.
 * Reference: compute_grid_cell_functor (src/geometry/cuda_nn_search.cu:36-59). */
int32_t gca_grid_cells(gca_ctx ctx, gca_cloud cloud, float radius, int32_t *cells_dev,
                       float origin[3], float *cell_size, int32_t dims[3]);

/* ------------------------------------------------------------------------------------------ */
/* Normals and colour gradient                                                               */

/* PCA normals over neighbours within radius (max_nn nearest if max_nn > 0), oriented
 * towards (0,0,-1); fewer than 3 neighbours -> zero normal.  Reference:
 * cuda_estimate_normals (src/geometry/cuda_estimate_normals.cuh:12-15, .cu:153-199) via
 * point_cloud::estimate_normals (point_cloud.cu:145-181). */
int32_t gca_estimate_normals(gca_ctx ctx, gca_cloud cloud, float radius, int32_t max_nn);

/* Intensity gradient d_p on the tangent plane (Park et al. eq. 10).  Reference:
 * cuda_compute_color_gradient (src/registration/cuda_compute_color_gradient.cuh:11-15,
 * .cu:92-127) via color_icp::compute_color_gradient_target (color_icp.cu:50-87). */
int32_t gca_compute_color_gradient(gca_ctx ctx, gca_cloud cloud, float radius, int32_t max_nn);

/* ------------------------------------------------------------------------------------------ */
/* Colored ICP                                                                               */

/* Rigid transform of points (and normals, gradients) in place.  Reference:
 * point_cloud::transform (point_cloud.cu:93-105), cuda_transform_point/normals
 * (geometry_util.cuh:133-177). */
int32_t gca_transform(gca_ctx ctx, gca_cloud cloud, const double T[16]);

/* Gauss-Newton system of colored ICP for source points mapped by T and correspondences
 * nn_dev (int32[n_src] device, -1 = none).  out[0..20] = upper triangle of J^T J (row-major),
 * out[21..26] = J^T r, out[27] = sum r^2, out[28] = number of correspondences.
 * Reference: cuda_build_gauss_newton_color_icp
 * (src/registration/cuda_color_icp_build_least_square.cuh:43-48, .cu:431-474). */
int32_t gca_build_gauss_newton(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn_dev,
                               float lambda_geometric, const double T[16],
                               int32_t accumulate_mode, double out[29]);

/* gca_build_gauss_newton for callers that cannot allocate (no counterpart in the reference; for
 * code under HIP graph capture, tight serving loops and harnesses that forbid hipMalloc): the same
 * kernels and sums, with the result left in device memory (out_dev: 29 doubles, same layout) and
 * no allocation, host read-back or synchronisation (the call only enqueues on the ctx stream).
 * The work is split by the source cloud's capacity (no size read-back), so the sums equal
 * gca_build_gauss_newton's bit for bit when the cloud is full, and to the FP64 summation order
 * otherwise.  Scratch: a static device array of 257 KiB per GPU (partials of up to 1024
 * workgroups and a ticket), which every GPU the library runs on holds until the process ends; the
 * workgroups stride over the source points, so any cloud fits.  Calls that overlap on one GPU
 * share that array and race. */
int32_t gca_build_gauss_newton_noalloc(gca_ctx ctx, gca_cloud src, gca_cloud tgt,
                                       const int32_t *nn_dev, float lambda_geometric,
                                       const double T[16], int32_t accumulate_mode, double *out_dev);

enum {
    GCA_RESIDUAL_PAIR = 0, /* out: float2 (sqrt(l) r_g, sqrt(1-l) r_c) per point, .cu:268-308 */
    GCA_RESIDUAL_ABS = 1,  /* out: |r_g|+|r_c| weighted, .cu:310-350 */
    GCA_RESIDUAL_RMSE = 2  /* out: l r_g^2 + (1-l) r_c^2, .cu:352-391 */
};
/* Per-point residuals (out_dev may be NULL) and their FP64 sum (*sum may be NULL).
 * Reference: cuda_compute_residual_color_icp / _abs_residual_ / _RMSE_
 * (src/registration/cuda_color_icp_build_least_square.cuh:11-41). */
int32_t gca_compute_residuals(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const int32_t *nn_dev,
                              float lambda_geometric, const double T[16], int32_t kind,
                              float *out_dev, double *sum);

/* Per-point absolute residuals of src (as it stands) against tgt: nearest target point within
 * max_corr, |sqrt(l) r_g| + |sqrt(1-l) r_c| per source point (0 without a correspondence) into
 * out_host (capacity floats; *n_points = the source count), and their mean over all source points
 * (*mean may be NULL).  The target's colour gradient is computed with gradient_radius if absent.
 * Synchronises.  Reference: color_icp::get_abs_residual / get_abs_residual_host
 * (src/registration/color_icp.cu:90-150), whose source has been moved to the final pose. */
int32_t gca_abs_residual_host(gca_ctx ctx, gca_cloud src, gca_cloud tgt, float max_corr,
                              float lambda_geometric, float gradient_radius, float *out_host,
                              int64_t capacity, int64_t *n_points, double *mean);

/* Host FP64 LDL^T solve of (J^T J) xi = -J^T r and the Euler (Rz Ry Rx) increment.  Identity
 * and GCA_ERR_NOT_SPD if the factorisation fails.  Reference: solve_JTJ_JTr<true>,
 * transform_vec6d_to_mat4f (src/registration/eigen_solver.hpp:30-85). */
int32_t gca_solve_gauss_newton(const double jtj_upper[21], const double jtr[6], double xi[6],
                               double T_step[16]);

/* Colored ICP of src onto tgt (tgt needs normals; its colour gradient is computed with
 * gradient_radius if absent).  The iteration loop runs on the device: correspondence search,
 * Gauss-Newton build, FP64 LDL^T solve and pose update without host round trips.
 * Reference: color_icp::align (src/registration/color_icp.cu:151-234). */
int32_t gca_color_icp_align(gca_ctx ctx, gca_cloud src, gca_cloud tgt, const gca_icp_params *p,
                            float gradient_radius, const double T0[16], double T_out[16],
                            gca_icp_report *report);

/* Objective E = sum lambda r_g^2 + (1-lambda) r_c^2 (unweighted residuals, reference
 * compute_RMSE_functor .cu:82-132) at pose Rz Ry Rx(xi)*T with frozen correspondences, and its
 * adjoint: dE/dxi (6, FP64) and per source point dE/d(x,y,z,intensity) (float4[n_src] device,
 * may be NULL).  No reference counterpart beyond the objective itself (the reference only
 * evaluates it); exposed for differentiable use. */
int32_t gca_color_icp_objective_grad(gca_ctx ctx, gca_cloud src, gca_cloud tgt,
                                     const int32_t *nn_dev, float lambda_geometric,
                                     const double T[16], const double xi[6], double *E,
                                     double dE_dxi[6], float *dE_dpts_dev);

/* Batched colored ICP: n_pairs independent (src, tgt) pairs registered together, each with
 * its own T0 (n_pairs x 16) and result (n_pairs x 16).  Every iteration of every pair runs in
 * one persistent launch that serves the pairs' work queues.  Reference: color_icp::align called once per pair
 * (src/registration/color_icp.cu:151-234). */
int32_t gca_color_icp_batch(gca_ctx ctx, int32_t n_pairs, const gca_cloud *srcs,
                            const gca_cloud *tgts, const gca_icp_params *p,
                            float gradient_radius, const double *T0, double *T_out,
                            gca_icp_report *reports);

/* Batched coarse-to-fine colored ICP: n_pairs chains of n_levels registrations, all in one
 * persistent launch.  srcs/tgts: n_pairs * n_levels clouds, pair k level l at k * n_levels + l
 * (coarse to fine); params and gradient_radius: one per level (accumulate_mode equal across
 * levels).  Level l + 1 of a pair starts from the pose level l ends with; levels with
 * max_iterations == 0 are skipped.  T0 / T_out: n_pairs x 16; reports: n_pairs * n_levels (may
 * be NULL).  Reference: color_icp::align per pair and level, as sequenced by
 * apps/test_color_icp.cpp:378-414. */
int32_t gca_color_icp_batch_multiscale(gca_ctx ctx, int32_t n_pairs, int32_t n_levels,
                                       const gca_cloud *srcs, const gca_cloud *tgts,
                                       const gca_icp_params *params, const float *gradient_radius,
                                       const double *T0, double *T_out, gca_icp_report *reports);

/* Batch plan: gca_color_icp_batch_multiscale set up once and run many times (same arguments,
 * same poses and reports).  create checks the pairs, readies the targets' neighbour grids and
 * allocates every buffer a run needs.  enqueue queues one run on ctx's stream: state set-up,
 * the persistent launch over every pair and level, the results into pinned host memory.  It
 * makes no host round trip and no allocation, so a caller may capture it into a graph of its
 * own.  With use_graph != 0 the plan captures the run as a graph of its own at the first
 * enqueue and replays it after (a stream the caller is capturing gets the plain launches).
 * A run reads the clouds as they are when it executes: sources rewritten in place are
 * registered afresh.  enqueue returns GCA_ERR_INVALID_VALUE, queueing nothing, once a cloud of
 * the plan was re-allocated, or a target's neighbour grid, normals or gradient was invalidated
 * since create: make a new plan then.  set_initial replaces the T0 table (n_pairs x 16) for
 * later runs; it synchronises ctx's stream.  results synchronises ctx's stream and reads the
 * last run's poses (n_pairs x 16) and reports (n_pairs * n_levels, may be NULL);
 * GCA_ERR_NOT_READY before any run.  Runs on one stream are ordered, and each overwrites the
 * previous one's results.  destroy synchronises and frees (before ctx is destroyed).
 * No reference counterpart: the reference registers pair by pair (color_icp::align,
 * src/registration/color_icp.cu:151-234); the batch is this library's. */
typedef struct gca_icp_batch_plan_s *gca_icp_batch_plan;
int32_t gca_icp_batch_plan_create(gca_ctx ctx, int32_t n_pairs, int32_t n_levels,
                                  const gca_cloud *srcs, const gca_cloud *tgts,
                                  const gca_icp_params *params, const float *gradient_radius,
                                  const double *T0, int32_t use_graph, gca_icp_batch_plan *plan);
int32_t gca_icp_batch_plan_enqueue(gca_icp_batch_plan plan);
int32_t gca_icp_batch_plan_set_initial(gca_icp_batch_plan plan, const double *T0);
int32_t gca_icp_batch_plan_results(gca_icp_batch_plan plan, double *T_out,
                                   gca_icp_report *reports);
int32_t gca_icp_batch_plan_captures(gca_icp_batch_plan plan, int64_t *captures);
int32_t gca_icp_batch_plan_destroy(gca_icp_batch_plan plan);

/* Multi-scale (coarse-to-fine) colored ICP between two full-resolution clouds: per level,
 * voxel downsample both, normals + colour gradient of the target, then align, carrying the
 * pose.  Reference: the coarse-to-fine schedule of apps/test_color_icp.cpp:378-414. */
int32_t gca_multiscale_icp(gca_ctx ctx, gca_cloud src_full, gca_cloud tgt_full, int32_t n_levels,
                           const gca_level_params *levels, const double T0[16],
                           double T_out[16], gca_icp_report *reports);

/* ------------------------------------------------------------------------------------------ */
/* Frame pipeline                                                                            */

/* The per-frame odometry step of the reference's demo (apps/app.cpp:111-372: create_from_rgbd
 * -> voxel_grid_down_sample -> estimate_normals -> color_icp.align against the previous frame,
 * coarse to fine as in apps/test_color_icp.cpp:378-414) as one object.  A step runs the front
 * end, every level's voxel downsample, normals and colour gradient, and the chained
 * registration of the new frame onto the previous one, all enqueued without a host round trip
 * on the pipeline's own stream (a private ctx on ctx's device).  Levels: up to 8, coarse to
 * fine (voxel_size, normal radius / max_nn, gradient radius / max_nn, ICP parameters). */
typedef struct gca_pipeline_s *gca_pipeline;
int32_t gca_pipeline_create(gca_ctx ctx, const gca_camera *camera, const gca_bilateral *filter,
                            float threshold_min_m, float threshold_max_m, int32_t n_levels,
                            const gca_level_params *levels, gca_pipeline *out);
int32_t gca_pipeline_destroy(gca_pipeline pipeline);
/* Enqueue one frame (device Z16 depth and BGR8 colour of camera size; the caller orders its
 * writes to them before the call).  Asynchronous. */
int32_t gca_pipeline_step(gca_pipeline pipeline, const uint16_t *depth_dev, const uint8_t *bgr_dev);
/* Wait for the last step; T = pose mapping the last frame onto the one before (row-major),
 * reports[n_levels] (may be NULL); *valid = 0 for the first frame (T = identity).
 * GCA_ERR_INVALID_VALUE if a level's voxel lattice was refused (point_cloud.cu:213-218). */
int32_t gca_pipeline_result(gca_pipeline pipeline, double T[16], gca_icp_report *reports,
                            int32_t *valid);
/* Replay each frame slot's step as a hipGraph (captured after one eager step per slot, and again
 * when the ctx tuning changes; two-lane overlapped steps are two graphs per slot, the build and
 * the registration).  A step with other frame buffers than the graph's updates the one kernel node
 * that reads them (hipGraphExecKernelNodeSetParams) instead of capturing. */
int32_t gca_pipeline_set_graph(gca_pipeline pipeline, int32_t enable);
/* Graph captures and front-end node updates since the pipeline was created (no reference
 * counterpart: replay bookkeeping). */
int32_t gca_pipeline_graph_stats(gca_pipeline pipeline, int64_t *captures, int64_t *node_updates);
/* The pipeline's own ctx (its stream, stage profiler and tuning; synchronising it waits for
 * every step enqueued). */
int32_t gca_pipeline_ctx(gca_pipeline pipeline, gca_ctx *ctx);
/* Frame slots of the current mode (no reference counterpart): 2 sequential, 4 overlapped (3 when
 * the levels cannot be built in lockstep; see gca_pipeline_set_overlap).  Step k uses slot
 * k % slots. */
int32_t gca_pipeline_slots(gca_pipeline pipeline, int32_t *slots);
/* Point counts of a frame slot (0 .. slots - 1; the last step used slot (frames - 1) % slots):
 * the full cloud and level_n[n_levels] (may be NULL).  Synchronises. */
int32_t gca_pipeline_sizes(gca_pipeline pipeline, int32_t slot, int64_t *full_n, int64_t *level_n);
/* Overlapped steps (no reference counterpart: the reference's loop runs one frame at a time).
 * enable != 0: each step builds the new frame (front end, every level) and, beside it on a
 * second stream, registers the previous frame, whose clouds the previous step built.  When the
 * pipeline has at most 4 levels and each uses the joint normals + gradient walk (gradient max_nn
 * 0, gradient radius <= normal radius), the levels are built in lockstep (each build kernel
 * launched once for all levels) and consecutive steps alternate between two lanes of streams
 * and scratch, so two frames' steps also run side by side; four frame slots (else one lane,
 * per-level streams, three slots).  No two concurrent parts touch the same clouds.  A step's
 * work is the same, and every pose equals the sequential pipeline's bit for bit, but
 * gca_pipeline_result after a step reports the registration of the frame before the one just
 * stepped; gca_pipeline_flush registers the last stepped frame on its own.  Changing the mode
 * synchronises and starts a new sequence (the next frame is a first frame). */
int32_t gca_pipeline_set_overlap(gca_pipeline pipeline, int32_t enable);
/* Overlapped steps: enqueue the registration still pending for the last stepped frame, so that
 * gca_pipeline_result reports it.  No-op in sequential mode or when nothing is pending. */
int32_t gca_pipeline_flush(gca_pipeline pipeline);

#ifdef __cplusplus
}
#endif

#endif /* GCA_GCA_H */
