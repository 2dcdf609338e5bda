/*
 * gca_oracle.h -- FP64 host restatement of the reference's RGB-D front end and colored-ICP
 * path.  TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.  The product
 * (cuda_3d_amd, libgca.so) never calls it.
 *
 * Discrete decisions (pixel validity, colour pixel, voxel keys, grid cells, neighbour sets,
 * nearest neighbour, K-nearest cap) follow fixed canonical rules that the GPU library
 * implements identically, so those outputs compare bit-exactly.  Continuous quantities are
 * computed in FP64 here.  Every function cites the reference file:line it restates.
 */
#ifndef GCA_ORACLE_H
#define GCA_ORACLE_H

#include <stdint.h>

#include "../include/gca/gca.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Largest number of cells a neighbour grid may have; the cell size doubles until it fits. */
#define ORC_GRID_CELL_CAP (1LL << 25)

typedef struct orc_grid {
    float origin[3];
    float cell;      /* cell edge, >= search radius */
    int32_t dims[3];
    int64_t ncells;
    int32_t *start;  /* ncells + 1 (CSR over points sorted by cell, then index) */
    int32_t *order;  /* n point indices */
    int32_t n;
} orc_grid;

/* sizeof of the structs the Python binding mirrors, in this order: gca_camera, gca_bilateral,
 * gca_icp_params, gca_icp_report, gca_level_params, orc_grid (a mirror that drifts from the
 * header would let the library write past the caller's arrays). */
void orc_abi_sizes(int64_t out[6]);

/* cuda_point_cloud_factory.cu:189-267 + geometry_util.cuh:124-130 (compaction).
 * xyzi/rgb: float4 per output point (capacity W*H).  pix: source pixel index (may be NULL).
 * Returns the number of points, or -1 on invalid arguments (.cu:286-290). */
int64_t orc_cloud_from_rgbd(const uint16_t *depth, const uint8_t *bgr, const gca_camera *cam,
                            const gca_bilateral *bil, float tmin, float tmax, float *xyzi,
                            float *rgb, int32_t *pix, double *filtered_raw);

/* geometry_util.cuh:88-105 (max initialised with FLT_MIN, as the reference). */
void orc_bounds(const float *xyzi, int64_t n, float mn[3], float mx[3]);

/* type.hpp:45-51 color3::to_intensity. */
float orc_intensity(float r, float g, float b);

/* cuda_voxel_grid_down_sample.cu:26-161 with point_cloud.cu:198-229.  Keys in FP32 as the
 * reference, ordered lexicographically on (kx, ky, kz) like its int3 comparison sort (:136);
 * per-voxel sums in 2^-30 fixed point.  keys_out: int32[3*M] (kx, ky, kz) per output voxel, may
 * be NULL.  Returns M, or -1 for voxel <= 0 / an axis beyond INT_MAX voxels (point_cloud.cu:213-218,
 * the reference's only size guard). */
int64_t orc_voxel_downsample(const float *xyzi, const float *rgb, int64_t n, const float bmin[3],
                             const float bmax[3], float voxel, float *out_xyzi, float *out_rgb,
                             int32_t *keys_out);

/* Grid over a cloud with bounds (bmin,bmax) for search radius r (cuda_nn_search.cu:557-608;
 * padding 1.5 r from point_cloud.cu:155-160,528-537). */
int orc_grid_build(orc_grid *g, const float *xyzi, int64_t n, const float bmin[3],
                   const float bmax[3], float r);
void orc_grid_free(orc_grid *g);
void orc_grid_cells(const orc_grid *g, const float *xyzi, int64_t n, int32_t *cells);

/* cuda_nn_search.cu:113-214,545-619 (canonical argmin of (d^2, index), FP64 predicate). */
void orc_nn_search(const float *q_xyzi, int64_t nq, const float *r_xyzi, int64_t nr,
                   const float rbmin[3], const float rbmax[3], float radius, int32_t *out);

/* cuda_nn_search.cu:216-414,622-723.  Lists sorted by (d^2, index), max_nn cap (0 = none).
 * offsets: nq+1.  idx may be NULL (count only).  Returns the total. */
int64_t orc_radius_search(const float *q_xyzi, int64_t nq, const float *r_xyzi, int64_t nr,
                          const float rbmin[3], const float rbmax[3], float radius, int32_t max_nn,
                          int32_t *offsets, int32_t *idx);

/* cuda_estimate_normals.cu:22-199 (centred moments, 2^-40 fixed point, FP64 Jacobi eigen). */
void orc_estimate_normals(const float *xyzi, int64_t n, const float bmin[3], const float bmax[3],
                          float radius, int32_t max_nn, float *normals4);

/* Symmetric 3x3 eigen-decomposition (cyclic Jacobi, FP64); eigenvectors in columns of V.
 * Replaces fast_eigen_compute_3x3_symm (util/fast_eigen_val_vec.cuh:115-205). */
void orc_eigen3(const double A[9], double evals[3], double V[9]);

/* cuda_compute_color_gradient.cu:13-127 + little_3x3_ldlt_solve.cuh:6-67. */
void orc_color_gradient(const float *xyzi, const float *normals4, int64_t n, const float bmin[3],
                        const float bmax[3], float radius, int32_t max_nn, float *grad4);

/* cuda_color_icp_build_least_square.cu:134-266,431-474, FP64 throughout.  T row-major 4x4.
 * out: 21 upper JTJ, 6 JTr, r^2, n_corr. */
void orc_build_gn(const float *src_xyzi, int64_t ns, const float *tgt_xyzi, const float *tgt_n4,
                  const float *tgt_g4, const int32_t *nn, float lambda, const double T[16],
                  double out[29]);

/* Residual variants (.cu:13-132,268-429).  kind GCA_RESIDUAL_*.  out: 2 per point for PAIR,
 * else 1.  Returns the FP64 sum of the per-point values (PAIR: sum of both). */
double orc_residuals(const float *src_xyzi, int64_t ns, const float *tgt_xyzi, const float *tgt_n4,
                     const float *tgt_g4, const int32_t *nn, float lambda, const double T[16],
                     int kind, double *out);

/* eigen_solver.hpp:44-85: LDL^T of JTJ, xi = -JTJ^-1 JTr, T_step = [Rz Ry Rx | xi3..5].
 * Returns 0, or 4 (not SPD -> identity). */
int orc_solve6(const double jtj21[21], const double jtr[6], double xi[6], double T_step[16]);
void orc_euler_to_T(const double xi[6], double T[16]);
void orc_matmul4(const double A[16], const double B[16], double C[16]);

/* color_icp.cu:151-234 (device loop semantics: NN, build, solve, T <- dT T, stop rule). */
void orc_icp_align(const float *src_xyzi, int64_t ns, const float *tgt_xyzi, const float *tgt_n4,
                   const float *tgt_g4, int64_t nt, const float tbmin[3], const float tbmax[3],
                   const gca_icp_params *p, const double T0[16], double T_out[16],
                   gca_icp_report *rep);

/* Objective of compute_RMSE_functor (.cu:82-132) at dT(xi) T with frozen nn, and adjoint. */
double orc_objective_grad(const float *src_xyzi, int64_t ns, const float *tgt_xyzi,
                          const float *tgt_n4, const float *tgt_g4, const int32_t *nn,
                          float lambda, const double T[16], const double xi[6], double dxi[6],
                          double *dpts4);

/* cuda_nn_search.cu:416-542,725-800 (keep iff #within radius incl. self > min_nn). */
int64_t orc_radius_outlier_removal(const float *xyzi, const float *rgb, int64_t n,
                                   const float bmin[3], const float bmax[3], float radius,
                                   int32_t min_nn, float *out_xyzi, float *out_rgb);

/* Coarse-to-fine driver (apps/test_color_icp.cpp:378-414 schedule). */
int orc_multiscale_icp(const float *src_xyzi, const float *src_rgb, int64_t ns,
                       const float *tgt_xyzi, const float *tgt_rgb, int64_t nt, int n_levels,
                       const gca_level_params *levels, const double T0[16], double T_out[16],
                       gca_icp_report *reps);

#ifdef __cplusplus
}
#endif
#endif
