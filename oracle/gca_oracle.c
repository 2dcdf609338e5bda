/*
 * gca_oracle.c -- FP64 host restatement of the reference path.  TEST INFRASTRUCTURE ONLY
 * (the checker and the CPU baseline); never linked into the product.
 *
 * Build: see oracle/Makefile (-O2 -ffp-contract=off: every FP64 operation is a single IEEE
 * rounding, so the discrete rules below are evaluated exactly as written).
 *
 * Canonical rules shared with the GPU library (each a deliberate, documented definition):
 *  - bilateral weights in FP64 from exp() of the exact tap arguments, summed in row-major tap
 *    order; the centre pixel decides validity before filtering;
 *  - back-projection, colour pixel and intensity in the reference's FP32/FP64 operation order;
 *  - voxel keys in FP32 exactly as the reference; per-voxel sums in 2^-30 fixed point;
 *  - grid cell = floor(((double)p - origin) / cell) and the query scans the cells overlapped by
 *    [q - r, q + r], so the neighbour set is exactly { j : d2_fp64(q, p_j) < r^2 };
 *  - nearest neighbour = argmin (d^2, index); K-cap keeps the K smallest (d^2, index);
 *  - neighbourhood moments and colour-gradient normal equations in 2^-40 fixed point
 *    (scaled by the radius), so they do not depend on traversal order.
 */
#include "gca_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FX30 1073741824.0          /* 2^30 */
#define FX40 1099511627776.0       /* 2^40 */

/* ORC_PLANT selects a deliberately wrong variant, compiled only by the test that checks the
 * pins reject it (tests/test_oracle.py::test_planted_mistakes_are_caught); 0 = correct. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

/* ------------------------------------------------------------------------------------------ */
/* Front end                                                                                 */

float orc_intensity(float r, float g, float b)
{
    /* type.hpp:45-51: evaluated in double, returned as float. */
    return (float)(0.2126 * (double)r + 0.7152 * (double)g + 0.0722 * (double)b);
}

/* Zero-skipping bilateral (cuda_point_cloud_factory.cu:64-98) in FP64. */
static double bilateral_px(const uint16_t *depth, int W, int H, int x, int y, int R, double ss,
                           double sd)
{
    uint16_t d0 = depth[(int64_t)y * W + x];
    double sw = 0.0, swd = 0.0;
    for (int dy = -R; dy <= R; ++dy) {
        for (int dx = -R; dx <= R; ++dx) {
            int nx = x + dx, ny = y + dy;
            if (nx < 0 || nx >= W || ny < 0 || ny >= H) continue;
            uint16_t dn = depth[(int64_t)ny * W + nx];
            if (!dn) continue;
            int diff = abs((int)d0 - (int)dn);
#if ORC_PLANT == 6 /* textbook Gaussian with the 1/2 factor (reference has none) */
            double ws = exp(-((double)(dx * dx + dy * dy)) / (2.0 * ss * ss));
#else
            double ws = exp(-((double)(dx * dx + dy * dy)) / (ss * ss));
#endif
            double wr = exp(-((double)diff * (double)diff) / (sd * sd));
            double w = ws * wr;
            sw += w;
            swd += w * (double)dn;
        }
    }
    return swd / sw;
}

int64_t orc_cloud_from_rgbd(const uint16_t *depth, const uint8_t *bgr, const gca_camera *cam,
                            const gca_bilateral *bil, float tmin, float tmax, float *xyzi,
                            float *rgb, int32_t *pix, double *filtered_raw)
{
    if (!cam || cam->width <= 0 || cam->height <= 0 || !(cam->depth_scale > 0.0f)) return -1;
    const int W = cam->width, H = cam->height;
    const float scale = cam->depth_scale;
    const gca_intrinsics di = cam->depth_intrin, ci = cam->color_intrin;
    const float *Rm = cam->depth_to_color.rotation, *tv = cam->depth_to_color.translation;
    const float inv255 = 1.0f / 255.0f;
    int64_t n = 0;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            uint16_t d0 = depth[(int64_t)y * W + x];
            if (filtered_raw) filtered_raw[(int64_t)y * W + x] = 0.0;
            if (!d0) continue; /* centre hole: invalid (documented deviation, quirk #2) */
            float z0 = (float)d0 * scale;
            if (z0 < tmin || z0 > tmax) continue;
            double f = (double)d0;
            if (bil && bil->mode == GCA_BILATERAL_FIXED) {
                f = bilateral_px(depth, W, H, x, y, bil->radius, (double)bil->sigma_space,
                                 (double)bil->sigma_depth);
            } else if (bil && bil->mode == GCA_BILATERAL_REFERENCE_ADAPTIVE) {
                /* cuda_point_cloud_factory.cu:100-129 */
                unsigned step = (unsigned)((z0 - tmin) / 0.5f);
                if (step > 5u) step = 5u;
                int R = (int)(1.0f * (float)step);
                f = bilateral_px(depth, W, H, x, y, R, (double)(8.0f + (float)step * 1.0f),
                                 (double)(100.0f + (float)step * 10.0f));
            }
            if (filtered_raw) filtered_raw[(int64_t)y * W + x] = f;
            float z = (float)f * scale;
            if (z < tmin || z > tmax) continue;
            /* __depth_uv_to_xyz (:172-180) with the (x - 0.5, y - 0.5) pixel centre (:232) */
            float u = (float)x - 0.5f, v = (float)y - 0.5f;
            float X = ((u - di.cx) * z) / di.fx;
            float Y = ((v - di.cy) * z) / di.fy;
            float Z = z;
            /* __transform_point_to_point (:160-170), column-major rotation */
            float c0 = ((Rm[0] * X + Rm[3] * Y) + Rm[6] * Z) + tv[0];
            float c1 = ((Rm[1] * X + Rm[4] * Y) + Rm[7] * Z) + tv[1];
            float c2 = ((Rm[2] * X + Rm[5] * Y) + Rm[8] * Z) + tv[2];
            if (!(c2 > 0.0f)) continue;
            /* __xyz_to_color_uv (:182-187) and int(u + 0.5) (:242-248) */
            float fu = ((c0 * ci.fx) / c2) + ci.cx;
            float fv = ((c1 * ci.fy) / c2) + ci.cy;
            fu = fu + 0.5f;
            fv = fv + 0.5f;
            if (!(fu > -1.0f && fu < (float)W && fv > -1.0f && fv < (float)H)) continue;
            int tx = (int)fu, ty = (int)fv;
            const uint8_t *c = bgr + 3 * ((int64_t)ty * W + tx);
            float b = (float)c[0] * inv255, g = (float)c[1] * inv255, r = (float)c[2] * inv255;
            xyzi[4 * n + 0] = X;
            xyzi[4 * n + 1] = Y;
            xyzi[4 * n + 2] = Z;
            xyzi[4 * n + 3] = orc_intensity(r, g, b);
            rgb[4 * n + 0] = r;
            rgb[4 * n + 1] = g;
            rgb[4 * n + 2] = b;
            rgb[4 * n + 3] = 0.0f;
            if (pix) pix[n] = y * W + x;
            ++n;
        }
    }
    return n;
}

void orc_bounds(const float *xyzi, int64_t n, float mn[3], float mx[3])
{
    for (int a = 0; a < 3; ++a) {
        mn[a] = FLT_MAX;
        mx[a] = FLT_MIN; /* geometry_util.cuh:92 */
    }
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = xyzi[4 * i + a];
            if (v < mn[a]) mn[a] = v;
            if (v > mx[a]) mx[a] = v;
        }
}

/* ------------------------------------------------------------------------------------------ */
/* Voxel grid downsample                                                                     */

typedef struct {
    int32_t k[3];
    int64_t idx;
} key_idx;

/* Lexicographic (kx, ky, kz), then source index: the reference's int3 comparison sort
 * (cuda_voxel_grid_down_sample.cu:50-62,136) made stable. */
static int cmp_key_idx(const void *a, const void *b)
{
    const key_idx *x = (const key_idx *)a, *y = (const key_idx *)b;
    for (int c = 0; c < 3; ++c)
        if (x->k[c] != y->k[c]) return x->k[c] < y->k[c] ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int same_key(const key_idx *x, const key_idx *y)
{
    return x->k[0] == y->k[0] && x->k[1] == y->k[1] && x->k[2] == y->k[2];
}

int64_t orc_voxel_downsample(const float *xyzi, const float *rgb, int64_t n, const float bmin[3],
                             const float bmax[3], float voxel, float *out_xyzi, float *out_rgb,
                             int32_t *keys_out)
{
    if (!(voxel > 0.0f)) return -1; /* point_cloud.cu:200-204 */
    for (int a = 0; a < 3; ++a) {
        float ext = bmax[a] - bmin[a];
        if ((double)voxel * 2147483647.0 < (double)ext) return -1; /* point_cloud.cu:213-218 */
    }
    if (n == 0) return 0;
    key_idx *ki = (key_idx *)malloc(sizeof(key_idx) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        /* __float2int_rd((p - min) / v) per axis, FP32 (cuda_voxel_grid_down_sample.cu:26-48) */
        for (int a = 0; a < 3; ++a) ki[i].k[a] = (int32_t)floorf((xyzi[4 * i + a] - bmin[a]) / voxel);
        ki[i].idx = i;
    }
    qsort(ki, (size_t)n, sizeof(key_idx), cmp_key_idx);
    int64_t m = 0;
    for (int64_t s = 0; s < n;) {
        int64_t e = s;
        int64_t S[6] = {0, 0, 0, 0, 0, 0};
        while (e < n && same_key(&ki[e], &ki[s])) {
            int64_t i = ki[e].idx;
            for (int a = 0; a < 3; ++a) {
                S[a] += llrint((double)xyzi[4 * i + a] * FX30);
                S[3 + a] += llrint((double)rgb[4 * i + a] * FX30);
            }
            ++e;
        }
        double cnt = (double)(e - s);
        float o[6];
        for (int c = 0; c < 6; ++c) o[c] = (float)(((double)S[c] / cnt) * (1.0 / FX30));
        out_xyzi[4 * m + 0] = o[0];
        out_xyzi[4 * m + 1] = o[1];
        out_xyzi[4 * m + 2] = o[2];
        out_xyzi[4 * m + 3] = orc_intensity(o[3], o[4], o[5]);
        out_rgb[4 * m + 0] = o[3];
        out_rgb[4 * m + 1] = o[4];
        out_rgb[4 * m + 2] = o[5];
        out_rgb[4 * m + 3] = 0.0f;
        if (keys_out)
            for (int a = 0; a < 3; ++a) keys_out[3 * m + a] = ki[s].k[a];
        ++m;
        s = e;
    }
    free(ki);
    return m;
}

/* ------------------------------------------------------------------------------------------ */
/* Neighbour grid                                                                            */

static void grid_params(const float bmin[3], const float bmax[3], float r, float origin[3],
                        float *cell, int32_t dims[3])
{
    float pad = 1.5f * r;
    float top[3];
    for (int a = 0; a < 3; ++a) {
        origin[a] = bmin[a] - pad;
        top[a] = bmax[a] + pad;
    }
    float s = r;
    for (;;) {
        double total = 1.0;
        for (int a = 0; a < 3; ++a) {
            float q = (top[a] - origin[a]) / s;
            dims[a] = (q < 2.0e9f) ? (int32_t)q + 1 : 2000000000;
            total *= (double)dims[a];
        }
        if (total <= (double)ORC_GRID_CELL_CAP) break;
        s = s * 2.0f;
    }
    *cell = s;
}

static inline int32_t cell_of(double p, float o, float s, int32_t dim)
{
    double c = floor(((double)p - (double)o) / (double)s);
    if (c < 0.0) c = 0.0;
    if (c > (double)(dim - 1)) c = (double)(dim - 1);
    return (int32_t)c;
}

int orc_grid_build(orc_grid *g, const float *xyzi, int64_t n, const float bmin[3],
                   const float bmax[3], float r)
{
    memset(g, 0, sizeof(*g));
    grid_params(bmin, bmax, r, g->origin, &g->cell, g->dims);
    g->ncells = (int64_t)g->dims[0] * g->dims[1] * g->dims[2];
    g->n = (int32_t)n;
    g->start = (int32_t *)calloc((size_t)g->ncells + 1, sizeof(int32_t));
    g->order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t *lin = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        int32_t c[3];
        for (int a = 0; a < 3; ++a) c[a] = cell_of(xyzi[4 * i + a], g->origin[a], g->cell, g->dims[a]);
        lin[i] = ((int64_t)c[0] * g->dims[1] + c[1]) * g->dims[2] + c[2];
        g->start[lin[i] + 1]++;
    }
    for (int64_t c = 0; c < g->ncells; ++c) g->start[c + 1] += g->start[c];
    int32_t *fill = (int32_t *)calloc((size_t)g->ncells, sizeof(int32_t));
    for (int64_t i = 0; i < n; ++i) g->order[g->start[lin[i]] + fill[lin[i]]++] = (int32_t)i;
    free(fill);
    free(lin);
    return 0;
}

void orc_grid_free(orc_grid *g)
{
    free(g->start);
    free(g->order);
    memset(g, 0, sizeof(*g));
}

void orc_grid_cells(const orc_grid *g, const float *xyzi, int64_t n, int32_t *cells)
{
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
            cells[3 * i + a] = cell_of(xyzi[4 * i + a], g->origin[a], g->cell, g->dims[a]);
}

/* Cell range overlapped by [q - r, q + r]; returns 0 if empty. */
static int query_range(const orc_grid *g, const double q[3], float r, int32_t lo[3], int32_t hi[3])
{
    for (int a = 0; a < 3; ++a) {
        if (!isfinite(q[a])) return 0;
        double l = floor((q[a] - (double)r - (double)g->origin[a]) / (double)g->cell);
        double h = floor((q[a] + (double)r - (double)g->origin[a]) / (double)g->cell);
        if (l < 0.0) l = 0.0;
        if (h > (double)(g->dims[a] - 1)) h = (double)(g->dims[a] - 1);
        if (l > h) return 0;
        lo[a] = (int32_t)l;
        hi[a] = (int32_t)h;
    }
    return 1;
}

static inline double dist2(const double q[3], const float *p)
{
    double dx = (double)p[0] - q[0], dy = (double)p[1] - q[1], dz = (double)p[2] - q[2];
    return (dx * dx + dy * dy) + dz * dz;
}

typedef struct {
    double d2;
    int32_t idx;
} cand;

static int cmp_cand(const void *a, const void *b)
{
    const cand *x = (const cand *)a, *y = (const cand *)b;
    if (x->d2 != y->d2) return x->d2 < y->d2 ? -1 : 1;
#if ORC_PLANT == 5 /* ties resolved toward the higher index */
    return x->idx > y->idx ? -1 : (x->idx < y->idx);
#else
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
#endif
}

/* All candidates with d2 < r^2 around q, sorted by (d2, idx) and capped. */
static int32_t gather(const orc_grid *g, const float *r_xyzi, const double q[3], float radius,
                      int32_t max_nn, cand **buf, int32_t *cap)
{
    int32_t lo[3], hi[3];
    if (!query_range(g, q, radius, lo, hi)) return 0;
    double r2 = (double)radius * (double)radius;
    int32_t k = 0;
    for (int32_t ix = lo[0]; ix <= hi[0]; ++ix)
        for (int32_t iy = lo[1]; iy <= hi[1]; ++iy)
            for (int32_t iz = lo[2]; iz <= hi[2]; ++iz) {
                int64_t c = ((int64_t)ix * g->dims[1] + iy) * g->dims[2] + iz;
                for (int32_t s = g->start[c]; s < g->start[c + 1]; ++s) {
                    int32_t j = g->order[s];
                    double d2 = dist2(q, r_xyzi + 4 * (int64_t)j);
                    if (d2 < r2) {
                        if (k == *cap) {
                            *cap = *cap ? 2 * *cap : 256;
                            *buf = (cand *)realloc(*buf, sizeof(cand) * (size_t)*cap);
                        }
                        (*buf)[k].d2 = d2;
                        (*buf)[k].idx = j;
                        ++k;
                    }
                }
            }
    qsort(*buf, (size_t)k, sizeof(cand), cmp_cand);
    if (max_nn > 0 && k > max_nn) k = max_nn;
    return k;
}

void orc_nn_search(const float *q_xyzi, int64_t nq, const float *r_xyzi, int64_t nr,
                   const float rbmin[3], const float rbmax[3], float radius, int32_t *out)
{
    orc_grid g;
    orc_grid_build(&g, r_xyzi, nr, rbmin, rbmax, radius);
    cand *buf = NULL;
    int32_t cap = 0;
    for (int64_t i = 0; i < nq; ++i) {
        double q[3] = {q_xyzi[4 * i], q_xyzi[4 * i + 1], q_xyzi[4 * i + 2]};
        int32_t k = gather(&g, r_xyzi, q, radius, 1, &buf, &cap);
        out[i] = k ? buf[0].idx : -1;
    }
    free(buf);
    orc_grid_free(&g);
}

int64_t orc_radius_search(const float *q_xyzi, int64_t nq, const float *r_xyzi, int64_t nr,
                          const float rbmin[3], const float rbmax[3], float radius, int32_t max_nn,
                          int32_t *offsets, int32_t *idx)
{
    orc_grid g;
    orc_grid_build(&g, r_xyzi, nr, rbmin, rbmax, radius);
    cand *buf = NULL;
    int32_t cap = 0;
    int64_t total = 0;
    for (int64_t i = 0; i < nq; ++i) {
        double q[3] = {q_xyzi[4 * i], q_xyzi[4 * i + 1], q_xyzi[4 * i + 2]};
        int32_t k = gather(&g, r_xyzi, q, radius, max_nn, &buf, &cap);
        if (offsets) offsets[i] = (int32_t)total;
        if (idx)
            for (int32_t t = 0; t < k; ++t) idx[total + t] = buf[t].idx;
        total += k;
    }
    if (offsets) offsets[nq] = (int32_t)total;
    free(buf);
    orc_grid_free(&g);
    return total;
}

/* ------------------------------------------------------------------------------------------ */
/* Normals                                                                                   */

void orc_eigen3(const double Ain[9], double evals[3], double V[9])
{
    double A[3][3], v[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[i][j] = Ain[3 * i + j];
            v[i][j] = (i == j) ? 1.0 : 0.0;
        }
    static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = (A[0][1] * A[0][1] + A[0][2] * A[0][2]) + A[1][2] * A[1][2];
        double dg = (A[0][0] * A[0][0] + A[1][1] * A[1][1]) + A[2][2] * A[2][2];
        if (!(off > 1e-36 * dg)) break;
        for (int r = 0; r < 3; ++r) {
            int p = P[r], q = Q[r];
            double apq = A[p][q];
            if (apq == 0.0) continue;
            double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
            if (theta < 0.0) t = -t;
            double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; ++k) { /* A <- A J (columns p, q) */
                double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - s * akq;
                A[k][q] = s * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) { /* A <- J^T A (rows p, q) */
                double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - s * aqk;
                A[q][k] = s * apk + c * aqk;
            }
            A[p][q] = 0.0;
            A[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    for (int i = 0; i < 3; ++i) {
        evals[i] = A[i][i];
        for (int j = 0; j < 3; ++j) V[3 * i + j] = v[i][j];
    }
}

/* Normal from fixed-point centred moments (count, S1[3], S2[6] = xx xy xz yy yz zz). */
static void normal_from_moments(int64_t cnt, const int64_t S1[3], const int64_t S2[6], float out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    if (cnt < 3) return; /* cuda_estimate_normals.cu:89-93 */
    double n = (double)cnt;
    double m[3], C[9];
    for (int a = 0; a < 3; ++a) m[a] = ((double)S1[a] / FX40) / n;
    static const int IA[6] = {0, 0, 0, 1, 1, 2}, IB[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
        double c = ((double)S2[k] / FX40) / n - m[IA[k]] * m[IB[k]];
        C[3 * IA[k] + IB[k]] = c;
        C[3 * IB[k] + IA[k]] = c;
    }
    double ev[3], V[9];
    orc_eigen3(C, ev, V);
    /* strict-< selection of the smallest eigenvalue (cuda_estimate_normals.cu:116-125) */
    int sel;
#if ORC_PLANT == 4 /* largest eigenvalue instead of the smallest */
    if (ev[0] > ev[1] && ev[0] > ev[2]) sel = 0;
    else if (ev[1] > ev[0] && ev[1] > ev[2]) sel = 1;
    else sel = 2;
#else
    if (ev[0] < ev[1] && ev[0] < ev[2]) sel = 0;
    else if (ev[1] < ev[0] && ev[1] < ev[2]) sel = 1;
    else sel = 2;
#endif
    double v[3] = {V[0 + sel], V[3 + sel], V[6 + sel]};
    double len = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (!(len > 0.0)) return;
    for (int a = 0; a < 3; ++a) v[a] /= len;
    /* orient toward (0,0,-1) (cuda_estimate_normals.cu:131-151) */
    if (-v[2] < 0.0)
        for (int a = 0; a < 3; ++a) v[a] = -v[a];
    out[0] = (float)v[0];
    out[1] = (float)v[1];
    out[2] = (float)v[2];
}

void orc_estimate_normals(const float *xyzi, int64_t n, const float bmin[3], const float bmax[3],
                          float radius, int32_t max_nn, float *normals4)
{
    orc_grid g;
    orc_grid_build(&g, xyzi, n, bmin, bmax, radius);
    cand *buf = NULL;
    int32_t cap = 0;
    const double inv_r = 1.0 / (double)radius;
    for (int64_t i = 0; i < n; ++i) {
        double q[3] = {xyzi[4 * i], xyzi[4 * i + 1], xyzi[4 * i + 2]};
        int32_t k = gather(&g, xyzi, q, radius, max_nn, &buf, &cap);
        int64_t S1[3] = {0, 0, 0}, S2[6] = {0, 0, 0, 0, 0, 0};
        for (int32_t t = 0; t < k; ++t) {
            const float *p = xyzi + 4 * (int64_t)buf[t].idx;
            double u[3];
            for (int a = 0; a < 3; ++a) u[a] = ((double)p[a] - q[a]) * inv_r;
            for (int a = 0; a < 3; ++a) S1[a] += llrint(u[a] * FX40);
            S2[0] += llrint((u[0] * u[0]) * FX40);
            S2[1] += llrint((u[0] * u[1]) * FX40);
            S2[2] += llrint((u[0] * u[2]) * FX40);
            S2[3] += llrint((u[1] * u[1]) * FX40);
            S2[4] += llrint((u[1] * u[2]) * FX40);
            S2[5] += llrint((u[2] * u[2]) * FX40);
        }
        normal_from_moments(k, S1, S2, normals4 + 4 * i);
    }
    free(buf);
    orc_grid_free(&g);
}

/* ------------------------------------------------------------------------------------------ */
/* Colour gradient                                                                           */

/* P E P and P F for P = I - n n^T, with E (xx xy xz yy yz zz) and F the fixed-point sums: the
 * canonical FP64 operation order (the device's gradient solve uses the same expressions). */
static void project_tangent(const int64_t SE[6], const int64_t SF[3], const double n[3], double PMP[6],
                            double Pf[3])
{
    static const int IA[6] = {0, 0, 0, 1, 1, 2}, IB[6] = {0, 1, 2, 1, 2, 2};
    double M[3][3], f[3], Mn[3];
    for (int c = 0; c < 6; ++c) {
        double v = (double)SE[c] / FX40;
        M[IA[c]][IB[c]] = v;
        M[IB[c]][IA[c]] = v;
    }
    for (int a = 0; a < 3; ++a) f[a] = (double)SF[a] / FX40;
    for (int a = 0; a < 3; ++a) Mn[a] = (M[a][0] * n[0] + M[a][1] * n[1]) + M[a][2] * n[2];
    double nMn = (n[0] * Mn[0] + n[1] * Mn[1]) + n[2] * Mn[2];
    for (int c = 0; c < 6; ++c) {
        int a = IA[c], b = IB[c];
        PMP[c] = ((M[a][b] - n[a] * Mn[b]) - Mn[a] * n[b]) + (n[a] * n[b]) * nMn;
    }
    double fn = (f[0] * n[0] + f[1] * n[1]) + f[2] * n[2];
    for (int a = 0; a < 3; ++a) Pf[a] = f[a] - fn * n[a];
}

void orc_color_gradient(const float *xyzi, const float *normals4, int64_t n, const float bmin[3],
                        const float bmax[3], float radius, int32_t max_nn, float *grad4)
{
    orc_grid g;
    orc_grid_build(&g, xyzi, n, bmin, bmax, radius);
    cand *buf = NULL;
    int32_t cap = 0;
    const double inv_r = 1.0 / (double)radius;
    const double rr = (double)radius * (double)radius;
    for (int64_t i = 0; i < n; ++i) {
        float *out = grad4 + 4 * i;
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        double q[3] = {xyzi[4 * i], xyzi[4 * i + 1], xyzi[4 * i + 2]};
        int32_t k = gather(&g, xyzi, q, radius, max_nn, &buf, &cap);
        if (k < 5) continue; /* cuda_compute_color_gradient.cu:41-44 */
        double nv[3] = {normals4[4 * i], normals4[4 * i + 1], normals4[4 * i + 2]};
        double Ip = (double)xyzi[4 * i + 3];
        /* Raw neighbourhood sums of u = (p - q) / r: E = sum u u^T, F = sum u dI (exact in fixed
         * point); the tangent projection P = I - n n^T of :48-60 is applied to the sums below
         * (sum (P u)(P u)^T = P E P, sum (P u) dI = P F), so the sums do not need the normal. */
        int64_t SE[6] = {0, 0, 0, 0, 0, 0}, SF[3] = {0, 0, 0};
        for (int32_t t = 0; t < k; ++t) {
            int32_t j = buf[t].idx;
            if (j == (int32_t)i) continue; /* :61-62 */
            const float *p = xyzi + 4 * (int64_t)j;
            double u[3] = {((double)p[0] - q[0]) * inv_r, ((double)p[1] - q[1]) * inv_r, ((double)p[2] - q[2]) * inv_r};
            double dI = (double)p[3] - Ip;
            SE[0] += llrint((u[0] * u[0]) * FX40);
            SE[1] += llrint((u[0] * u[1]) * FX40);
            SE[2] += llrint((u[0] * u[2]) * FX40);
            SE[3] += llrint((u[1] * u[1]) * FX40);
            SE[4] += llrint((u[1] * u[2]) * FX40);
            SE[5] += llrint((u[2] * u[2]) * FX40);
            for (int a = 0; a < 3; ++a) SF[a] += llrint((u[a] * dI) * FX40);
        }
        double PMP[6], Pf[3];
        project_tangent(SE, SF, nv, PMP, Pf);
        /* :75-84: + (k-1)^2 n n^T + 1e-6 I */
#if ORC_PLANT == 2 /* orthogonality constraint term dropped */
        double kk = 0.0;
#else
        double kk = (double)(k - 1) * (double)(k - 1);
#endif
        double A[3][3];
        static const int IA[6] = {0, 0, 0, 1, 1, 2}, IB[6] = {0, 1, 2, 1, 2, 2};
        for (int c = 0; c < 6; ++c) {
            double v = PMP[c] * rr + kk * (nv[IA[c]] * nv[IB[c]]);
            A[IA[c]][IB[c]] = v;
            A[IB[c]][IA[c]] = v;
        }
        for (int a = 0; a < 3; ++a) A[a][a] += 1.0e-6;
        double b[3];
        for (int a = 0; a < 3; ++a) b[a] = Pf[a] * (double)radius;
        /* unpivoted LDL^T (little_3x3_ldlt_solve.cuh:11-60) */
        double D0 = A[0][0];
        double L10 = A[1][0] / D0, L20 = A[2][0] / D0;
        double D1 = A[1][1] - L10 * L10 * D0;
        double L21 = (A[2][1] - L10 * L20 * D0) / D1;
        double D2 = A[2][2] - L20 * L20 * D0 - L21 * L21 * D1;
        double x0 = b[0];
        double x1 = b[1] - L10 * x0;
        double x2 = b[2] - L20 * x0 - L21 * x1;
        x0 /= D0;
        x1 /= D1;
        x2 /= D2;
        x1 -= L21 * x2;
        x0 -= L10 * x1 + L20 * x2;
        out[0] = (float)x0;
        out[1] = (float)x1;
        out[2] = (float)x2;
    }
    free(buf);
    orc_grid_free(&g);
}

/* ------------------------------------------------------------------------------------------ */
/* Colored ICP                                                                               */

static inline void xform(const double T[16], const float *p, double y[3])
{
    for (int a = 0; a < 3; ++a)
        y[a] = ((T[4 * a] * (double)p[0] + T[4 * a + 1] * (double)p[1]) + T[4 * a + 2] * (double)p[2]) +
               T[4 * a + 3];
}

static inline void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static inline double dot3(const double a[3], const double b[3])
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* Residuals and Jacobians of one correspondence, FP64
 * (cuda_color_icp_build_least_square.cu:184-253). */
static void residual_jac(const double p[3], const float *q4, const float *n4, const float *g4,
                         double Ip, double sl, double sc, double *rg, double *rc, double Jg[6],
                         double Jc[6], double *rg_raw, double *rc_raw, double gvec[3])
{
    double q[3] = {q4[0], q4[1], q4[2]}, n[3] = {n4[0], n4[1], n4[2]}, d[3] = {g4[0], g4[1], g4[2]};
    double dq[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    double r_g = dot3(dq, n);
    double c[3];
#if ORC_PLANT == 1 /* transposed cross product: n x p */
    cross3(n, p, c);
#else
    cross3(p, n, c);
#endif
    for (int a = 0; a < 3; ++a) {
        Jg[a] = sl * c[a];
        Jg[3 + a] = sl * n[a];
    }
    double dn = dot3(d, n);
    double g[3] = {d[0] - dn * n[0], d[1] - dn * n[1], d[2] - dn * n[2]};
    cross3(p, g, c);
    for (int a = 0; a < 3; ++a) {
        Jc[a] = sc * c[a];
        Jc[3 + a] = sc * g[a];
    }
    double proj[3] = {p[0] - r_g * n[0] - q[0], p[1] - r_g * n[1] - q[1], p[2] - r_g * n[2] - q[2]};
    double Iproj = (double)q4[3] + dot3(d, proj);
    double r_c = Iproj - Ip;
    *rg = sl * r_g;
    *rc = sc * r_c;
    if (rg_raw) *rg_raw = r_g;
    if (rc_raw) *rc_raw = r_c;
    if (gvec)
        for (int a = 0; a < 3; ++a) gvec[a] = g[a];
}

void orc_build_gn(const float *src_xyzi, int64_t ns, const float *tgt_xyzi, const float *tgt_n4,
                  const float *tgt_g4, const int32_t *nn, float lambda, const double T[16],
                  double out[29])
{
    memset(out, 0, sizeof(double) * 29);
    double sl = sqrt((double)lambda), sc = sqrt(1.0 - (double)lambda);
    for (int64_t i = 0; i < ns; ++i) {
        int32_t j = nn[i];
        if (j < 0) continue;
        double p[3], rg, rc, Jg[6], Jc[6];
        xform(T, src_xyzi + 4 * i, p);
        residual_jac(p, tgt_xyzi + 4 * (int64_t)j, tgt_n4 + 4 * (int64_t)j, tgt_g4 + 4 * (int64_t)j,
                     (double)src_xyzi[4 * i + 3], sl, sc, &rg, &rc, Jg, Jc, NULL, NULL, NULL);
        int k = 0;
        for (int a = 0; a < 6; ++a)
            for (int b = a; b < 6; ++b) out[k++] += Jg[a] * Jg[b] + Jc[a] * Jc[b];
        for (int a = 0; a < 6; ++a) out[21 + a] += Jg[a] * rg + Jc[a] * rc;
        out[27] += rg * rg + rc * rc;
        out[28] += 1.0;
    }
}

double orc_residuals(const float *src_xyzi, int64_t ns, const float *tgt_xyzi, const float *tgt_n4,
                     const float *tgt_g4, const int32_t *nn, float lambda, const double T[16],
                     int kind, double *out)
{
    double sl = sqrt((double)lambda), sc = sqrt(1.0 - (double)lambda), total = 0.0;
    for (int64_t i = 0; i < ns; ++i) {
        int32_t j = nn[i];
        double rg = 0.0, rc = 0.0, rgr = 0.0, rcr = 0.0, Jg[6], Jc[6];
        if (j >= 0) {
            double p[3];
            xform(T, src_xyzi + 4 * i, p);
            residual_jac(p, tgt_xyzi + 4 * (int64_t)j, tgt_n4 + 4 * (int64_t)j,
                         tgt_g4 + 4 * (int64_t)j, (double)src_xyzi[4 * i + 3], sl, sc, &rg, &rc, Jg,
                         Jc, &rgr, &rcr, NULL);
        }
        if (kind == GCA_RESIDUAL_PAIR) {
            if (out) {
                out[2 * i] = rg;
                out[2 * i + 1] = rc;
            }
            total += rg + rc;
        } else if (kind == GCA_RESIDUAL_ABS) {
            double v = fabs(rg) + fabs(rc);
            if (out) out[i] = v;
            total += v;
        } else {
#if ORC_PLANT == 7 /* squares the already-weighted residuals */
            double v = (double)lambda * rg * rg + (1.0 - (double)lambda) * rc * rc;
#else
            /* compute_RMSE_functor (cuda_color_icp_build_least_square.cu:82-132): unweighted */
            double v = (double)lambda * rgr * rgr + (1.0 - (double)lambda) * rcr * rcr;
#endif
            if (out) out[i] = v;
            total += v;
        }
    }
    return total;
}

void orc_matmul4(const double A[16], const double B[16], double C[16])
{
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            t[4 * i + j] = ((A[4 * i] * B[j] + A[4 * i + 1] * B[4 + j]) + A[4 * i + 2] * B[8 + j]) +
                           A[4 * i + 3] * B[12 + j];
    memcpy(C, t, sizeof(t));
}

void orc_euler_to_T(const double xi[6], double T[16])
{
    /* transform_vec6d_to_mat4f (eigen_solver.hpp:30-40): R = Rz(xi2) Ry(xi1) Rx(xi0) */
#if ORC_PLANT == 3 /* angles applied in the wrong order: Rx(xi2) Ry(xi1) Rz(xi0) */
    double ca = cos(xi[2]), sa = sin(xi[2]);
    double cb = cos(xi[1]), sb = sin(xi[1]);
    double cc = cos(xi[0]), sc = sin(xi[0]);
#else
    double ca = cos(xi[0]), sa = sin(xi[0]);
    double cb = cos(xi[1]), sb = sin(xi[1]);
    double cc = cos(xi[2]), sc = sin(xi[2]);
#endif
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

int orc_solve6(const double jtj21[21], const double jtr[6], double xi[6], double T_step[16])
{
    double A[6][6], L[6][6], D[6];
    int k = 0;
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) {
            A[a][b] = jtj21[k];
            A[b][a] = jtj21[k];
            ++k;
        }
    int ok = 1;
    for (int j = 0; j < 6 && ok; ++j) {
        double dj = A[j][j];
        for (int m = 0; m < j; ++m) dj -= L[j][m] * L[j][m] * D[m];
        if (!(dj > 0.0) || !isfinite(dj)) {
            ok = 0;
            break;
        }
        D[j] = dj;
        L[j][j] = 1.0;
        for (int i = j + 1; i < 6; ++i) {
            double v = A[i][j];
            for (int m = 0; m < j; ++m) v -= L[i][m] * L[j][m] * D[m];
            L[i][j] = v / dj;
        }
    }
    if (ok) {
        double y[6];
        for (int i = 0; i < 6; ++i) {
            double v = -jtr[i];
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
            if (!isfinite(xi[i])) ok = 0;
    }
    if (!ok)
        for (int i = 0; i < 6; ++i) xi[i] = 0.0;
    orc_euler_to_T(xi, T_step);
    return ok ? 0 : 4;
}

void orc_icp_align(const float *src_xyzi, int64_t ns, const float *tgt_xyzi, const float *tgt_n4,
                   const float *tgt_g4, int64_t nt, const float tbmin[3], const float tbmax[3],
                   const gca_icp_params *p, const double T0[16], double T_out[16],
                   gca_icp_report *rep)
{
    double T[16];
    memcpy(T, T0, sizeof(T));
    memset(rep, 0, sizeof(*rep));
    orc_grid g;
    orc_grid_build(&g, tgt_xyzi, nt, tbmin, tbmax, p->max_correspondence_distance);
    int32_t *nn = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ns > 0 ? ns : 1));
    cand *buf = NULL;
    int32_t cap = 0;
    double sys[29] = {0};
    for (int it = 0; it < p->max_iterations; ++it) {
        for (int64_t i = 0; i < ns; ++i) {
            double q[3];
            xform(T, src_xyzi + 4 * i, q);
            int32_t k = gather(&g, tgt_xyzi, q, p->max_correspondence_distance, 1, &buf, &cap);
            nn[i] = k ? buf[0].idx : -1;
        }
        orc_build_gn(src_xyzi, ns, tgt_xyzi, tgt_n4, tgt_g4, nn, p->lambda_geometric, T, sys);
        double xi[6], dT[16];
        if (orc_solve6(sys, sys + 21, xi, dT) != 0) rep->solver_failures++;
        orc_matmul4(dT, T, T);
        rep->iterations = it + 1;
        rep->n_correspondences = (int32_t)sys[28];
        double cos_t = 0.5 * (dT[0] + dT[5] + dT[10] - 1.0);
        double tt = (dT[3] * dT[3] + dT[7] * dT[7]) + dT[11] * dT[11];
        if (cos_t >= 1.0 - p->conv_rotation && tt <= p->conv_translation_sq) {
            rep->converged = 1;
            break;
        }
    }
    rep->rmse = ns > 0 ? sys[27] / (double)ns : 0.0;
    rep->fitness = ns > 0 ? sys[28] / (double)ns : 0.0;
    memcpy(T_out, T, sizeof(T));
    free(buf);
    free(nn);
    orc_grid_free(&g);
}

double orc_objective_grad(const float *src_xyzi, int64_t ns, const float *tgt_xyzi,
                          const float *tgt_n4, const float *tgt_g4, const int32_t *nn,
                          float lambda, const double T[16], const double xi[6], double dxi[6],
                          double *dpts4)
{
    double dT[16];
    orc_euler_to_T(xi, dT);
    double ca = cos(xi[0]), sa = sin(xi[0]), cb = cos(xi[1]), sb = sin(xi[1]);
    double cc = cos(xi[2]), sc = sin(xi[2]);
    double RzRy[9] = {cc * cb, -sc, cc * sb, sc * cb, cc, sc * sb, -sb, 0.0, cb};
    double Rx[9] = {1, 0, 0, 0, ca, -sa, 0, sa, ca};
    double M[9]; /* dR * R_T (rotation part of the full pose) */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = (dT[4 * i] * T[j] + dT[4 * i + 1] * T[4 + j]) + dT[4 * i + 2] * T[8 + j];
    double lam = (double)lambda, E = 0.0;
    for (int a = 0; a < 6; ++a) dxi[a] = 0.0;
    for (int64_t i = 0; i < ns; ++i) {
        if (dpts4)
            for (int a = 0; a < 4; ++a) dpts4[4 * i + a] = 0.0;
        int32_t j = nn[i];
        if (j < 0) continue;
        double y[3], z[3], p[3];
        xform(T, src_xyzi + 4 * i, y);
        for (int a = 0; a < 3; ++a) z[a] = (dT[4 * a] * y[0] + dT[4 * a + 1] * y[1]) + dT[4 * a + 2] * y[2];
        for (int a = 0; a < 3; ++a) p[a] = z[a] + dT[4 * a + 3];
        double rg, rc, Jg[6], Jc[6], rgr, rcr, g[3];
        residual_jac(p, tgt_xyzi + 4 * (int64_t)j, tgt_n4 + 4 * (int64_t)j, tgt_g4 + 4 * (int64_t)j,
                     (double)src_xyzi[4 * i + 3], 1.0, 1.0, &rg, &rc, Jg, Jc, &rgr, &rcr, g);
        E += lam * rgr * rgr + (1.0 - lam) * rcr * rcr;
        const float *n4 = tgt_n4 + 4 * (int64_t)j;
        double u[3];
        for (int a = 0; a < 3; ++a) u[a] = 2.0 * lam * rgr * (double)n4[a] + 2.0 * (1.0 - lam) * rcr * g[a];
        double w[3], w2[3], xy[3], c[3];
        for (int a = 0; a < 3; ++a) /* w = dR^T u */
            w[a] = (dT[a] * u[0] + dT[4 + a] * u[1]) + dT[8 + a] * u[2];
        cross3(y, w, c);
        dxi[0] += c[0];
        for (int a = 0; a < 3; ++a) /* w2 = (Rz Ry)^T u */
            w2[a] = (RzRy[a] * u[0] + RzRy[3 + a] * u[1]) + RzRy[6 + a] * u[2];
        for (int a = 0; a < 3; ++a) xy[a] = (Rx[3 * a] * y[0] + Rx[3 * a + 1] * y[1]) + Rx[3 * a + 2] * y[2];
        cross3(xy, w2, c);
        dxi[1] += c[1];
        cross3(z, u, c);
        dxi[2] += c[2];
        for (int a = 0; a < 3; ++a) dxi[3 + a] += u[a];
        if (dpts4) {
            for (int a = 0; a < 3; ++a)
                dpts4[4 * i + a] = (M[a] * u[0] + M[3 + a] * u[1]) + M[6 + a] * u[2];
            dpts4[4 * i + 3] = -2.0 * (1.0 - lam) * rcr;
        }
    }
    return E;
}

int64_t orc_radius_outlier_removal(const float *xyzi, const float *rgb, int64_t n,
                                   const float bmin[3], const float bmax[3], float radius,
                                   int32_t min_nn, float *out_xyzi, float *out_rgb)
{
    if (!(radius > 0.0f)) return -1;
    orc_grid g;
    orc_grid_build(&g, xyzi, n, bmin, bmax, radius);
    cand *buf = NULL;
    int32_t cap = 0;
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
        double q[3] = {xyzi[4 * i], xyzi[4 * i + 1], xyzi[4 * i + 2]};
        int32_t k = gather(&g, xyzi, q, radius, 0, &buf, &cap);
        if (k > min_nn) {
            memcpy(out_xyzi + 4 * m, xyzi + 4 * i, 16);
            memcpy(out_rgb + 4 * m, rgb + 4 * i, 16);
            ++m;
        }
    }
    free(buf);
    orc_grid_free(&g);
    return m;
}

int orc_multiscale_icp(const float *src_xyzi, const float *src_rgb, int64_t ns,
                       const float *tgt_xyzi, const float *tgt_rgb, int64_t nt, int n_levels,
                       const gca_level_params *levels, const double T0[16], double T_out[16],
                       gca_icp_report *reps)
{
    double T[16];
    memcpy(T, T0, sizeof(T));
    float sbmin[3], sbmax[3], tbmin[3], tbmax[3];
    orc_bounds(src_xyzi, ns, sbmin, sbmax);
    orc_bounds(tgt_xyzi, nt, tbmin, tbmax);
    float *sx = (float *)malloc(16 * (size_t)(ns + 1)), *sr = (float *)malloc(16 * (size_t)(ns + 1));
    float *tx = (float *)malloc(16 * (size_t)(nt + 1)), *tr = (float *)malloc(16 * (size_t)(nt + 1));
    float *tn = (float *)malloc(16 * (size_t)(nt + 1)), *tg = (float *)malloc(16 * (size_t)(nt + 1));
    int rc = 0;
    for (int l = 0; l < n_levels; ++l) {
        const gca_level_params *L = &levels[l];
        int64_t ms = orc_voxel_downsample(src_xyzi, src_rgb, ns, sbmin, sbmax, L->voxel_size, sx, sr, NULL);
        int64_t mt = orc_voxel_downsample(tgt_xyzi, tgt_rgb, nt, tbmin, tbmax, L->voxel_size, tx, tr, NULL);
        if (ms < 0 || mt < 0) {
            rc = 1;
            break;
        }
        /* the downsampled target keeps the input's bounds (point_cloud.cu:224-226) */
        orc_estimate_normals(tx, mt, tbmin, tbmax, L->normal_radius, L->normal_max_nn, tn);
        orc_color_gradient(tx, tn, mt, tbmin, tbmax, L->gradient_radius, L->gradient_max_nn, tg);
        orc_icp_align(sx, ms, tx, tn, tg, mt, tbmin, tbmax, &L->icp, T, T, &reps[l]);
    }
    memcpy(T_out, T, sizeof(T));
    free(sx);
    free(sr);
    free(tx);
    free(tr);
    free(tn);
    free(tg);
    return rc;
}

void orc_abi_sizes(int64_t out[6])
{
    out[0] = (int64_t)sizeof(gca_camera);
    out[1] = (int64_t)sizeof(gca_bilateral);
    out[2] = (int64_t)sizeof(gca_icp_params);
    out[3] = (int64_t)sizeof(gca_icp_report);
    out[4] = (int64_t)sizeof(gca_level_params);
    out[5] = (int64_t)sizeof(orc_grid);
}
