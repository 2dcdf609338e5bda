// cpu_pipeline.cpp -- the benchmark's frame step as a multithreaded FP32 C++ program on the host
// cores: the CPU baseline that bench.py times beside the GPU.  TEST INFRASTRUCTURE ONLY (like the
// rest of oracle/): only bench.py's cpu_baseline leg and tests/ load it; the product never does.
//
// A step is what one frame costs the GPU frame pipeline: the new frame's front end (5x5
// bilateral, back-projection, colour, stable compaction, bounds), then for each pyramid level its
// voxel downsample, neighbour grid, PCA normals (K nearest within the radius) and colour gradient,
// and the colored-ICP registration of the new frame onto the previous one, coarse to fine, with
// the reference's stopping rule.  The algorithms are the reference's (the functions the oracle
// restates, cited there): cuda_point_cloud_factory.cu:64-267 (front end),
// cuda_voxel_grid_down_sample.cu:123-161 (voxel means), cuda_nn_search.cu:545-723 (27-cell grid
// search), cuda_estimate_normals.cu:82-151 (normals), cuda_compute_color_gradient.cu:13-90
// (gradient), cuda_color_icp_build_least_square.cu:134-266 (Gauss-Newton terms),
// eigen_solver.hpp:30-85 (solve, Euler increment) and color_icp.cu:151-234 (loop, stop rule).
// Arithmetic is FP32 except the per-thread Gauss-Newton sums, the 6x6 solve and the pose (FP64),
// and the data structures are the plain CPU ones: an LSD radix sort of packed voxel keys, a dense
// counting-sort grid, and per-thread candidate buffers.  Loops over pixels, points and voxels are
// OpenMP-parallel.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

// Workload parameters (mirrored by oracle.cpu_pipeline_params in Python).
struct cpb_params {
    int W, H;
    float fx, fy, cx, cy, scale, tmin, tmax;
    int radius;
    float sigma_space, sigma_depth;
    int n_levels;
    float voxel[8];
    int max_iter;
    float lambda;
    double conv_rot, conv_t2;
};

namespace {

using Params = cpb_params;

struct Cloud {
    std::vector<float> x, y, z, I, r, g, b;  // SoA
    std::vector<float> nx, ny, nz, gx, gy, gz;
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    size_t size() const { return x.size(); }
    void resize(size_t n)
    {
        for (auto *v : {&x, &y, &z, &I, &r, &g, &b}) v->resize(n);
    }
};

struct Grid {
    float origin[3] = {0, 0, 0};
    float cell = 1.f;
    int dims[3] = {1, 1, 1};
    std::vector<int32_t> start;  // ncells + 1
    std::vector<int32_t> order;
    std::vector<float> sx, sy, sz;  // points in cell order
};

float intensity_of(float r, float g, float b) { return (float)(0.2126 * (double)r + 0.7152 * (double)g + 0.0722 * (double)b); }

// ---- front end ------------------------------------------------------------------------------
void front_end(const uint16_t *depth, const uint8_t *bgr, const Params &p, Cloud &out)
{
    const int W = p.W, H = p.H, R = p.radius;
    std::vector<float> ws((2 * R + 1) * (2 * R + 1));
    for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx)
            ws[(dy + R) * (2 * R + 1) + dx + R] = expf(-(float)(dx * dx + dy * dy) / (p.sigma_space * p.sigma_space));
    const float inv_sd2 = 1.f / (p.sigma_depth * p.sigma_depth), inv255 = 1.f / 255.f;
    std::vector<int32_t> row_cnt(H + 1, 0);
    std::vector<float> px((size_t)W * H), py((size_t)W * H), pz((size_t)W * H), pr((size_t)W * H), pg((size_t)W * H),
        pb((size_t)W * H);
    std::vector<uint8_t> ok((size_t)W * H);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        int cnt = 0;
        for (int x = 0; x < W; ++x) {
            const size_t k = (size_t)y * W + x;
            ok[k] = 0;
            const uint16_t d0 = depth[k];
            if (!d0) continue;
            const float z0 = (float)d0 * p.scale;
            if (z0 < p.tmin || z0 > p.tmax) continue;
            float sw = 0.f, swd = 0.f;
            for (int dy = -R; dy <= R; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= H) continue;
                for (int dx = -R; dx <= R; ++dx) {
                    const int xx = x + dx;
                    if (xx < 0 || xx >= W) continue;
                    const uint16_t dn = depth[(size_t)yy * W + xx];
                    if (!dn) continue;
                    const float df = (float)((int)d0 - (int)dn);
                    const float w = ws[(dy + R) * (2 * R + 1) + dx + R] * expf(-df * df * inv_sd2);
                    sw += w;
                    swd += w * (float)dn;
                }
            }
            const float z = (swd / sw) * p.scale;
            if (z < p.tmin || z > p.tmax) continue;
            const float X = (((float)x - 0.5f - p.cx) * z) / p.fx, Y = (((float)y - 0.5f - p.cy) * z) / p.fy;
            const float fu = ((X * p.fx) / z) + p.cx + 0.5f, fv = ((Y * p.fy) / z) + p.cy + 0.5f;
            if (!(fu > -1.f && fu < (float)W && fv > -1.f && fv < (float)H)) continue;
            const uint8_t *c = bgr + 3 * ((size_t)(int)fv * W + (int)fu);
            px[k] = X;
            py[k] = Y;
            pz[k] = z;
            pb[k] = c[0] * inv255;
            pg[k] = c[1] * inv255;
            pr[k] = c[2] * inv255;
            ok[k] = 1;
            ++cnt;
        }
        row_cnt[y + 1] = cnt;
    }
    for (int y = 0; y < H; ++y) row_cnt[y + 1] += row_cnt[y];
    out.resize((size_t)row_cnt[H]);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        size_t o = (size_t)row_cnt[y];
        for (int x = 0; x < W; ++x) {
            const size_t k = (size_t)y * W + x;
            if (!ok[k]) continue;
            out.x[o] = px[k];
            out.y[o] = py[k];
            out.z[o] = pz[k];
            out.r[o] = pr[k];
            out.g[o] = pg[k];
            out.b[o] = pb[k];
            out.I[o] = intensity_of(pr[k], pg[k], pb[k]);
            ++o;
        }
    }
}

void bounds(Cloud &c)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    const int64_t n = (int64_t)c.size();
#pragma omp parallel
    {
        float a[3] = {INFINITY, INFINITY, INFINITY}, b[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma omp for schedule(static) nowait
        for (int64_t i = 0; i < n; ++i) {
            const float v[3] = {c.x[i], c.y[i], c.z[i]};
            for (int k = 0; k < 3; ++k) {
                a[k] = std::min(a[k], v[k]);
                b[k] = std::max(b[k], v[k]);
            }
        }
#pragma omp critical
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], a[k]);
            mx[k] = std::max(mx[k], b[k]);
        }
    }
    memcpy(c.mn, mn, sizeof(mn));
    memcpy(c.mx, mx, sizeof(mx));
}

// ---- voxel downsample: packed 63-bit keys, parallel LSD radix sort, segment means -------------
void radix_sort(std::vector<uint64_t> &key, std::vector<int32_t> &val, int bits)
{
    const size_t n = key.size();
    std::vector<uint64_t> k2(n);
    std::vector<int32_t> v2(n);
    const int nt = omp_get_max_threads();
    for (int shift = 0; shift < bits; shift += 11) {
        std::vector<size_t> hist((size_t)nt * 2048, 0);
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const size_t lo = n * t / nt, hi = n * (t + 1) / nt;
            size_t *h = &hist[(size_t)t * 2048];
            for (size_t i = lo; i < hi; ++i) ++h[(key[i] >> shift) & 2047];
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < 2048; ++d)
                    for (int u = 0; u < nt; ++u) {
                        const size_t c = hist[(size_t)u * 2048 + d];
                        hist[(size_t)u * 2048 + d] = run;
                        run += c;
                    }
            }
            for (size_t i = lo; i < hi; ++i) {
                const size_t o = h[(key[i] >> shift) & 2047]++;
                k2[o] = key[i];
                v2[o] = val[i];
            }
        }
        key.swap(k2);
        val.swap(v2);
    }
}

void voxel_downsample(const Cloud &in, float v, Cloud &out)
{
    const int64_t n = (int64_t)in.size();
    int64_t dims[3];
    for (int a = 0; a < 3; ++a) dims[a] = (int64_t)std::floor((in.mx[a] - in.mn[a]) / v) + 1;
    int bits[3];
    for (int a = 0; a < 3; ++a) {
        bits[a] = 1;
        while ((int64_t(1) << bits[a]) < dims[a]) ++bits[a];
    }
    std::vector<uint64_t> key((size_t)n);
    std::vector<int32_t> idx((size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t kx = (uint64_t)(int64_t)std::floor((in.x[i] - in.mn[0]) / v);
        const uint64_t ky = (uint64_t)(int64_t)std::floor((in.y[i] - in.mn[1]) / v);
        const uint64_t kz = (uint64_t)(int64_t)std::floor((in.z[i] - in.mn[2]) / v);
        key[i] = (kx << (bits[1] + bits[2])) | (ky << bits[2]) | kz;  // lexicographic (kx, ky, kz)
        idx[i] = (int32_t)i;
    }
    radix_sort(key, idx, bits[0] + bits[1] + bits[2]);
    std::vector<int64_t> seg;
    seg.reserve((size_t)n / 4 + 2);
    for (int64_t i = 0; i < n; ++i)
        if (i == 0 || key[i] != key[i - 1]) seg.push_back(i);
    seg.push_back(n);
    const int64_t m = (int64_t)seg.size() - 1;
    out.resize((size_t)m);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < m; ++s) {
        float acc[6] = {0, 0, 0, 0, 0, 0};
        for (int64_t k = seg[s]; k < seg[s + 1]; ++k) {
            const int32_t i = idx[k];
            acc[0] += in.x[i];
            acc[1] += in.y[i];
            acc[2] += in.z[i];
            acc[3] += in.r[i];
            acc[4] += in.g[i];
            acc[5] += in.b[i];
        }
        const float c = (float)(seg[s + 1] - seg[s]);
        out.x[s] = acc[0] / c;
        out.y[s] = acc[1] / c;
        out.z[s] = acc[2] / c;
        out.r[s] = acc[3] / c;
        out.g[s] = acc[4] / c;
        out.b[s] = acc[5] / c;
        out.I[s] = intensity_of(out.r[s], out.g[s], out.b[s]);
    }
    memcpy(out.mn, in.mn, sizeof(out.mn));  // the output keeps the input's bounds (point_cloud.cu:224-226)
    memcpy(out.mx, in.mx, sizeof(out.mx));
}

// ---- dense grid (cell = radius, origin = min bound padded by 1.5 r) ---------------------------
void grid_build(const Cloud &c, float r, Grid &G)
{
    const float pad = 1.5f * r;
    G.cell = r;
    for (int a = 0; a < 3; ++a) {
        G.origin[a] = c.mn[a] - pad;
        G.dims[a] = (int)((c.mx[a] + pad - G.origin[a]) / r) + 1;
    }
    const int64_t ncell = (int64_t)G.dims[0] * G.dims[1] * G.dims[2];
    const int64_t n = (int64_t)c.size();
    std::vector<int32_t> cell((size_t)n);
    G.start.assign((size_t)ncell + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int q[3];
        const float p[3] = {c.x[i], c.y[i], c.z[i]};
        for (int a = 0; a < 3; ++a) q[a] = std::min(std::max((int)((p[a] - G.origin[a]) / r), 0), G.dims[a] - 1);
        cell[i] = (q[0] * G.dims[1] + q[1]) * G.dims[2] + q[2];
    }
    for (int64_t i = 0; i < n; ++i) ++G.start[(size_t)cell[i] + 1];
    for (int64_t k = 0; k < ncell; ++k) G.start[k + 1] += G.start[k];
    std::vector<int32_t> fill(G.start.begin(), G.start.end() - 1);
    G.order.resize((size_t)n);
    G.sx.resize((size_t)n);
    G.sy.resize((size_t)n);
    G.sz.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t s = fill[cell[i]]++;
        G.order[s] = (int32_t)i;
        G.sx[s] = c.x[i];
        G.sy[s] = c.y[i];
        G.sz[s] = c.z[i];
    }
}

// neighbours of q within r (d2 < r^2): calls f(slot, d2)
template <class F>
inline void for_neighbours(const Grid &G, float qx, float qy, float qz, float r, F &&f)
{
    const float q[3] = {qx, qy, qz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = std::max((int)std::floor((q[a] - r - G.origin[a]) / G.cell), 0);
        hi[a] = std::min((int)std::floor((q[a] + r - G.origin[a]) / G.cell), G.dims[a] - 1);
        if (lo[a] > hi[a]) return;
    }
    const float r2 = r * r;
    for (int ix = lo[0]; ix <= hi[0]; ++ix)
        for (int iy = lo[1]; iy <= hi[1]; ++iy) {
            const int64_t row = ((int64_t)ix * G.dims[1] + iy) * G.dims[2];
            const int32_t s1 = G.start[row + hi[2] + 1];
            for (int32_t s = G.start[row + lo[2]]; s < s1; ++s) {
                const float dx = G.sx[s] - qx, dy = G.sy[s] - qy, dz = G.sz[s] - qz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 < r2) f(s, d2);
            }
        }
}

// symmetric 3x3 eigenvector of the smallest eigenvalue (cyclic Jacobi, FP32)
void smallest_eigvec(float A[3][3], float v[3])
{
    float V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 8; ++sweep) {
        const float off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        if (off < 1e-30f) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (std::fabs(A[p][q]) < 1e-30f) continue;
                const float th = (A[q][q] - A[p][p]) / (2.f * A[p][q]);
                const float t = (th >= 0 ? 1.f : -1.f) / (std::fabs(th) + std::sqrt(th * th + 1.f));
                const float c = 1.f / std::sqrt(t * t + 1.f), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const float akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const float apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const float vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int m = 0;
    for (int k = 1; k < 3; ++k)
        if (A[k][k] < A[m][m]) m = k;
    for (int k = 0; k < 3; ++k) v[k] = V[k][m];
}

struct Cand {
    float d2;
    int32_t idx;
};

void normals_and_gradient(Cloud &c, const Grid &G, float rn, int K, float rg)
{
    const int64_t n = (int64_t)c.size();
    for (auto *v : {&c.nx, &c.ny, &c.nz, &c.gx, &c.gy, &c.gz}) v->assign((size_t)n, 0.f);
#pragma omp parallel
    {
        std::vector<Cand> cand;
        cand.reserve(256);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            const float qx = c.x[i], qy = c.y[i], qz = c.z[i];
            cand.clear();
            for_neighbours(G, qx, qy, qz, rn, [&](int32_t s, float d2) { cand.push_back({d2, G.order[s]}); });
            size_t k = cand.size();
            if (K > 0 && k > (size_t)K) {
                std::nth_element(cand.begin(), cand.begin() + K, cand.end(), [](const Cand &a, const Cand &b) {
                    return a.d2 < b.d2 || (a.d2 == b.d2 && a.idx < b.idx);
                });
                k = (size_t)K;
            }
            float nrm[3] = {0, 0, 0};
            if (k >= 3) {
                float m[3] = {0, 0, 0};
                for (size_t t = 0; t < k; ++t) {
                    const int32_t j = cand[t].idx;
                    m[0] += c.x[j] - qx;
                    m[1] += c.y[j] - qy;
                    m[2] += c.z[j] - qz;
                }
                for (float &v : m) v /= (float)k;
                float A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (size_t t = 0; t < k; ++t) {
                    const int32_t j = cand[t].idx;
                    const float d[3] = {c.x[j] - qx - m[0], c.y[j] - qy - m[1], c.z[j] - qz - m[2]};
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) A[a][b] += d[a] * d[b];
                }
                smallest_eigvec(A, nrm);
                if (nrm[2] > 0.f)  // orient towards (0, 0, -1) (cuda_estimate_normals.cu:146-149)
                    for (float &v : nrm) v = -v;
            }
            c.nx[i] = nrm[0];
            c.ny[i] = nrm[1];
            c.nz[i] = nrm[2];
        }
        // colour gradient over every neighbour within rg (after all normals exist)
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            const float qx = c.x[i], qy = c.y[i], qz = c.z[i], Iq = c.I[i];
            const float nv[3] = {c.nx[i], c.ny[i], c.nz[i]};
            float A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, bb[3] = {0, 0, 0};
            int k = 0;
            for_neighbours(G, qx, qy, qz, rg, [&](int32_t s, float) {
                const int32_t j = G.order[s];
                ++k;
                if (j == i) return;
                const float e[3] = {c.x[j] - qx, c.y[j] - qy, c.z[j] - qz};
                const float en = e[0] * nv[0] + e[1] * nv[1] + e[2] * nv[2];
                const float v[3] = {e[0] - en * nv[0], e[1] - en * nv[1], e[2] - en * nv[2]};
                const float dI = c.I[j] - Iq;
                for (int a = 0; a < 3; ++a) {
                    bb[a] += v[a] * dI;
                    for (int b = 0; b < 3; ++b) A[a][b] += v[a] * v[b];
                }
            });
            float gout[3] = {0, 0, 0};
            if (k >= 5) {
                const float w = (float)((k - 1) * (k - 1));
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) A[a][b] += w * nv[a] * nv[b] + (a == b ? 1e-6f : 0.f);
                // unpivoted LDL^T (little_3x3_ldlt_solve.cuh:11-60)
                const float D0 = A[0][0], L10 = A[1][0] / D0, L20 = A[2][0] / D0;
                const float D1 = A[1][1] - L10 * L10 * D0, L21 = (A[2][1] - L10 * L20 * D0) / D1;
                const float D2 = A[2][2] - L20 * L20 * D0 - L21 * L21 * D1;
                float x0 = bb[0], x1 = bb[1] - L10 * x0, x2 = bb[2] - L20 * x0 - L21 * x1;
                x0 /= D0;
                x1 /= D1;
                x2 /= D2;
                x1 -= L21 * x2;
                x0 -= L10 * x1 + L20 * x2;
                gout[0] = x0;
                gout[1] = x1;
                gout[2] = x2;
            }
            c.gx[i] = gout[0];
            c.gy[i] = gout[1];
            c.gz[i] = gout[2];
        }
    }
}

// ---- colored ICP ----------------------------------------------------------------------------
bool solve6(const double A21[21], const double b[6], double xi[6])
{
    double A[6][6], L[6][6] = {{0}}, D[6];
    int k = 0;
    for (int a = 0; a < 6; ++a)
        for (int c = a; c < 6; ++c) A[a][c] = A[c][a] = A21[k++];
    for (int j = 0; j < 6; ++j) {
        double dj = A[j][j];
        for (int m = 0; m < j; ++m) dj -= L[j][m] * L[j][m] * D[m];
        if (!(dj > 0.0)) return false;
        D[j] = dj;
        for (int i = j + 1; i < 6; ++i) {
            double v = A[i][j];
            for (int m = 0; m < j; ++m) v -= L[i][m] * L[j][m] * D[m];
            L[i][j] = v / dj;
        }
    }
    double y[6];
    for (int i = 0; i < 6; ++i) {  // L z = -b, then y = z / D
        double v = -b[i];
        for (int m = 0; m < i; ++m) v -= L[i][m] * y[m];
        y[i] = v;
    }
    for (int i = 0; i < 6; ++i) y[i] /= D[i];
    for (int i = 5; i >= 0; --i) {
        double v = y[i];
        for (int m = i + 1; m < 6; ++m) v -= L[m][i] * xi[m];
        xi[i] = v;
    }
    return true;
}

void euler_to_T(const double xi[6], double T[16])
{
    const double ca = cos(xi[0]), sa = sin(xi[0]), cb = cos(xi[1]), sb = sin(xi[1]), cc = cos(xi[2]), sc = sin(xi[2]);
    const double M[16] = {cc * cb, cc * sb * sa - sc * ca, cc * sb * ca + sc * sa, xi[3],
                          sc * cb, sc * sb * sa + cc * ca, sc * sb * ca - cc * sa, xi[4],
                          -sb,     cb * sa,                cb * ca,                xi[5],
                          0,       0,                      0,                      1};
    memcpy(T, M, sizeof(M));
}

void matmul4(const double A[16], const double B[16], double C[16])
{
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            t[4 * i + j] = A[4 * i] * B[j] + A[4 * i + 1] * B[4 + j] + A[4 * i + 2] * B[8 + j] + A[4 * i + 3] * B[12 + j];
    memcpy(C, t, sizeof(t));
}

int icp(const Cloud &src, const Cloud &tgt, const Grid &G, float corr, const Params &p, double T[16])
{
    const int64_t n = (int64_t)src.size();
    const float sl = std::sqrt(p.lambda), sc = std::sqrt(1.f - p.lambda);
    int it = 0;
    for (; it < p.max_iter; ++it) {
        const float Tf[12] = {(float)T[0], (float)T[1], (float)T[2],  (float)T[3], (float)T[4],  (float)T[5],
                              (float)T[6], (float)T[7], (float)T[8], (float)T[9], (float)T[10], (float)T[11]};
        double sys[28] = {0};
#pragma omp parallel
        {
            double acc[28] = {0};
#pragma omp for schedule(dynamic, 512) nowait
            for (int64_t i = 0; i < n; ++i) {
                const float x = Tf[0] * src.x[i] + Tf[1] * src.y[i] + Tf[2] * src.z[i] + Tf[3];
                const float y = Tf[4] * src.x[i] + Tf[5] * src.y[i] + Tf[6] * src.z[i] + Tf[7];
                const float z = Tf[8] * src.x[i] + Tf[9] * src.y[i] + Tf[10] * src.z[i] + Tf[11];
                float bd = INFINITY;
                int32_t bj = -1;
                for_neighbours(G, x, y, z, corr, [&](int32_t s, float d2) {
                    const int32_t j = G.order[s];
                    if (d2 < bd || (d2 == bd && j < bj)) {
                        bd = d2;
                        bj = j;
                    }
                });
                if (bj < 0) continue;
                const float qx = tgt.x[bj], qy = tgt.y[bj], qz = tgt.z[bj];
                const float nx = tgt.nx[bj], ny = tgt.ny[bj], nz = tgt.nz[bj];
                const float dx = tgt.gx[bj], dy = tgt.gy[bj], dz = tgt.gz[bj];
                const float rg = (x - qx) * nx + (y - qy) * ny + (z - qz) * nz;
                const float dn = dx * nx + dy * ny + dz * nz;
                const float gx = dx - dn * nx, gy = dy - dn * ny, gz = dz - dn * nz;
                const float Iproj = tgt.I[bj] + dx * (x - rg * nx - qx) + dy * (y - rg * ny - qy) + dz * (z - rg * nz - qz);
                const float J[2][6] = {{sl * (y * nz - z * ny), sl * (z * nx - x * nz), sl * (x * ny - y * nx), sl * nx,
                                        sl * ny, sl * nz},
                                       {sc * (y * gz - z * gy), sc * (z * gx - x * gz), sc * (x * gy - y * gx), sc * gx,
                                        sc * gy, sc * gz}};
                const float r[2] = {sl * rg, sc * (Iproj - src.I[i])};
                int k = 0;
                for (int a = 0; a < 6; ++a)
                    for (int b = a; b < 6; ++b) acc[k++] += J[0][a] * J[0][b] + J[1][a] * J[1][b];
                for (int a = 0; a < 6; ++a) acc[21 + a] += J[0][a] * r[0] + J[1][a] * r[1];
                acc[27] += r[0] * r[0] + r[1] * r[1];
            }
#pragma omp critical
            for (int k = 0; k < 28; ++k) sys[k] += acc[k];
        }
        double xi[6] = {0, 0, 0, 0, 0, 0}, dT[16];
        if (!solve6(sys, sys + 21, xi)) memset(xi, 0, sizeof(xi));
        euler_to_T(xi, dT);
        matmul4(dT, T, T);
        const double cos_t = 0.5 * (dT[0] + dT[5] + dT[10] - 1.0);
        const double tt = dT[3] * dT[3] + dT[7] * dT[7] + dT[11] * dT[11];
        if (cos_t >= 1.0 - p.conv_rot && tt <= p.conv_t2) {
            ++it;
            break;
        }
    }
    return it;
}

struct Frame {
    Cloud full;
    Cloud lv[8];
    Grid grid[8];
};

// One frame step: everything the new frame needs, then its registration onto prev (if any).
void step(const uint16_t *depth, const uint8_t *bgr, const Params &p, Frame &f, const Frame *prev, double T[16],
          int *iters)
{
    front_end(depth, bgr, p, f.full);
    bounds(f.full);
    for (int l = 0; l < p.n_levels; ++l) {
        const float v = p.voxel[l];
        voxel_downsample(f.full, v, f.lv[l]);
        grid_build(f.lv[l], 2.5f * v, f.grid[l]);
        normals_and_gradient(f.lv[l], f.grid[l], 2.5f * v, 30, 2.0f * v);
    }
    if (!prev) return;
    static const double I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    memcpy(T, I, sizeof(I));
    for (int l = 0; l < p.n_levels; ++l) {
        const int it = icp(f.lv[l], prev->lv[l], prev->grid[l], 2.5f * p.voxel[l], p, T);
        if (iters) iters[l] = it;
    }
}

}  // namespace

extern "C" {

// Registers frame B onto frame A (each depth uint16 W*H, BGR8 W*H*3) `warmup` + `reps` times
// with `threads` OpenMP threads, after preparing frame A once (untimed).  ms[reps]: each timed
// step (frame B's front end, levels and registration); T_out[16]: the last pose; iters[n_levels].
// Returns the number of threads used.
int cpb_run(const uint16_t *depth_a, const uint8_t *bgr_a, const uint16_t *depth_b, const uint8_t *bgr_b,
            const cpb_params *p, int threads, int warmup, int reps, double *ms, double *T_out, int *iters, int64_t *points)
{
    if (threads > 0) omp_set_num_threads(threads);
    const int used = omp_get_max_threads();
    Frame a, b;
    double T[16];
    step(depth_a, bgr_a, *p, a, nullptr, T, nullptr);
    for (int k = 0; k < warmup + reps; ++k) {
        const auto t0 = std::chrono::steady_clock::now();
        step(depth_b, bgr_b, *p, b, &a, T, iters);
        const auto t1 = std::chrono::steady_clock::now();
        if (k >= warmup) ms[k - warmup] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    memcpy(T_out, T, sizeof(T));
    if (points) {
        points[0] = (int64_t)b.full.size();
        for (int l = 0; l < p->n_levels; ++l) points[1 + l] = (int64_t)b.lv[l].size();
    }
    return used;
}

}  // extern "C"
