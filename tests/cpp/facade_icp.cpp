// Registers two clouds through the C++ facade (include/gca/gca.hpp), the way a port of the
// reference's apps/test_color_icp.cpp:378-414 would: create clouds, estimate the target's
// normals, color_icp::align.  Input: a binary file
//   int64 n_src, int64 n_tgt, float32 src_xyz[n_src*3], src_rgb[n_src*3],
//   float32 tgt_xyz[n_tgt*3], tgt_rgb[n_tgt*3]
// Output: one JSON line with the pose (row-major), iterations, the reference "RMSE", the
// per-point absolute residuals at the final pose (color_icp::get_abs_residual_host), the target's
// Euclidean clusters and the moved source's residual segments ([size, first index] each).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/gca/gca.hpp"

using pts = std::vector<std::array<float, 3>>;

static bool read_cloud(FILE *f, int64_t n, pts &xyz, pts &rgb)
{
    xyz.resize((size_t)n);
    rgb.resize((size_t)n);
    if (n == 0) return true;
    return fread(xyz[0].data(), sizeof(float) * 3, (size_t)n, f) == (size_t)n &&
           fread(rgb[0].data(), sizeof(float) * 3, (size_t)n, f) == (size_t)n;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s clouds.bin [normal_radius max_nn corr_dist grad_radius iters]\n", argv[0]);
        return 2;
    }
    const float nr = argc > 2 ? (float)atof(argv[2]) : 0.05f;
    const int nk = argc > 3 ? atoi(argv[3]) : 30;
    const float corr = argc > 4 ? (float)atof(argv[4]) : 0.05f;
    const float gr = argc > 5 ? (float)atof(argv[5]) : 0.04f;
    const int iters = argc > 6 ? atoi(argv[6]) : 30;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t ns = 0, nt = 0;
    pts sx, sr, tx, tr;
    const bool ok = fread(&ns, 8, 1, f) == 1 && fread(&nt, 8, 1, f) == 1 && read_cloud(f, ns, sx, sr) &&
                    read_cloud(f, nt, tx, tr);
    fclose(f);
    if (!ok) return 2;

    auto src = gca::point_cloud::create_from_host(sx, sr);
    auto tgt = gca::point_cloud::create_from_host(tx, tr);
    if (!src || !tgt || !tgt->estimate_normals(nr, nk)) {
        fprintf(stderr, "setup failed: %s\n", gca_status_string(gca::last_status()));
        return 1;
    }
    gca::color_icp icp((size_t)iters, corr, gr);
    icp.set_source_point_cloud(src);
    icp.set_target_point_cloud(tgt);
    if (!icp.align()) {
        fprintf(stderr, "align failed: %s\n", gca_status_string(gca::last_status()));
        return 1;
    }
    const auto &T = icp.get_final_transformation_matrix();
    const auto res = icp.get_abs_residual_host();
    const auto clusters = tgt->euclidean_clustering(0.02f, 20, 2000);
    const auto moved = icp.get_transformed_source_point_cloud();
    const auto segments = moved->icp_residual_segmentation(0.02f, res.first, 10, 1500);
    printf("{\"T\": [");
    for (int k = 0; k < 16; ++k) printf("%s%.17g", k ? ", " : "", T[k]);
    printf("], \"iterations\": %d, \"rmse\": %.17g, \"n_src\": %zu, \"moved_src\": %zu", icp.report().iterations,
           (double)icp.get_RSME(), src->points_number(), moved->points_number());
    printf(", \"abs_mean\": %.9g, \"abs\": [", (double)res.second);
    for (size_t k = 0; k < res.first.size(); ++k) printf("%s%.9g", k ? ", " : "", (double)res.first[k]);
    printf("], \"clusters\": [");
    for (size_t k = 0; k < clusters.size(); ++k)
        printf("%s[%zu, %d]", k ? ", " : "", clusters[k].size(), clusters[k].empty() ? -1 : clusters[k][0]);
    printf("], \"segments\": [");
    for (size_t k = 0; k < segments.size(); ++k)
        printf("%s[%zu, %d]", k ? ", " : "", segments[k].size(), segments[k].empty() ? -1 : segments[k][0]);
    printf("]}\n");
    return 0;
}
