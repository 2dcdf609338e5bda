/*
 * gca.hpp -- header-only C++ facade over the C ABI (gca.h), mirroring the reference's host
 * classes so that their call sites port by changing includes:
 *
 *   gca::point_cloud  <- gca::point_cloud   (src/geometry/point_cloud.hpp:22-131)
 *   gca::color_icp    <- gca::color_icp     (src/registration/color_icp.hpp:16-63)
 *
 * Error behaviour follows the reference (no exceptions by design, point_cloud.cu /
 * color_icp.cu): factories return nullptr, predicates return false, and the status of the
 * last failing call is kept in gca::last_status().  A context (device + HIP stream + scratch)
 * replaces the reference's implicit CUDA context; gca::context::default_context() is the one
 * every object uses unless given another.
 */
#ifndef GCA_GCA_HPP
#define GCA_GCA_HPP

#include <array>
#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "gca.h"

namespace gca
{

inline int32_t &last_status()
{
    static thread_local int32_t s = GCA_OK;
    return s;
}

inline bool ok(int32_t s)
{
    if (s != GCA_OK) last_status() = s;
    return s == GCA_OK;
}

using mat4x4d = std::array<double, 16>;  // row-major rigid transform

inline mat4x4d identity4()
{
    return {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
}

class context
{
public:
    explicit context(int32_t device = 0) { ok(gca_ctx_create(device, &m_ctx)); }
    ~context()
    {
        if (m_ctx) gca_ctx_destroy(m_ctx);
    }
    context(const context &) = delete;
    context &operator=(const context &) = delete;

    bool valid() const { return m_ctx != nullptr; }
    gca_ctx handle() const { return m_ctx; }
    bool synchronize() const { return ok(gca_ctx_synchronize(m_ctx)); }
    void *stream() const
    {
        void *s = nullptr;
        ok(gca_ctx_get_stream(m_ctx, &s));
        return s;
    }
    bool set_tuning(int32_t key, int32_t value) { return ok(gca_ctx_set_tuning(m_ctx, key, value)); }

    static std::shared_ptr<context> default_context()
    {
        static std::shared_ptr<context> c = std::make_shared<context>(0);
        return c;
    }

private:
    gca_ctx m_ctx = nullptr;
};

class point_cloud
{
public:
    explicit point_cloud(size_t n_points = 1, std::shared_ptr<context> ctx = context::default_context())
        : m_ctx(std::move(ctx))
    {
        ok(gca_cloud_create(m_ctx->handle(), (int64_t)(n_points ? n_points : 1), &m_cloud));
    }
    ~point_cloud()
    {
        if (m_cloud) gca_cloud_destroy(m_cloud);
    }
    point_cloud(const point_cloud &other) : point_cloud(other.points_number(), other.m_ctx)
    {
        ok(gca_cloud_copy(m_ctx->handle(), other.m_cloud, m_cloud));  // point_cloud.cu:29-36
    }
    point_cloud &operator=(const point_cloud &other)
    {
        if (this != &other) ok(gca_cloud_copy(m_ctx->handle(), other.m_cloud, m_cloud));
        return *this;
    }
    point_cloud &operator+=(const point_cloud &other)  // point_cloud.cu:55-73
    {
        ok(gca_cloud_append(m_ctx->handle(), m_cloud, other.m_cloud));
        return *this;
    }

    gca_cloud handle() const { return m_cloud; }
    const std::shared_ptr<context> &ctx() const { return m_ctx; }

    size_t points_number() const
    {
        int64_t n = 0;
        ok(gca_cloud_size(m_cloud, &n));
        return (size_t)n;
    }
    bool has_points() const { return points_number() > 0; }
    bool has_normals() const
    {
        int32_t f = 0;
        ok(gca_cloud_has_normals(m_cloud, &f));
        return f != 0;
    }

    // point_cloud::create_from_rgbd (point_cloud.cu:455-479); depth (Z16) and bgr (BGR8) are
    // device buffers of camera.width * camera.height pixels.
    static std::shared_ptr<point_cloud> create_from_rgbd(const uint16_t *depth_dev, const uint8_t *bgr_dev,
                                                         const gca_camera &camera, float threshold_min_in_meter = 0.0f,
                                                         float threshold_max_in_meter = 10.0f,
                                                         const gca_bilateral *filter = nullptr,
                                                         std::shared_ptr<context> ctx = context::default_context())
    {
        gca_bilateral none{GCA_BILATERAL_NONE, 0, 1.0f, 1.0f};
        auto pc = std::make_shared<point_cloud>((size_t)camera.width * camera.height, ctx);
        if (!ok(gca_cloud_from_rgbd(ctx->handle(), depth_dev, bgr_dev, &camera, filter ? filter : &none,
                                    threshold_min_in_meter, threshold_max_in_meter, pc->m_cloud)))
            return nullptr;
        return pc;
    }

    // point_cloud::create_from_pcl (point_cloud.cu:481-504): host xyz and rgb in [0, 1].
    static std::shared_ptr<point_cloud> create_from_host(const std::vector<std::array<float, 3>> &xyz,
                                                         const std::vector<std::array<float, 3>> &rgb,
                                                         std::shared_ptr<context> ctx = context::default_context())
    {
        if (xyz.size() != rgb.size()) return nullptr;
        auto pc = std::make_shared<point_cloud>(xyz.size(), ctx);
        if (!ok(gca_cloud_upload(pc->m_cloud, xyz.empty() ? nullptr : xyz[0].data(),
                                 rgb.empty() ? nullptr : rgb[0].data(), (int64_t)xyz.size())))
            return nullptr;
        return pc;
    }

    // point_cloud::download (point_cloud.cu:75-83) as SoA host arrays.
    void download(std::vector<std::array<float, 3>> &xyz, std::vector<std::array<float, 3>> &rgb) const
    {
        const size_t n = points_number();
        xyz.resize(n);
        rgb.resize(n);
        if (n) ok(gca_cloud_download(m_cloud, xyz[0].data(), rgb[0].data(), nullptr, nullptr, nullptr));
    }
    std::vector<std::array<float, 3>> download_normals() const  // point_cloud.cu:85-91
    {
        std::vector<std::array<float, 3>> nrm(points_number());
        if (!nrm.empty()) ok(gca_cloud_download(m_cloud, nullptr, nullptr, nullptr, nrm[0].data(), nullptr));
        return nrm;
    }

    std::array<float, 3> get_min_bound() const { return bounds().first; }   // point_cloud.cu:122-126
    std::array<float, 3> get_max_bound() const { return bounds().second; }  // point_cloud.cu:128-132

    // point_cloud::estimate_normals (point_cloud.cu:145-181); max_nn = 0: all within radius
    bool estimate_normals(float search_radius, int32_t max_nn = 0)
    {
        return ok(gca_estimate_normals(m_ctx->handle(), m_cloud, search_radius, max_nn));
    }

    // point_cloud::transform (point_cloud.cu:93-105)
    void transform(const mat4x4d &trans_mat) { ok(gca_transform(m_ctx->handle(), m_cloud, trans_mat.data())); }

    // point_cloud::voxel_grid_down_sample (point_cloud.cu:198-229): nullptr on invalid input
    std::shared_ptr<point_cloud> voxel_grid_down_sample(float voxel_size) const
    {
        auto out = std::make_shared<point_cloud>(1, m_ctx);
        if (!ok(gca_voxel_downsample(m_ctx->handle(), m_cloud, voxel_size, out->m_cloud))) return nullptr;
        return out;
    }

    // point_cloud::radius_outlier_removal (point_cloud.cu:238-282)
    std::shared_ptr<point_cloud> radius_outlier_removal(float radius, int32_t min_neighbors_in_radius) const
    {
        auto out = std::make_shared<point_cloud>(1, m_ctx);
        if (!ok(gca_radius_outlier_removal(m_ctx->handle(), m_cloud, radius, min_neighbors_in_radius, out->m_cloud)))
            return nullptr;
        return out;
    }

    // point_cloud::nn_search (point_cloud.cu:506-557): index into reference_pc or -1, per query
    // point (host copy; empty on invalid input, point_cloud.cu:512-516).
    static std::vector<int32_t> nn_search(const point_cloud &query_pc, const point_cloud &reference_pc, float radius);

    // point_cloud::euclidean_clustering (point_cloud.cu:284-335): clusters in seed order, each as
    // ascending point indices (the reference lists a cluster in its search-queue order); empty
    // where the reference returns nothing.
    std::vector<std::vector<int32_t>> euclidean_clustering(float cluster_tolerance, int32_t min_cluster_size,
                                                           int32_t max_cluster_size) const
    {
        std::vector<int32_t> lab(points_number());
        int32_t nc = 0;
        if (gca_euclidean_clustering_host(m_ctx->handle(), m_cloud, cluster_tolerance, min_cluster_size,
                                          max_cluster_size, lab.empty() ? nullptr : lab.data(),
                                          (int64_t)lab.size(), &nc, nullptr) != GCA_OK)
            return {};
        return clusters_from_labels(lab, nc);
    }

    // point_cloud::icp_residual_segmentation (point_cloud.cu:391-435): segments of the radius
    // graph whose edges join points with absolute residuals less than 0.003 apart, as
    // euclidean_clustering returns clusters; abs_residual has one value per point (as from
    // color_icp::get_abs_residual_host).  Empty where the reference returns nothing.
    std::vector<std::vector<int32_t>> icp_residual_segmentation(float cluster_tolerance,
                                                                const std::vector<float> &abs_residual,
                                                                int32_t min_cluster_size,
                                                                int32_t max_cluster_size) const
    {
        std::vector<int32_t> lab(points_number());
        int32_t nc = 0;
        if (abs_residual.size() < lab.size()) return {};
        if (gca_residual_segmentation_host(m_ctx->handle(), m_cloud, cluster_tolerance,
                                           abs_residual.empty() ? nullptr : abs_residual.data(), min_cluster_size,
                                           max_cluster_size, lab.empty() ? nullptr : lab.data(), (int64_t)lab.size(),
                                           &nc, nullptr) != GCA_OK)
            return {};
        return clusters_from_labels(lab, nc);
    }

private:
    static std::vector<std::vector<int32_t>> clusters_from_labels(const std::vector<int32_t> &lab, int32_t nc)
    {
        std::vector<std::vector<int32_t>> out((size_t)nc);
        for (size_t i = 0; i < lab.size(); ++i)
            if (lab[i] >= 0) out[(size_t)lab[i]].push_back((int32_t)i);
        return out;
    }

    std::pair<std::array<float, 3>, std::array<float, 3>> bounds() const
    {
        std::pair<std::array<float, 3>, std::array<float, 3>> b{};
        ok(gca_cloud_bounds(m_cloud, b.first.data(), b.second.data()));
        return b;
    }

    std::shared_ptr<context> m_ctx;
    gca_cloud m_cloud = nullptr;
};

class color_icp
{
public:
    color_icp() = delete;
    // color_icp::color_icp (color_icp.cu:17-25); the stop thresholds are the reference's
    // constants (color_icp.hpp:45-47).
    color_icp(size_t maximum_iterations, float max_correspondence_distance, float search_radius_color_gradient)
        : m_grad_r(search_radius_color_gradient)
    {
        m_params.max_iterations = (int32_t)maximum_iterations;
        m_params.max_correspondence_distance = max_correspondence_distance;
        m_params.lambda_geometric = 0.968f;
        m_params.accumulate_mode = GCA_ACCUM_COMPENSATED;
        m_params.conv_rotation = 1e-6;
        m_params.conv_translation_sq = 1e-6;
    }

    // color_icp.cu:27-48: the source is copied, the target is shared
    void set_source_point_cloud(const std::shared_ptr<const point_cloud> &new_pc)
    {
        m_source = new_pc ? std::make_shared<point_cloud>(*new_pc) : nullptr;
        m_done = false;
    }
    void set_target_point_cloud(const std::shared_ptr<const point_cloud> &new_pc)
    {
        m_target = new_pc;
        m_done = false;
    }

    // color_icp::align (color_icp.cu:151-234); on success the stored source copy is moved by
    // the final transformation, as the reference leaves it.
    bool align()
    {
        if (!m_source || !m_target) return false;
        m_T = identity4();
        mat4x4d T0 = identity4();
        if (!ok(gca_color_icp_align(m_source->ctx()->handle(), m_source->handle(), m_target->handle(), &m_params,
                                    m_grad_r, T0.data(), m_T.data(), &m_report)))
            return false;
        m_source->transform(m_T);
        m_done = true;
        return true;
    }

    // color_icp::get_abs_residual_host (color_icp.cu:90-150): per source point |r_g| + |r_c|
    // (weighted) against its nearest target point within the correspondence distance, the source
    // standing where align left it, and the mean over the source points; empty on invalid state.
    std::pair<std::vector<float>, float> get_abs_residual_host() const
    {
        if (!m_source || !m_target) return {};
        std::vector<float> out(m_source->points_number());
        int64_t n = 0;
        double mean = 0.0;
        if (!ok(gca_abs_residual_host(m_source->ctx()->handle(), m_source->handle(), m_target->handle(),
                                      m_params.max_correspondence_distance, m_params.lambda_geometric, m_grad_r,
                                      out.empty() ? nullptr : out.data(), (int64_t)out.size(), &n, &mean)))
            return {};
        out.resize((size_t)n);
        return {std::move(out), (float)mean};
    }

    const mat4x4d &get_final_transformation_matrix() const { return m_T; }
    float get_RSME() const { return (float)m_report.rmse; }  // (sic) color_icp.hpp:33
    const gca_icp_report &report() const { return m_report; }
    std::shared_ptr<const point_cloud> get_transformed_source_point_cloud() const
    {
        return m_done ? m_source : nullptr;
    }

private:
    gca_icp_params m_params{};
    float m_grad_r;
    std::shared_ptr<point_cloud> m_source;
    std::shared_ptr<const point_cloud> m_target;
    mat4x4d m_T = identity4();
    gca_icp_report m_report{};
    bool m_done = false;
};

inline std::vector<int32_t> point_cloud::nn_search(const point_cloud &query_pc, const point_cloud &reference_pc,
                                                   float radius)
{
    std::vector<int32_t> idx(query_pc.points_number());
    int64_t n = 0;
    if (!ok(gca_nn_search_host(query_pc.ctx()->handle(), query_pc.handle(), reference_pc.handle(), radius,
                               idx.empty() ? nullptr : idx.data(), (int64_t)idx.size(), &n)))
        return {};
    idx.resize((size_t)n);
    return idx;
}

}  // namespace gca

#endif /* GCA_GCA_HPP */
