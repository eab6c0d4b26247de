// vren/debug_consumers.hpp -- drop-ins for the demo's debug views of the clustered-shading outputs:
// vren_demo::show_clusters (clustered_shading_debug.hpp:17-34, clustered_shading_debug.cpp:52-100) and
// vren_demo::visualize_bvh (visualize_bvh.hpp:10-26, visualize_bvh.cpp:21-90). Both enqueue on a stream and write
// device buffers (an RGBA32F image, a line-vertex list) that can be dumped and compared.
#pragma once

#include "pipeline/clustered_shading.hpp"

namespace vren_demo
{
    class show_clusters
    {
    public:
        // show_clusters.comp:16-21
        enum mode : int32_t
        {
            k_key = 0,
            k_ijk = 1,
            k_normal = 2,
            k_light_assignment_counts = 3,
            k_light_assignment_offsets = 4,
            k_light_assignment_indices = 5,
        };

        explicit show_clusters(vren::context const& context) : m_context(&context) {}

        // reads cluster_and_shade's cluster references, keys and light lists; writes W*H RGBA32F pixels
        void operator()(vren::stream stream, vren::uvec2 const& screen, int32_t mode,
                        vren::cluster_and_shade const& cluster_and_shade, vren::device_buffer const& output) const
        {
            if (output.m_size < (size_t) screen.x * screen.y * 16)
                throw std::invalid_argument("show_clusters: output buffer too small");
            vren_ctx* ctx = m_context->handle();
            vren::check_status(
                ctx,
                vren_show_clusters(ctx, stream.get(), screen.x, screen.y, mode,
                                   cluster_and_shade.m_cluster_reference_buffer.at<uint32_t>(),
                                   cluster_and_shade.m_cluster_key_buffer.at<uint32_t>(),
                                   (uint32_t) (cluster_and_shade.m_cluster_key_buffer.m_size / 4),
                                   cluster_and_shade.m_assigned_light_counts_buffer.at<uint32_t>(),
                                   cluster_and_shade.m_assigned_light_offsets_buffer.at<uint32_t>(),
                                   cluster_and_shade.m_assigned_light_indices_buffer.at<uint32_t>(),
                                   (uint32_t) (cluster_and_shade.m_assigned_light_indices_buffer.m_size / 4),
                                   output.at<float>()),
                "vren_demo::show_clusters");
        }

    private:
        vren::context const* m_context;
    };

    class visualize_bvh
    {
    public:
        explicit visualize_bvh(vren::context const& context) : m_context(&context) {}

        // 24 line vertices of 16 bytes per node of a BVH of level_count + 1 levels
        static size_t get_required_draw_buffer_size(uint32_t level_count)
        {
            size_t n = 0;
            for (uint32_t l = 0; l <= level_count; l++) n += (size_t) 1 << (5 * l);
            return n * 24 * 16;
        }

        void write(vren::stream stream, vren::resource_container& resource_container, vren::device_buffer const& bvh,
                   uint32_t level_count, vren::device_buffer const& draw_buffer) const
        {
            (void) resource_container;
            if (level_count > 5 || draw_buffer.m_size < get_required_draw_buffer_size(level_count) ||
                bvh.m_size < get_required_draw_buffer_size(level_count) / 12)
                throw std::invalid_argument("visualize_bvh: level_count above 5 or a buffer too small");
            vren_ctx* ctx = m_context->handle();
            vren::check_status(ctx,
                               vren_visualize_bvh(ctx, stream.get(), bvh.at<vren_bvh_node>(), level_count,
                                                  draw_buffer.at<float>()),
                               "vren_demo::visualize_bvh");
        }

    private:
        vren::context const* m_context;
    };
} // namespace vren_demo
