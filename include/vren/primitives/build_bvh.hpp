// vren/primitives/build_bvh.hpp -- drop-in for build_bvh.hpp:8-56 (build_bvh.cpp:38-136).
#pragma once

#include "../base.hpp"

namespace vren
{
    struct bvh_node
    {
        inline static const uint32_t k_leaf_node = VREN_BVH_LEAF_NODE;
        inline static const uint32_t k_invalid_node = VREN_BVH_INVALID_NODE;

        vren::vec3 m_min;
        uint32_t m_next;
        vren::vec3 m_max;
        uint32_t _pad;

        bool is_leaf() const { return m_next == k_leaf_node; }
        bool is_invalid() const { return m_next == k_invalid_node; }
    };
    static_assert(sizeof(bvh_node) == sizeof(vren_bvh_node) && sizeof(bvh_node) == 32, "32-byte BVH node");

    inline uint32_t calc_bvh_padded_leaf_count(uint32_t leaf_count) { return vren_calc_bvh_padded_leaf_count(leaf_count); }
    inline uint32_t calc_bvh_buffer_length(uint32_t leaf_count) { return vren_calc_bvh_buffer_length(leaf_count); }
    inline size_t calc_bvh_buffer_size(uint32_t leaf_count) { return vren_calc_bvh_buffer_size(leaf_count); }
    inline uint32_t calc_bvh_root_index(uint32_t leaf_count) { return vren_calc_bvh_root_index(leaf_count); }
    inline uint32_t calc_bvh_level_count(uint32_t leaf_count) { return vren_calc_bvh_level_count(leaf_count); }

    class build_bvh
    {
    public:
        explicit build_bvh(vren::context const& context) : m_context(&context) {}

        static size_t get_required_buffer_size(uint32_t leaf_count) { return calc_bvh_buffer_size(leaf_count); }

        // Builds the parent levels over the `leaf_count` (padded: a power of 32, >= 32) leaves at the front of
        // `buffer` (build_bvh.cpp:38-99).
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& buffer, uint32_t leaf_count) const
        {
            (void) resource_container;
            vren_ctx* ctx = m_context->handle();
            check_status(ctx, vren_build_bvh(ctx, stream.get(), buffer.at<vren_bvh_node>(), leaf_count),
                         "vren::build_bvh");
        }

    private:
        vren::context const* m_context;
    };
    // Karras LBVH over sorted (Morton code, light index) pairs (no reference counterpart; a binary BVH
    // of the lights beside the reference's 32-ary one). Node layout: vren_lbvh_node; internal nodes 0 .. n-2, leaf i at n - 1 + i (capi.h, vren_build_lbvh).
    class build_lbvh
    {
    public:
        explicit build_lbvh(vren::context const& context) : m_context(&context) {}

        static size_t get_required_node_buffer_size(uint32_t leaf_count)
        {
            return (size_t) (2 * (size_t) leaf_count - 1) * sizeof(vren_lbvh_node);
        }
        static size_t get_required_parent_buffer_size(uint32_t leaf_count) { return (size_t) (2 * (size_t) leaf_count - 1) * 4; }

        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& point_light_index_buffer, uint32_t leaf_count,
                        vren::device_buffer const& view_space_point_light_position_buffer,
                        vren::device_buffer const& point_light_buffer, vren::device_buffer const& node_buffer,
                        vren::device_buffer const& parent_buffer) const
        {
            (void) resource_container;
            if (leaf_count == 0 || node_buffer.m_size < get_required_node_buffer_size(leaf_count) ||
                parent_buffer.m_size < get_required_parent_buffer_size(leaf_count))
                throw std::invalid_argument("build_lbvh: no leaves or output buffer too small");
            vren_ctx* ctx = m_context->handle();
            check_status(ctx,
                         vren_build_lbvh(ctx, stream.get(), point_light_index_buffer.at<uint32_t>(), leaf_count,
                                         view_space_point_light_position_buffer.at<float>(),
                                         point_light_buffer.at<float>(), node_buffer.at<vren_lbvh_node>(),
                                         parent_buffer.at<uint32_t>()),
                         "vren::build_lbvh");
        }

    private:
        vren::context const* m_context;
    };
} // namespace vren
