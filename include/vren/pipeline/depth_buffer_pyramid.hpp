// vren/pipeline/depth_buffer_pyramid.hpp -- drop-in for depth_buffer_pyramid.hpp:19-107
// (depth_buffer_pyramid.cpp:14-23, 177-299): the Hi-Z max pyramid of a D32 depth buffer.
//
// The reference's pyramid is an R32F image with one mip level per pyramid level; here it is one device buffer
// holding the levels one after another, row-major, with level_offset() giving each level's first texel.
// depth_buffer_reductor::copy_and_reduce enqueues the build on a stream instead of returning a render-graph node.
#pragma once

#include "clustered_shading.hpp"

namespace vren
{
    class depth_buffer_pyramid
    {
    public:
        static const uint32_t k_max_depth_buffer_pyramid_level_count = 16;

        depth_buffer_pyramid(vren::context const& context, uint32_t width, uint32_t height) :
            m_base_width(width), m_base_height(height),
            m_level_count(vren_depth_pyramid_level_count(width, height)),
            m_buffer(check_size(width, height) * sizeof(float))
        {
            (void) context;
        }

        uint32_t get_image_width(uint32_t level) const { return (m_base_width >> level) > 0 ? m_base_width >> level : 1u; }
        uint32_t get_image_height(uint32_t level) const
        {
            return (m_base_height >> level) > 0 ? m_base_height >> level : 1u;
        }
        uint32_t get_level_count() const { return m_level_count; }
        // first texel of `level` in the buffer (floats)
        uint64_t level_offset(uint32_t level) const
        {
            return vren_depth_pyramid_level_offset(m_base_width, m_base_height, level);
        }
        vren::device_buffer const& buffer() const { return m_buffer; }

    private:
        static uint64_t check_size(uint32_t width, uint32_t height)
        {
            if (width == 0 || height == 0 || width > 65535 || height > 65535)
                throw std::invalid_argument("depth_buffer_pyramid: size must be in [1, 65535]^2");
            return vren_depth_pyramid_length(width, height);
        }

        uint32_t m_base_width, m_base_height;
        uint32_t m_level_count;
        vren::owned_device_buffer m_buffer;

        friend class depth_buffer_reductor;
    };

    class depth_buffer_reductor
    {
    public:
        explicit depth_buffer_reductor(vren::context const& context) : m_context(&context) {}

        // depth_buffer_pyramid.cpp:216-299: level 0 = the depth buffer, then the 2x2 max reduction level by level
        void copy_and_reduce(vren::stream stream, vren::depth_buffer_t const& depth_buffer,
                             vren::depth_buffer_pyramid const& depth_buffer_pyramid) const
        {
            vren::depth_buffer_pyramid const& p = depth_buffer_pyramid;
            if (depth_buffer.m_buffer.m_size < (size_t) p.m_base_width * p.m_base_height * sizeof(float))
                throw std::invalid_argument("depth_buffer_reductor: depth buffer smaller than the pyramid's base");
            vren_ctx* ctx = m_context->handle();
            check_status(ctx,
                         vren_depth_pyramid(ctx, stream.get(), depth_buffer.m_buffer.at<float>(), p.m_base_width,
                                            p.m_base_height, 1, p.m_buffer.at<float>()),
                         "vren::depth_buffer_reductor::copy_and_reduce");
        }

    private:
        vren::context const* m_context;
    };
} // namespace vren
