// vren/point_light_bouncer.hpp -- drop-in for vren_demo::point_light_bouncer (point_light_bouncer.hpp,
// point_light_bouncer.cpp:17-57), the per-frame point-light animation ahead of the light BVH build.
#pragma once

#include "base.hpp"

namespace vren_demo
{
    class point_light_bouncer
    {
    public:
        explicit point_light_bouncer(vren::context const& context) : m_context(&context) {}

        // Moves `point_light_count` lights in place by speed * dt inside [aabb_min, aabb_max], reflecting at the
        // walls (bounce_point_lights.comp:33-73). Positions and directions are vec4 buffers; w is written as 1 and
        // 0. frame_idx and resource_container are kept for the call shape (the reference uses them to keep its
        // descriptor set alive; nothing here needs them).
        void bounce(uint32_t frame_idx, vren::stream stream, vren::resource_container& resource_container,
                    vren::device_buffer const& point_light_position_buffer,
                    vren::device_buffer const& point_light_direction_buffer, uint32_t point_light_count,
                    vren::vec3 const& aabb_min, vren::vec3 const& aabb_max, float speed, float dt) const
        {
            (void) frame_idx;
            (void) resource_container;
            const size_t need = (size_t) point_light_count * 16;
            if (point_light_position_buffer.m_size < need || point_light_direction_buffer.m_size < need)
                throw std::invalid_argument("vren_demo::point_light_bouncer: position and direction buffers need "
                                            "point_light_count * 16 bytes");
            vren_ctx* ctx = m_context->handle();
            const float lo[3] = {aabb_min.x, aabb_min.y, aabb_min.z};
            const float hi[3] = {aabb_max.x, aabb_max.y, aabb_max.z};
            vren::check_status(ctx,
                               vren_bounce_point_lights(ctx, stream.get(), point_light_position_buffer.at<float>(0),
                                                        point_light_direction_buffer.at<float>(0), point_light_count,
                                                        lo, hi, speed, dt),
                               "vren_demo::point_light_bouncer");
        }

    private:
        vren::context const* m_context;
    };
} // namespace vren_demo
