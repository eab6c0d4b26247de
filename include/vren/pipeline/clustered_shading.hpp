// vren/pipeline/clustered_shading.hpp -- drop-in for clustered_shading.hpp:12-182 (clustered_shading.cpp:29-1172).
//
// The four stages keep the reference's class names and operator() argument order. Differences at the boundary:
//   - images become device buffers: the D32 depth (f32[W*H]), the RGBA16F normals (f16x4[W*H]), the R32UI cluster
//     reference image (u32[W*H]) and the RGBA32F output (f32x4[W*H]), all row-major;
//   - shade and cluster_and_shade take the reference's material path (texcoord + material index G-buffer planes,
//     a vren::material_buffer and the context's texture manager, resolved as shade.comp:89-96 does) and, as a
//     second overload, per-pixel resolved materials (albedo + roughness f32x4, metallic f32);
//   - capacities are the sizes of the buffers passed in (the reference assumes its config.hpp maxima), and an
//     overflow is reported by vren::context::sync() instead of writing out of bounds;
//   - each stage takes the cluster configuration (vren_cluster_cfg) in m_config, defaulting to the reference's.
#pragma once

#include "../material.hpp"
#include "../primitives/build_bvh.hpp"

namespace vren
{
    struct uvec2
    {
        uint32_t x, y;
    };

    // camera.hpp fields (camera.cpp:9-50 builds the matrices)
    struct camera
    {
        vren::vec3 m_position = {0.0f, 0.0f, 0.0f};
        float m_yaw = 0.0f;
        float m_pitch = 0.0f;
        float m_fov_y = 0.78539819f; // 45 degrees
        float m_aspect_ratio = 1.0f;
        float m_near_plane = 0.01f;
        float m_far_plane = 1000.0f;

        vren_camera to_c() const
        {
            vren_camera c;
            c.position[0] = m_position.x;
            c.position[1] = m_position.y;
            c.position[2] = m_position.z;
            c.fov_y = m_fov_y;
            c.aspect_ratio = m_aspect_ratio;
            c.near_plane = m_near_plane;
            c.far_plane = m_far_plane;
            c.yaw = m_yaw;
            c.pitch = m_pitch;
            return c;
        }
        // column-major 4x4
        void get_view(float out[16]) const
        {
            float proj[16];
            vren_camera c = to_c();
            vren_camera_matrices(&c, out, proj);
        }
        void get_projection(float out[16]) const
        {
            float view[16];
            vren_camera c = to_c();
            vren_camera_matrices(&c, view, out);
        }
    };

    // light.hpp:21-38: point-light positions (vec4, w unused), point lights {vec3 colour x intensity; float radius}
    // and directional lights {vec3 direction; pad; vec3 colour; pad}
    struct light_array
    {
        vren::device_buffer m_point_light_position_buffer;
        vren::device_buffer m_point_light_buffer;
        uint32_t m_point_light_count = 0;
        vren::device_buffer m_directional_light_buffer;
        uint32_t m_directional_light_count = 0;
    };

    // gbuffer.hpp / gbuffer.cpp:7-27: normals (RGBA16F), texcoords (RG32F) and material indices (R16_UINT); the
    // resolved-material planes serve the overloads that take no material buffer
    struct gbuffer
    {
        vren::device_buffer m_normal_buffer;           // f16x4[W*H]
        vren::device_buffer m_texcoord_buffer;         // f32x2[W*H]
        vren::device_buffer m_material_index_buffer;   // u16[W*H]
        vren::device_buffer m_albedo_roughness_buffer; // f32x4[W*H]
        vren::device_buffer m_metallic_buffer;         // f32[W*H]
    };

    struct depth_buffer_t
    {
        vren::device_buffer m_buffer; // f32[W*H] NDC depth
    };

    namespace clustered_shading
    {
        inline vren_cluster_cfg reference_config()
        {
            vren_cluster_cfg c;
            vren_cluster_cfg_default(&c);
            return c;
        }

        // clustered_shading.hpp:18-42, clustered_shading.cpp:29-346
        class construct_point_light_bvh
        {
        public:
            vren_cluster_cfg m_config = reference_config();

            explicit construct_point_light_bvh(vren::context const& context) : m_context(&context) {}

            static size_t get_required_bvh_buffer_size(uint32_t point_light_count)
            {
                return vren::calc_bvh_buffer_size(point_light_count);
            }
            // (Morton code, light index) pairs, sorted by code
            static size_t get_required_point_light_index_buffer_size(uint32_t point_light_count)
            {
                return (size_t) point_light_count * 8;
            }
            static size_t get_required_view_space_point_light_position_buffer_size(uint32_t point_light_count)
            {
                return (size_t) point_light_count * 16;
            }

            void operator()(vren::stream stream, vren::resource_container& resource_container,
                            vren::light_array const& light_array,
                            vren::device_buffer const& view_space_point_light_position_buffer,
                            vren::camera const& camera, vren::device_buffer const& bvh_buffer,
                            vren::device_buffer const& point_light_index_buffer) const
            {
                (void) resource_container;
                vren_ctx* ctx = m_context->handle();
                const uint32_t L = light_array.m_point_light_count;
                if (bvh_buffer.m_size < get_required_bvh_buffer_size(L) ||
                    point_light_index_buffer.m_size < get_required_point_light_index_buffer_size(L) ||
                    view_space_point_light_position_buffer.m_size <
                        get_required_view_space_point_light_position_buffer_size(L))
                    throw std::invalid_argument("construct_point_light_bvh: output buffer too small");
                vren_camera c = camera.to_c();
                check_status(ctx,
                             vren_construct_point_light_bvh(
                                 ctx, stream.get(), light_array.m_point_light_position_buffer.at<float>(),
                                 light_array.m_point_light_buffer.at<float>(), L, &c, &m_config,
                                 view_space_point_light_position_buffer.at<float>(),
                                 point_light_index_buffer.at<uint32_t>(), bvh_buffer.at<vren_bvh_node>()),
                             "vren::clustered_shading::construct_point_light_bvh");
            }

        private:
            vren::context const* m_context;
        };

        // clustered_shading.hpp:47-74, clustered_shading.cpp:363-448
        class find_unique_cluster_list
        {
        public:
            vren_cluster_cfg m_config = reference_config();

            explicit find_unique_cluster_list(vren::context const& context) : m_context(&context) {}

            // cluster_key_buffer: u32[capacity] unique keys; allocation_index_buffer: u32 unique-key count (device);
            // cluster_reference_buffer: u32[W*H] index of each pixel's key
            void operator()(uint32_t frame_idx, vren::stream stream, vren::resource_container& resource_container,
                            vren::uvec2 const& screen, vren::camera const& camera, vren::gbuffer const& gbuffer,
                            vren::depth_buffer_t const& depth_buffer, vren::device_buffer const& cluster_key_buffer,
                            vren::device_buffer const& allocation_index_buffer,
                            vren::device_buffer const& cluster_reference_buffer) const
            {
                (void) frame_idx;
                (void) resource_container;
                if (cluster_reference_buffer.m_size < (size_t) screen.x * screen.y * 4 || allocation_index_buffer.m_size < 4)
                    throw std::invalid_argument("find_unique_cluster_list: output buffer too small");
                vren_ctx* ctx = m_context->handle();
                vren_camera c = camera.to_c();
                check_status(ctx,
                             vren_find_unique_cluster_list(
                                 ctx, stream.get(), screen.x, screen.y, &c, &m_config,
                                 depth_buffer.m_buffer.at<float>(), gbuffer.m_normal_buffer.at<uint16_t>(),
                                 cluster_key_buffer.at<uint32_t>(), (uint32_t) (cluster_key_buffer.m_size / 4),
                                 allocation_index_buffer.at<uint32_t>(), cluster_reference_buffer.at<uint32_t>()),
                             "vren::clustered_shading::find_unique_cluster_list");
            }

        private:
            vren::context const* m_context;
        };

        // clustered_shading.hpp:79-108, clustered_shading.cpp:473-690
        class assign_lights
        {
        public:
            vren_cluster_cfg m_config = reference_config();

            explicit assign_lights(vren::context const& context) : m_context(&context), m_assigned_total(4) {}

            void operator()(uint32_t frame_idx, vren::stream stream, vren::resource_container& resource_container,
                            vren::uvec2 const& screen, vren::camera const& camera,
                            vren::device_buffer const& cluster_key_buffer,
                            vren::device_buffer const& allocation_index_buffer,
                            vren::device_buffer const& light_bvh_buffer, uint32_t light_bvh_root_index,
                            uint32_t light_count, vren::device_buffer const& light_index_buffer,
                            vren::device_buffer const& assigned_light_indices_buffer,
                            vren::device_buffer const& assigned_light_counts_buffer,
                            vren::device_buffer const& assigned_light_offsets_buffer,
                            vren::device_buffer const& view_space_point_light_position_buffer) const
            {
                (void) frame_idx;
                (void) resource_container;
                if (light_count > 0 && light_bvh_root_index != vren::calc_bvh_root_index(light_count))
                    throw std::invalid_argument("assign_lights: light_bvh_root_index does not match light_count");
                size_t key_cap = cluster_key_buffer.m_size;
                if (assigned_light_counts_buffer.m_size < key_cap) key_cap = assigned_light_counts_buffer.m_size;
                if (assigned_light_offsets_buffer.m_size < key_cap) key_cap = assigned_light_offsets_buffer.m_size;
                vren_ctx* ctx = m_context->handle();
                vren_camera c = camera.to_c();
                check_status(ctx,
                             vren_assign_lights(ctx, stream.get(), screen.x, screen.y, &c, &m_config,
                                                cluster_key_buffer.at<uint32_t>(),
                                                allocation_index_buffer.at<uint32_t>(), (uint32_t) (key_cap / 4),
                                                light_bvh_buffer.at<vren_bvh_node>(), light_count,
                                                light_index_buffer.at<uint32_t>(),
                                                view_space_point_light_position_buffer.at<float>(),
                                                assigned_light_counts_buffer.at<uint32_t>(),
                                                assigned_light_offsets_buffer.at<uint32_t>(),
                                                assigned_light_indices_buffer.at<uint32_t>(),
                                                (uint32_t) (assigned_light_indices_buffer.m_size / 4),
                                                m_assigned_total.at<uint32_t>()),
                             "vren::clustered_shading::assign_lights");
            }

            // total number of assigned light indices of the last call (device u32)
            vren::device_buffer const& assigned_total() const { return m_assigned_total; }

        private:
            vren::context const* m_context;
            vren::owned_device_buffer m_assigned_total;
        };

        // clustered_shading.hpp:113-140, clustered_shading.cpp:709-811
        class shade
        {
        public:
            vren_cluster_cfg m_config = reference_config();

            explicit shade(vren::context const& context) : m_context(&context) {}

            void operator()(uint32_t frame_idx, vren::stream stream, vren::resource_container& resource_container,
                            vren::uvec2 const& screen, vren::camera const& camera, vren::gbuffer const& gbuffer,
                            vren::depth_buffer_t const& depth_buffer,
                            vren::device_buffer const& cluster_reference_buffer,
                            vren::device_buffer const& assigned_light_indices_buffer,
                            vren::device_buffer const& assigned_light_counts_buffer,
                            vren::device_buffer const& assigned_light_offsets_buffer,
                            vren::light_array const& light_array, vren::device_buffer const& output) const
            {
                (void) frame_idx;
                (void) resource_container;
                if (output.m_size < (size_t) screen.x * screen.y * 16)
                    throw std::invalid_argument("shade: output buffer too small");
                size_t key_cap = assigned_light_counts_buffer.m_size < assigned_light_offsets_buffer.m_size
                                     ? assigned_light_counts_buffer.m_size
                                     : assigned_light_offsets_buffer.m_size;
                vren_ctx* ctx = m_context->handle();
                vren_camera c = camera.to_c();
                check_status(
                    ctx,
                    vren_shade(ctx, stream.get(), screen.x, screen.y, &c, &m_config, depth_buffer.m_buffer.at<float>(),
                               gbuffer.m_normal_buffer.at<uint16_t>(), gbuffer.m_albedo_roughness_buffer.at<float>(),
                               gbuffer.m_metallic_buffer.at<float>(), cluster_reference_buffer.at<uint32_t>(),
                               (uint32_t) (key_cap / 4), assigned_light_counts_buffer.at<uint32_t>(),
                               assigned_light_offsets_buffer.at<uint32_t>(),
                               assigned_light_indices_buffer.at<uint32_t>(),
                               (uint32_t) (assigned_light_indices_buffer.m_size / 4),
                               light_array.m_point_light_position_buffer.at<float>(),
                               light_array.m_point_light_buffer.at<float>(), light_array.m_point_light_count,
                               light_array.m_directional_light_count ? light_array.m_directional_light_buffer.at<float>()
                                                                     : nullptr,
                               light_array.m_directional_light_count, output.at<float>()),
                    "vren::clustered_shading::shade");
            }

            // the reference's signature (clustered_shading.hpp:129-140): materials resolved from the G-buffer's
            // texcoords and material indices, the material buffer and the context's textures (shade.comp:89-96)
            void operator()(uint32_t frame_idx, vren::stream stream, vren::resource_container& resource_container,
                            vren::uvec2 const& screen, vren::camera const& camera, vren::gbuffer const& gbuffer,
                            vren::depth_buffer_t const& depth_buffer,
                            vren::device_buffer const& cluster_reference_buffer,
                            vren::device_buffer const& assigned_light_indices_buffer,
                            vren::device_buffer const& assigned_light_counts_buffer,
                            vren::device_buffer const& assigned_light_offsets_buffer,
                            vren::light_array const& light_array, vren::material_buffer const& material_buffer,
                            vren::device_buffer const& output) const
            {
                (void) frame_idx;
                (void) resource_container;
                const size_t px = (size_t) screen.x * screen.y;
                if (output.m_size < px * 16 || gbuffer.m_texcoord_buffer.m_size < px * 8 ||
                    gbuffer.m_material_index_buffer.m_size < px * 2)
                    throw std::invalid_argument("shade: output, texcoord or material-index buffer too small");
                size_t key_cap = assigned_light_counts_buffer.m_size < assigned_light_offsets_buffer.m_size
                                     ? assigned_light_counts_buffer.m_size
                                     : assigned_light_offsets_buffer.m_size;
                vren_material_set ms;
                ms.texcoords = gbuffer.m_texcoord_buffer.at<float>();
                ms.material_indices = gbuffer.m_material_index_buffer.at<uint16_t>();
                ms.materials = material_buffer.m_material_count ? material_buffer.m_buffer.at<vren_material>() : nullptr;
                ms.material_count = material_buffer.m_material_count;
                ms.texture_count = m_context->m_texture_manager.count();
                ms.textures = m_context->m_texture_manager.device_table();
                vren_ctx* ctx = m_context->handle();
                vren_camera c = camera.to_c();
                check_status(
                    ctx,
                    vren_shade_materials(
                        ctx, stream.get(), screen.x, screen.y, &c, &m_config, depth_buffer.m_buffer.at<float>(),
                        gbuffer.m_normal_buffer.at<uint16_t>(), &ms, cluster_reference_buffer.at<uint32_t>(),
                        (uint32_t) (key_cap / 4), assigned_light_counts_buffer.at<uint32_t>(),
                        assigned_light_offsets_buffer.at<uint32_t>(), assigned_light_indices_buffer.at<uint32_t>(),
                        (uint32_t) (assigned_light_indices_buffer.m_size / 4),
                        light_array.m_point_light_position_buffer.at<float>(),
                        light_array.m_point_light_buffer.at<float>(), light_array.m_point_light_count,
                        light_array.m_directional_light_count ? light_array.m_directional_light_buffer.at<float>()
                                                              : nullptr,
                        light_array.m_directional_light_count, output.at<float>()),
                    "vren::clustered_shading::shade");
            }

        private:
            vren::context const* m_context;
        };
    } // namespace clustered_shading

    // clustered_shading.hpp:147-182, clustered_shading.cpp:817-1172: owns the intermediate buffers, sized at the
    // reference's maxima (config.hpp:15-24) unless other limits are given, and runs the four stages in order on
    // one stream. The reference returns a render-graph node; here the work is enqueued directly.
    class cluster_and_shade
    {
    public:
        struct limits
        {
            uint32_t m_max_screen_width = 1920;
            uint32_t m_max_screen_height = 1080;
            uint32_t m_max_point_light_count = 1u << 20;
            uint32_t m_max_unique_cluster_key_count = 1u << 17;
            uint32_t m_max_assigned_light_count = 1u << 23;
        };

        vren::clustered_shading::construct_point_light_bvh m_construct_point_light_bvh;
        vren::clustered_shading::find_unique_cluster_list m_find_unique_cluster_list;
        vren::clustered_shading::assign_lights m_assign_lights;
        vren::clustered_shading::shade m_shade;

        vren::owned_device_buffer m_view_space_point_light_position_buffer;
        vren::owned_device_buffer m_point_light_bvh_buffer;
        vren::owned_device_buffer m_point_light_index_buffer;
        vren::owned_device_buffer m_cluster_key_buffer;
        vren::owned_device_buffer m_cluster_key_dispatch_params_buffer; // the unique-key count
        vren::owned_device_buffer m_cluster_reference_buffer;
        vren::owned_device_buffer m_assigned_light_indices_buffer;
        vren::owned_device_buffer m_assigned_light_counts_buffer;
        vren::owned_device_buffer m_assigned_light_offsets_buffer;

        explicit cluster_and_shade(vren::context const& context) : cluster_and_shade(context, limits()) {}
        cluster_and_shade(vren::context const& context, limits const& l) :
            m_construct_point_light_bvh(context), m_find_unique_cluster_list(context), m_assign_lights(context),
            m_shade(context),
            m_view_space_point_light_position_buffer((size_t) l.m_max_point_light_count * 16),
            m_point_light_bvh_buffer(vren::calc_bvh_buffer_size(l.m_max_point_light_count)),
            m_point_light_index_buffer((size_t) l.m_max_point_light_count * 8),
            m_cluster_key_buffer((size_t) l.m_max_unique_cluster_key_count * 4),
            m_cluster_key_dispatch_params_buffer(4),
            m_cluster_reference_buffer((size_t) l.m_max_screen_width * l.m_max_screen_height * 4),
            m_assigned_light_indices_buffer((size_t) l.m_max_assigned_light_count * 4),
            m_assigned_light_counts_buffer((size_t) l.m_max_unique_cluster_key_count * 4),
            m_assigned_light_offsets_buffer((size_t) l.m_max_unique_cluster_key_count * 4), m_limits(l)
        {
        }

        void set_config(vren_cluster_cfg const& cfg)
        {
            m_construct_point_light_bvh.m_config = cfg;
            m_find_unique_cluster_list.m_config = cfg;
            m_assign_lights.m_config = cfg;
            m_shade.m_config = cfg;
        }

        // the reference's signature (clustered_shading.hpp:173-181): shade resolves materials from the material
        // buffer and the context's textures
        void operator()(vren::stream stream, vren::uvec2 const& screen, vren::camera const& camera,
                        vren::gbuffer const& gbuffer, vren::depth_buffer_t const& depth_buffer,
                        vren::light_array const& light_array, vren::material_buffer const& material_buffer,
                        vren::device_buffer const& output)
        {
            run(stream, screen, camera, gbuffer, depth_buffer, light_array, &material_buffer, output);
        }

        // per-pixel resolved materials (gbuffer.m_albedo_roughness_buffer, gbuffer.m_metallic_buffer)
        void operator()(vren::stream stream, vren::uvec2 const& screen, vren::camera const& camera,
                        vren::gbuffer const& gbuffer, vren::depth_buffer_t const& depth_buffer,
                        vren::light_array const& light_array, vren::device_buffer const& output)
        {
            run(stream, screen, camera, gbuffer, depth_buffer, light_array, nullptr, output);
        }

    private:
        void run(vren::stream stream, vren::uvec2 const& screen, vren::camera const& camera,
                 vren::gbuffer const& gbuffer, vren::depth_buffer_t const& depth_buffer,
                 vren::light_array const& light_array, vren::material_buffer const* material_buffer,
                 vren::device_buffer const& output)
        {
            if (screen.x > m_limits.m_max_screen_width || screen.y > m_limits.m_max_screen_height ||
                light_array.m_point_light_count > m_limits.m_max_point_light_count)
                throw std::invalid_argument("cluster_and_shade: screen or light count above the configured limits");
            vren::resource_container rc;
            const uint32_t L = light_array.m_point_light_count;
            vren::device_buffer bvh = sized(m_point_light_bvh_buffer, vren::calc_bvh_buffer_size(L));
            vren::device_buffer codes = sized(m_point_light_index_buffer, (size_t) L * 8);
            vren::device_buffer view_pos = sized(m_view_space_point_light_position_buffer, (size_t) L * 16);
            vren::device_buffer refs = sized(m_cluster_reference_buffer, (size_t) screen.x * screen.y * 4);
            m_construct_point_light_bvh(stream, rc, light_array, view_pos, camera, bvh, codes);
            m_find_unique_cluster_list(0, stream, rc, screen, camera, gbuffer, depth_buffer, m_cluster_key_buffer,
                                       m_cluster_key_dispatch_params_buffer, refs);
            m_assign_lights(0, stream, rc, screen, camera, m_cluster_key_buffer, m_cluster_key_dispatch_params_buffer,
                            bvh, L ? vren::calc_bvh_root_index(L) : 0, L, codes, m_assigned_light_indices_buffer,
                            m_assigned_light_counts_buffer, m_assigned_light_offsets_buffer, view_pos);
            if (material_buffer)
                m_shade(0, stream, rc, screen, camera, gbuffer, depth_buffer, refs, m_assigned_light_indices_buffer,
                        m_assigned_light_counts_buffer, m_assigned_light_offsets_buffer, light_array, *material_buffer,
                        output);
            else
                m_shade(0, stream, rc, screen, camera, gbuffer, depth_buffer, refs, m_assigned_light_indices_buffer,
                        m_assigned_light_counts_buffer, m_assigned_light_offsets_buffer, light_array, output);
        }

        static vren::device_buffer sized(vren::device_buffer const& b, size_t size)
        {
            vren::device_buffer r;
            r.m_ptr = b.m_ptr;
            r.m_size = size;
            return r;
        }
        limits m_limits;
    };
} // namespace vren
