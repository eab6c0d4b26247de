// vren/material.hpp -- drop-in for material.hpp:19-58: the material record shade resolves per pixel
// (shade.comp:89-96) and the device buffer holding them.
#pragma once

#include "base.hpp"

namespace vren
{
    // GLSL Material (common.glsl:61-68), std430: 32 bytes
    struct material
    {
        uint32_t m_base_color_texture_idx;
        uint32_t m_metallic_roughness_texture_idx;
        float m_metallic_factor;
        float m_roughness_factor;
        vren::vec4 m_base_color_factor;
    };
    static_assert(sizeof(material) == sizeof(vren_material), "vren::material is the 32-byte GLSL Material");

    // material.hpp:34-58: a device buffer of up to `capacity` materials; m_material_count of them are read
    class material_buffer
    {
    public:
        vren::owned_device_buffer m_buffer;
        uint32_t m_material_count = 0;

        explicit material_buffer(vren::context const& context, uint32_t capacity = 1024) :
            m_buffer((size_t) capacity * sizeof(vren::material))
        {
            (void) context;
        }

        // copies `count` materials to the device (the reference writes its host-visible buffer directly)
        void write(vren::material const* materials, uint32_t count)
        {
            if ((size_t) count * sizeof(vren::material) > m_buffer.m_size)
                throw std::invalid_argument("material_buffer: more materials than its capacity");
            if (count) {
                hipError_t e = hipMemcpy(m_buffer.m_ptr, materials, (size_t) count * sizeof(vren::material),
                                         hipMemcpyHostToDevice);
                if (e != hipSuccess) throw std::runtime_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
            }
            m_material_count = count;
        }
    };
} // namespace vren
