// vren/base.hpp -- the host-side types every drop-in operator takes instead of the reference's Vulkan ones.
//
//   reference (vren/vren)                        here
//   vren::context const& (context.hpp:37-91)     vren::context const&   (owns a vren_ctx on one HIP device)
//   VkCommandBuffer                              vren::stream           (a hipStream_t; work is enqueued, not recorded)
//   vren::resource_container&                    vren::resource_container& (kept for the call shape: buffers are
//                                                caller-owned device memory, so nothing has to outlive the call)
//   vren::vk_utils::buffer const& (+ offset)     vren::device_buffer const& (+ the same byte offset)
//
// Errors follow the reference: a violated precondition throws std::invalid_argument (the reference asserts, or
// throws it for the radix-sort length, radix_sort.cpp:158-161); a failed HIP call throws std::runtime_error
// (VREN_CHECK, misc.cpp:22-40). Every operator is asynchronous on its stream, like a recorded command buffer
// that the caller submits; vren::context::sync() is the fence wait.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../vren_hip/capi.h"

namespace vren
{
    inline void check_status(vren_ctx* ctx, int status, const char* what)
    {
        if (status == VREN_OK) return;
        std::string msg = std::string(what) + ": " + vren_status_string(status);
        if (ctx) {
            const char* detail = vren_last_error(ctx);
            if (detail && *detail) msg += std::string(" (") + detail + ")";
        }
        if (status == VREN_EINVAL) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }

    struct stream
    {
        hipStream_t m_handle = nullptr;

        stream() = default;
        stream(hipStream_t handle) : m_handle(handle) {}
        vren_stream_t get() const { return reinterpret_cast<vren_stream_t>(m_handle); }
    };

    class resource_container
    {
    };

    struct device_buffer
    {
        void* m_ptr = nullptr;
        size_t m_size = 0;

        template <typename T>
        T* at(size_t byte_offset = 0) const
        {
            return reinterpret_cast<T*>(static_cast<char*>(m_ptr) + byte_offset);
        }
    };

    // device memory owned by the object (what the reference's vk_utils::buffer is: an RAII allocation)
    class owned_device_buffer : public device_buffer
    {
    public:
        owned_device_buffer() = default;
        explicit owned_device_buffer(size_t size)
        {
            if (size == 0) return;
            hipError_t e = hipMalloc(&m_ptr, size);
            if (e != hipSuccess) throw std::runtime_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            m_size = size;
        }
        owned_device_buffer(owned_device_buffer const&) = delete;
        owned_device_buffer& operator=(owned_device_buffer const&) = delete;
        owned_device_buffer(owned_device_buffer&& o) noexcept { swap(o); }
        owned_device_buffer& operator=(owned_device_buffer&& o) noexcept
        {
            swap(o);
            return *this;
        }
        ~owned_device_buffer()
        {
            if (m_ptr) (void) hipFree(m_ptr);
        }

    private:
        void swap(owned_device_buffer& o)
        {
            std::swap(m_ptr, o.m_ptr);
            std::swap(m_size, o.m_size);
        }
    };

    // texture_manager.hpp:29-47: the bindless texture array shade samples. Each texture is an RGBA8 UNORM device
    // image (tinygltf_parser.cpp:157) with its sampler state; rewrite_descriptor_set() uploads the descriptor
    // table after textures are added, as the reference rewrites its descriptor set.
    struct texture
    {
        vren::device_buffer m_texels; // height rows of width RGBA8 texels, 4-byte aligned
        uint32_t m_width = 0, m_height = 0;
        uint32_t m_filter = VREN_FILTER_LINEAR;
        uint32_t m_address_u = VREN_ADDRESS_REPEAT, m_address_v = VREN_ADDRESS_REPEAT;
    };

    class texture_manager
    {
    public:
        static constexpr uint32_t k_max_texture_count = 65536;

        std::vector<vren::texture> m_textures;

        void rewrite_descriptor_set()
        {
            if (m_textures.size() > k_max_texture_count)
                throw std::invalid_argument("texture_manager: more than 65536 textures");
            std::vector<vren_texture> d(m_textures.size());
            for (size_t i = 0; i < m_textures.size(); i++) {
                vren::texture const& t = m_textures[i];
                if (t.m_texels.m_size < (size_t) t.m_width * t.m_height * 4)
                    throw std::invalid_argument("texture_manager: texel buffer smaller than width * height * 4");
                d[i] = vren_texture{t.m_texels.at<uint8_t>(), t.m_width, t.m_height, t.m_filter, t.m_address_u,
                                    t.m_address_v, 0u};
            }
            if (m_table.m_size < d.size() * sizeof(vren_texture)) m_table = owned_device_buffer(d.size() * sizeof(vren_texture));
            if (!d.empty()) {
                hipError_t e = hipMemcpy(m_table.m_ptr, d.data(), d.size() * sizeof(vren_texture), hipMemcpyHostToDevice);
                if (e != hipSuccess) throw std::runtime_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
            }
            m_count = (uint32_t) d.size();
        }

        vren_texture const* device_table() const { return m_count ? m_table.at<vren_texture>() : nullptr; }
        uint32_t count() const { return m_count; }

    private:
        owned_device_buffer m_table;
        uint32_t m_count = 0;
    };

    class context
    {
    public:
        // the reference keeps its texture manager in the context's toolbox (m_context->m_toolbox->m_texture_manager)
        mutable vren::texture_manager m_texture_manager;

        explicit context(int hip_device = 0)
        {
            check_status(nullptr, vren_create(hip_device, &m_handle), "vren_create");
        }
        context(context const&) = delete;
        context& operator=(context const&) = delete;
        ~context() { vren_destroy(m_handle); }

        vren_ctx* handle() const { return m_handle; }

        // waits for `s` and reports device-side capacity overflows of earlier calls (VREN_ECAPACITY)
        void sync(vren::stream s = {}) const { check_status(m_handle, vren_sync_status(m_handle, s.get()), "sync"); }

    private:
        vren_ctx* m_handle = nullptr;
    };

    struct vec3
    {
        float x, y, z;
    };
    struct vec4
    {
        float x, y, z, w;
    };

    inline uint32_t round_to_next_power_of_2(uint32_t v)
    {
        if (v <= 1) return 1;
        v--;
        v |= v >> 1;
        v |= v >> 2;
        v |= v >> 4;
        v |= v >> 8;
        v |= v >> 16;
        return v + 1;
    }
    inline bool is_power_of_2(uint32_t v) { return v != 0 && (v & (v - 1)) == 0; }
} // namespace vren
