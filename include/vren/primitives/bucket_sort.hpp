// vren/primitives/bucket_sort.hpp -- drop-in for bucket_sort.hpp:9-45 (bucket_sort.cpp:62-161).
#pragma once

#include "../base.hpp"

namespace vren
{
    class bucket_sort
    {
    public:
        inline static const uint32_t k_key_size = 1 << 16;
        inline static const uint32_t k_key_mask = k_key_size - 1;
        inline static const size_t k_offset_alignment = 256;

        explicit bucket_sort(vren::context const& context) : m_context(&context) {}

        // bucket_sort.cpp:67-70: the sorted uvec2 pairs, padded to 256 bytes, then k_key_size uint32 bucket offsets
        static size_t get_required_output_buffer_size(uint32_t length)
        {
            return pairs_bytes(length) + k_key_size * sizeof(uint32_t);
        }

        // Counting sort of `input_buffer_length` (key, value) uint32 pairs by key & 0xFFFF. Stable here (the
        // reference's atomic scatter is not), which is one of the orders the reference may produce.
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& input_buffer, uint32_t input_buffer_length,
                        size_t input_buffer_offset, vren::device_buffer const& output_buffer,
                        size_t output_buffer_offset) const
        {
            (void) resource_container;
            vren_ctx* ctx = m_context->handle();
            uint32_t* out = output_buffer.at<uint32_t>(output_buffer_offset);
            uint32_t* offsets = output_buffer.at<uint32_t>(output_buffer_offset + pairs_bytes(input_buffer_length));
            check_status(ctx,
                         vren_bucket_sort(ctx, stream.get(), input_buffer.at<uint32_t>(input_buffer_offset),
                                          input_buffer_length, out, offsets, k_key_size, 0),
                         "vren::bucket_sort");
        }

    private:
        static size_t pairs_bytes(uint32_t length)
        {
            return ((size_t) length * 8 + k_offset_alignment - 1) / k_offset_alignment * k_offset_alignment;
        }
        vren::context const* m_context;
    };
} // namespace vren
