// vren/primitives/radix_sort.hpp -- drop-in for radix_sort.hpp:9-53 (radix_sort.cpp:124-338).
#pragma once

#include "../base.hpp"

namespace vren
{
    class radix_sort
    {
    public:
        inline static const uint32_t k_workgroup_size = 1024; // the reference's minimum length (radix_sort.cpp:158)

        explicit radix_sort(vren::context const& context) : m_context(&context) {}

        // radix_sort.cpp:124-147. Scratch 1 is the onesweep sort's whole workspace (ping-pong keys, digit
        // histograms, look-back statuses); scratch 2 is not used and may be empty.
        vren::owned_device_buffer create_scratch_buffer_1(uint32_t length) const
        {
            return vren::owned_device_buffer(vren_radix_sort_scratch_bytes(length, 0));
        }
        vren::owned_device_buffer create_scratch_buffer_2(uint32_t length) const
        {
            (void) length;
            return vren::owned_device_buffer();
        }
        static size_t get_required_scratch_buffer_size(uint32_t length, bool with_values)
        {
            return vren_radix_sort_scratch_bytes(length, with_values ? 1 : 0);
        }

        // Stable ascending sort of `length` uint32 keys in place (radix_sort.cpp:149-338), with the reference's
        // precondition: length >= 1024 and a power of 2, else std::invalid_argument.
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& buffer, uint32_t length,
                        vren::device_buffer const& scratch_buffer_1, vren::device_buffer const& scratch_buffer_2) const
        {
            (void) resource_container;
            (void) scratch_buffer_2;
            if (!(length >= k_workgroup_size && vren::is_power_of_2(length)))
                throw std::invalid_argument("Length must be higher than 1024 and a power of 2");
            vren_ctx* ctx = m_context->handle();
            check_status(ctx,
                         vren_radix_sort(ctx, stream.get(), buffer.at<uint32_t>(), nullptr, length, 0, 32,
                                         scratch_buffer_1.m_ptr, scratch_buffer_1.m_size),
                         "vren::radix_sort");
        }

        // Key-value form (no reference counterpart): any length >= 1, values carried stably with their keys,
        // keys compared on bits [begin_bit, end_bit).
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& keys, vren::device_buffer const& values, uint32_t length,
                        vren::device_buffer const& scratch, uint32_t begin_bit = 0, uint32_t end_bit = 32) const
        {
            (void) resource_container;
            vren_ctx* ctx = m_context->handle();
            check_status(ctx,
                         vren_radix_sort(ctx, stream.get(), keys.at<uint32_t>(), values.at<uint32_t>(), length,
                                         begin_bit, end_bit, scratch.m_ptr, scratch.m_size),
                         "vren::radix_sort");
        }

    private:
        vren::context const* m_context;
    };
} // namespace vren
