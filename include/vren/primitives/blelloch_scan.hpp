// vren/primitives/blelloch_scan.hpp -- drop-in for blelloch_scan.hpp:9-49 (blelloch_scan.cpp:141-166).
#pragma once

#include "../base.hpp"

namespace vren
{
    class blelloch_scan
    {
    public:
        explicit blelloch_scan(vren::context const& context) : m_context(&context) {}

        // blelloch_scan.cpp:57-139: the down-sweep half of the two-phase scan, in place on `blocks_num` segments of
        // `length` (a power of 2) uint32 each, starting at byte `offset`. On an up-sweep tree (vren::reduce add)
        // with clear_last = true this leaves the exclusive prefix sum of the tree's leaves, as the reference's radix
        // sort uses it (radix_sort.cpp:281-289). With clear_last = false the segment's last element is the prefix
        // of the whole segment, except in segments shorter than 1024 (the reference's LDS pass zero-pads them, and
        // they are scanned from 0).
        void downsweep(vren::stream stream, vren::resource_container& resource_container,
                       vren::device_buffer const& buffer, uint32_t length, uint32_t offset, uint32_t blocks_num,
                       bool clear_last) const
        {
            (void) resource_container;
            vren_ctx* ctx = m_context->handle();
            check_status(ctx,
                         vren_blelloch_downsweep_u32(ctx, stream.get(), buffer.at<uint32_t>(offset), length,
                                                     blocks_num, clear_last ? 1 : 0),
                         "vren::blelloch_scan::downsweep");
        }

        // In-place exclusive prefix sum (mod 2^32) of `blocks_num` segments of `length` uint32 each, starting at byte
        // `offset`. The reference needs a power-of-2 length; any length >= 1 works here. It runs as one
        // single-pass look-back scan (one read and one write per element), not as a reduce followed by
        // downsweep().
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& buffer, uint32_t length, uint32_t offset, uint32_t blocks_num) const
        {
            (void) resource_container;
            vren_ctx* ctx = m_context->handle();
            uint32_t* p = buffer.at<uint32_t>(offset);
            check_status(ctx, vren_exclusive_scan_u32(ctx, stream.get(), p, p, length, blocks_num),
                         "vren::blelloch_scan");
        }

    private:
        vren::context const* m_context;
    };
} // namespace vren
