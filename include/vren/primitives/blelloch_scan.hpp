// vren/primitives/blelloch_scan.hpp -- drop-in for blelloch_scan.hpp:9-49 (blelloch_scan.cpp:141-166).
#pragma once

#include "../base.hpp"

namespace vren
{
    class blelloch_scan
    {
    public:
        explicit blelloch_scan(vren::context const& context) : m_context(&context) {}

        // In-place exclusive prefix sum (mod 2^32) of `blocks_num` segments of `length` uint32 each, starting at byte
        // `offset`. The reference needs a power-of-2 length; any length >= 1 works here. (The reference's public
        // downsweep() is the second half of its two-phase algorithm; the single-pass scan here has no such half.)
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
