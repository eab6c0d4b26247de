// vren/primitives/reduce.hpp -- drop-in for reduce.hpp:15-49 (reduce.cpp:32-140).
#pragma once

#include "../base.hpp"

namespace vren
{
    enum reduce_operation
    {
        ReduceOperationAdd,
        ReduceOperationMin,
        ReduceOperationMax
    };

    // Writes the up-sweep tree of the input (padded to the next power of 2 with the operation's identity) into
    // the output: tree[i] = OP over the aligned block of 2^ctz(i+1) elements ending at i, the total at the last
    // element. _data_type_t is uint32_t or vren::vec4 (component-wise). `blocks_num` independent segments of
    // `input_buffer_length` elements each, like the reference's grid.y.
    template <typename _data_type_t, vren::reduce_operation _operation_t>
    class reduce
    {
        static_assert(std::is_same<_data_type_t, uint32_t>::value || std::is_same<_data_type_t, vren::vec4>::value,
                      "reduce is instantiated for uint32_t and vren::vec4");

    public:
        explicit reduce(vren::context const& context) : m_context(&context) {}

        // reduce.cpp:32-114
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& input_buffer, uint32_t input_buffer_length,
                        size_t input_buffer_offset, vren::device_buffer const& output_buffer,
                        size_t output_buffer_offset, uint32_t blocks_num) const
        {
            (void) resource_container;
            int op = _operation_t == ReduceOperationAdd ? VREN_REDUCE_ADD
                     : _operation_t == ReduceOperationMin ? VREN_REDUCE_MIN
                                                          : VREN_REDUCE_MAX;
            vren_ctx* ctx = m_context->handle();
            if (std::is_same<_data_type_t, uint32_t>::value) {
                check_status(ctx,
                             vren_reduce_u32(ctx, stream.get(), input_buffer.at<uint32_t>(input_buffer_offset),
                                             input_buffer_length, output_buffer.at<uint32_t>(output_buffer_offset), op,
                                             blocks_num),
                             "vren::reduce");
            } else {
                check_status(ctx,
                             vren_reduce_vec4(ctx, stream.get(), input_buffer.at<float>(input_buffer_offset),
                                              input_buffer_length, output_buffer.at<float>(output_buffer_offset), op,
                                              blocks_num),
                             "vren::reduce");
            }
        }

        // in place (reduce.cpp:116-128)
        void operator()(vren::stream stream, vren::resource_container& resource_container,
                        vren::device_buffer const& buffer, uint32_t length, size_t offset, uint32_t blocks_num) const
        {
            (*this)(stream, resource_container, buffer, length, offset, buffer, offset, blocks_num);
        }

    private:
        vren::context const* m_context;
    };

    // reduce.cpp:137-140
    inline uint32_t calc_reduce_output_buffer_length(uint32_t count) { return vren::round_to_next_power_of_2(count); }
} // namespace vren
