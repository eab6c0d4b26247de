// pyramid.hip -- depth_buffer_pyramid + depth_buffer_reductor::copy_and_reduce (depth_buffer_pyramid.cpp:14-23,
// 177-299; depth_buffer_copy.comp; depth_buffer_reduce.comp): the Hi-Z max pyramid of a D32 depth buffer.
//
// Level 0 is a copy of the depth buffer; level l + 1 has max(W >> (l + 1), 1) x max(H >> (l + 1), 1) texels and
// texel (x, y) is max(0, the level-l texels (2x + i, 2y + j) that lie inside level l), taken in the reference's
// order i = 0, 1 outer, j = 0, 1 inner with GLSL max(a, b) = a < b ? b : a (so NaN and negative inputs never win,
// and an odd last row or column of a level is not read by the next one). The reference dispatches one 32x32 pass
// per level with a barrier between levels; here one workgroup takes a 64x64 region of a level into LDS and writes
// the six levels below it (64 -> 1 texels), so the depth is read once and every level written once:
//   launch 1: depth -> level 0 (copy) and levels 1..6;  launch 2: level 6 -> levels 7..12;  launch 3: 12 -> 13..15.
// A region's texels at every level below depend on that region alone (level sizes halve with the floor), so the
// workgroups are independent. Images of a batch are stacked on the grid's z axis.
#include "common.hpp"

namespace {

constexpr int kRegion = 64;   // texels of the input level per workgroup side
constexpr int kThreads = 256;
constexpr int kLevelsPerLaunch = 6;

__host__ __device__ inline uint32_t level_dim(uint32_t d, uint32_t l) { return (d >> l) > 0 ? (d >> l) : 1u; }

struct PyramidLaunch {
    const float* src;        // the input level of image 0
    uint64_t src_image_stride; // floats between images of `src`
    float* pyr;              // the pyramid of image 0
    uint64_t pyr_image_stride;
    uint32_t W, H;           // level 0 size
    uint32_t lin;            // input level
    uint32_t n_out;          // output levels lin + 1 .. lin + n_out
    uint32_t copy;           // write the input to level lin (launch 1: the depth copy to level 0)
    uint64_t level_off[17];  // float offset of each level in one image's pyramid
};

__device__ __forceinline__ float glsl_max(float a, float b) { return a < b ? b : a; }

__global__ __launch_bounds__(kThreads) void pyramid_kernel(const PyramidLaunch a)
{
    __shared__ float s_a[kRegion * kRegion];
    __shared__ float s_b[(kRegion / 2) * (kRegion / 2)];
    const uint32_t tid = threadIdx.x, lane = tid % VREN_WAVE, wave = tid / VREN_WAVE;
    const uint32_t img = blockIdx.z;
    const float* src = a.src + (uint64_t)img * a.src_image_stride;
    float* pyr = a.pyr + (uint64_t)img * a.pyr_image_stride;
    const uint32_t Wi = level_dim(a.W, a.lin), Hi = level_dim(a.H, a.lin);
    const uint32_t x0 = blockIdx.x * kRegion, y0 = blockIdx.y * kRegion;
    // the region of the input level: one 64-texel row per wave and load, 16 rows per wave
    float* copy_dst = a.copy ? pyr + a.level_off[a.lin] : nullptr;
#pragma unroll 4
    for (int r = 0; r < kRegion / 4; r++) {
        const uint32_t ly = wave * (kRegion / 4) + r;
        const uint32_t gx = x0 + lane, gy = y0 + ly;
        float v = 0.0f;
        if (gx < Wi && gy < Hi) {
            v = src[(uint64_t)gy * Wi + gx];
            if (copy_dst) copy_dst[(uint64_t)gy * Wi + gx] = v;
        }
        s_a[ly * kRegion + lane] = v;
    }
    __syncthreads();
    // levels lin + 1 .. lin + n_out, ping-ponging between the two LDS arrays
    float* in = s_a;
    float* out = s_b;
    uint32_t side = kRegion; // input texels per side of the region at the current level
    for (uint32_t k = 1; k <= a.n_out; k++) {
        const uint32_t lp = a.lin + k - 1, l = a.lin + k;
        const uint32_t Wp = level_dim(a.W, lp), Hp = level_dim(a.H, lp);
        const uint32_t Wl = level_dim(a.W, l), Hl = level_dim(a.H, l);
        const uint32_t half = side / 2;
        const uint32_t px0 = x0 >> (k - 1), py0 = y0 >> (k - 1); // region origin at the input level
        const uint32_t ox0 = x0 >> k, oy0 = y0 >> k;              // and at the output level
        float* dst = pyr + a.level_off[l];
        for (uint32_t t = tid; t < half * half; t += kThreads) {
            const uint32_t x = t % half, y = t / half;
            float m = 0.0f;
#pragma unroll
            for (uint32_t i = 0; i < 2; i++)
#pragma unroll
                for (uint32_t j = 0; j < 2; j++) {
                    const uint32_t fx = 2 * x + i, fy = 2 * y + j;
                    if (px0 + fx < Wp && py0 + fy < Hp) m = glsl_max(m, in[fy * side + fx]);
                }
            out[y * half + x] = m;
            if (ox0 + x < Wl && oy0 + y < Hl) dst[(uint64_t)(oy0 + y) * Wl + (ox0 + x)] = m;
        }
        __syncthreads();
        float* tmp = in;
        in = out;
        out = tmp;
        side = half;
    }
}

uint32_t pyramid_levels(uint32_t W, uint32_t H)
{
    uint32_t m = W > H ? W : H, l = 0;
    while (m > 1) {
        m >>= 1;
        l++;
    }
    return l + 1; // glm::log2(max(W, H)) + 1 on unsigned integers (depth_buffer_pyramid.cpp:18)
}

uint64_t pyramid_offset(uint32_t W, uint32_t H, uint32_t level)
{
    uint64_t o = 0;
    for (uint32_t l = 0; l < level; l++) o += (uint64_t)level_dim(W, l) * level_dim(H, l);
    return o;
}

} // namespace

extern "C" {

VREN_EXPORT uint32_t vren_depth_pyramid_level_count(uint32_t width, uint32_t height)
{
    return (width == 0 || height == 0) ? 0u : pyramid_levels(width, height);
}

VREN_EXPORT uint64_t vren_depth_pyramid_level_offset(uint32_t width, uint32_t height, uint32_t level)
{
    return pyramid_offset(width, height, level);
}

VREN_EXPORT uint64_t vren_depth_pyramid_length(uint32_t width, uint32_t height)
{
    return pyramid_offset(width, height, vren_depth_pyramid_level_count(width, height));
}

VREN_EXPORT int vren_depth_pyramid(vren_ctx* ctx, vren_stream_t stream, const float* depth, uint32_t width,
                                   uint32_t height, uint32_t images, float* pyramid)
{
    if (!ctx) return VREN_EINVAL;
    if (!depth || !pyramid || width == 0 || height == 0 || images == 0)
        return vren_fail(ctx, VREN_EINVAL, "depth_pyramid: null buffer or zero size");
    if (width > 65535 || height > 65535 || images > 65535)
        return vren_fail(ctx, VREN_EINVAL, "depth_pyramid: width, height and images must be at most 65535");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "depth_pyramid: context has no device");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t levels = pyramid_levels(width, height);
    PyramidLaunch a;
    memset(&a, 0, sizeof(a));
    for (uint32_t l = 0; l <= levels && l < 17; l++) a.level_off[l] = pyramid_offset(width, height, l);
    a.pyr = pyramid;
    a.pyr_image_stride = a.level_off[levels];
    a.W = width;
    a.H = height;
    for (uint32_t lin = 0; lin == 0 || lin + 1 < levels; lin += kLevelsPerLaunch) {
        a.lin = lin;
        a.copy = lin == 0;
        a.src = lin == 0 ? depth : pyramid + a.level_off[lin];
        a.src_image_stride = lin == 0 ? (uint64_t)width * height : a.pyr_image_stride;
        a.n_out = std::min<uint32_t>(kLevelsPerLaunch, levels - 1 - lin);
        const uint32_t Wi = level_dim(width, lin), Hi = level_dim(height, lin);
        dim3 grid((Wi + kRegion - 1) / kRegion, (Hi + kRegion - 1) / kRegion, images);
        pyramid_kernel<<<grid, kThreads, 0, s>>>(a);
        VREN_LAUNCH_CHECK(ctx);
        if (lin + kLevelsPerLaunch >= levels) break;
    }
    return VREN_OK;
}

} // extern "C"
