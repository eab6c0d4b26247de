// pyramid.hip -- depth_buffer_pyramid + depth_buffer_reductor::copy_and_reduce (depth_buffer_pyramid.cpp:14-23,
// 177-299; depth_buffer_copy.comp; depth_buffer_reduce.comp): the Hi-Z max pyramid of a D32 depth buffer.
//
// Level 0 is a copy of the depth buffer; level l + 1 has max(W >> (l + 1), 1) x max(H >> (l + 1), 1) texels and
// texel (x, y) is max(0, the level-l texels (2x + i, 2y + j) that lie inside level l), taken in the reference's
// order i = 0, 1 outer, j = 0, 1 inner with GLSL max(a, b) = a < b ? b : a. Starting from 0, that max only ever
// replaces the running value by a strictly larger one, so NaN and negative inputs never win, the order of the four
// texels does not matter, and a texel outside level l can be replaced by 0 (an odd last row or column of a level is
// not read by the next one). The reference dispatches one 32x32 pass per level with a barrier between levels; here
// one workgroup takes a 64x64 region of a level and writes the six levels below it (64 -> 1 texels), so the depth
// is read once and every level written once:
//   launch 1: depth -> level 0 (copy) and levels 1..6;  launch 2: level 6 -> levels 7..12;  launch 3: 12 -> 13..15.
// Each of the 256 threads loads a 4x4 block of the region (four 16-byte rows) and reduces it in registers to its
// 2x2 texels of the first level below and its one texel of the second; one wave finishes the last four levels
// from those 16x16 texels through LDS and lane shuffles. A texel outside its level is zero at every step, which
// is exactly the bounds rule above. A region's texels at every level below depend on that region alone (level
// sizes halve with the floor), so the workgroups are independent. Images of a batch are stacked one after another
// in the regions; an image's pyramid need not start 16-byte aligned (the rows are accessed through 4-byte-aligned vector
// types, which the compiler lowers to accesses legal at that alignment).
#include "common.hpp"

namespace {

constexpr int kRegion = 64;   // texels of the input level per workgroup side
constexpr int kThreads = 256; // 16 x 16 threads, a 4x4 block each
constexpr int kLevelsPerLaunch = 6;

__host__ __device__ inline uint32_t level_dim(uint32_t d, uint32_t l) { return (d >> l) > 0 ? (d >> l) : 1u; }

typedef float f32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float f32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));

struct PyramidLaunch {
    const float* src;          // the input level of image 0
    uint64_t src_image_stride; // floats between images of `src`
    float* pyr;                // the pyramid of image 0
    uint64_t pyr_image_stride;
    uint32_t W, H;             // level 0 size
    uint32_t lin;              // input level
    uint32_t n_out;            // output levels lin + 1 .. lin + n_out (0 .. 6)
    uint32_t copy;             // write the input to level lin (launch 1: the depth copy to level 0)
    uint32_t rx, ry;           // regions per row and per column of the input level
    uint32_t n_regions;        // rx * ry * images: the grid, one workgroup each
    uint64_t level_off[17];    // float offset of each level in one image's pyramid
};

__device__ __forceinline__ float glsl_max(float a, float b) { return a < b ? b : a; }

// max(0, a, b, c, d) as the reference's chain; every argument is already 0 outside its level
__device__ __forceinline__ float max4(float a, float b, float c, float d)
{
    return glsl_max(glsl_max(glsl_max(glsl_max(0.0f, a), b), c), d);
}

__global__ __launch_bounds__(kThreads) void pyramid_kernel(const PyramidLaunch a)
{
    __shared__ float s_l2[16 * 16];
    const uint32_t tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    // Workgroups go to the eight XCDs in turn (workgroup b to XCD b mod 8). The regions are dealt out so that each
    // XCD takes one contiguous run of them: horizontally adjacent regions then share an L2, and a cache line of a
    // level that two regions write in part (every line of levels whose region rows are under 128 bytes, and every
    // line of an image whose pyramid does not start on a line) is merged there instead of being written back in
    // pieces from two L2s.
    const uint32_t b = blockIdx.x, q = a.n_regions / 8, rem = a.n_regions % 8, xcd = b % 8, slot = b / 8;
    const uint32_t region = (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + slot;
    const uint32_t img = region / (a.rx * a.ry), rxy = region - img * (a.rx * a.ry);
    const uint32_t rgx = rxy % a.rx, rgy = rxy / a.rx;
    const float* src = a.src + (uint64_t)img * a.src_image_stride;
    float* pyr = a.pyr + (uint64_t)img * a.pyr_image_stride;
    const uint32_t Wi = level_dim(a.W, a.lin), Hi = level_dim(a.H, a.lin);
    const uint32_t x0 = rgx * kRegion, y0 = rgy * kRegion;

    // the thread's 4x4 block of the input level, zero outside it
    float v[4][4];
    const uint32_t bx = x0 + 4 * tx, by = y0 + 4 * ty;
    float* copy_dst = a.copy ? pyr + a.level_off[a.lin] : nullptr;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const uint32_t gy = by + r;
        const uint64_t row = (uint64_t)gy * Wi;
        if (gy < Hi && bx + 3 < Wi) {
            const f32x4_a4 q = *(const f32x4_a4*)(src + row + bx);
            v[r][0] = q.x;
            v[r][1] = q.y;
            v[r][2] = q.z;
            v[r][3] = q.w;
            if (copy_dst) *(f32x4_a4*)(copy_dst + row + bx) = q;
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                v[r][c] = 0.0f;
                if (gy < Hi && bx + c < Wi) {
                    v[r][c] = src[row + bx + c];
                    if (copy_dst) copy_dst[row + bx + c] = v[r][c];
                }
            }
        }
    }
    if (a.n_out == 0) return;

    // first level below: the thread's 2x2 texels
    const uint32_t l1 = a.lin + 1;
    const uint32_t W1 = level_dim(a.W, l1), H1 = level_dim(a.H, l1);
    const uint32_t ox = (x0 >> 1) + 2 * tx, oy = (y0 >> 1) + 2 * ty;
    float m1[2][2];
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const float m = max4(v[2 * b][2 * c], v[2 * b + 1][2 * c], v[2 * b][2 * c + 1], v[2 * b + 1][2 * c + 1]);
            m1[b][c] = (ox + c < W1 && oy + b < H1) ? m : 0.0f;
        }
    {
        float* dst = pyr + a.level_off[l1];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            if (oy + b >= H1) continue;
            const uint64_t row = (uint64_t)(oy + b) * W1 + ox;
            if (ox + 1 < W1) {
                f32x2_a4 q;
                q.x = m1[b][0];
                q.y = m1[b][1];
                *(f32x2_a4*)(dst + row) = q;
            } else if (ox < W1) {
                dst[row] = m1[b][0];
            }
        }
    }
    if (a.n_out == 1) return;

    // second level below: the thread's one texel
    const uint32_t l2 = a.lin + 2;
    const uint32_t W2 = level_dim(a.W, l2), H2 = level_dim(a.H, l2);
    const uint32_t px = (x0 >> 2) + tx, py = (y0 >> 2) + ty;
    const bool in2 = px < W2 && py < H2;
    const float m2 = in2 ? max4(m1[0][0], m1[1][0], m1[0][1], m1[1][1]) : 0.0f;
    if (in2) pyr[a.level_off[l2] + (uint64_t)py * W2 + px] = m2;
    if (a.n_out == 2) return;
    s_l2[ty * 16 + tx] = m2;
    __syncthreads();
    if (tid >= VREN_WAVE) return;

    // the last levels in one wave: 8x8 texels (one per lane), then 4x4, 2x2 and 1x1 over lane shuffles
    const uint32_t lane = tid;
    uint32_t lx = lane % 8, ly = lane / 8;
    float m = max4(s_l2[(2 * ly) * 16 + 2 * lx], s_l2[(2 * ly + 1) * 16 + 2 * lx], s_l2[(2 * ly) * 16 + 2 * lx + 1],
                   s_l2[(2 * ly + 1) * 16 + 2 * lx + 1]);
    uint32_t side = 8; // texels per side of the region at level k
    for (uint32_t k = 3; k <= a.n_out; k++) {
        const uint32_t l = a.lin + k;
        const uint32_t Wl = level_dim(a.W, l), Hl = level_dim(a.H, l);
        if (k > 3) {
            // lane (lx, ly) of the side x side grid takes the four texels of the previous 2 side grid
            const uint32_t ps = 2 * side;
            const uint32_t sx = (lane % side) * 2, sy = (lane / side) * 2;
            const float t00 = shfl(m, (int)(sy * ps + sx)), t01 = shfl(m, (int)((sy + 1) * ps + sx));
            const float t10 = shfl(m, (int)(sy * ps + sx + 1)), t11 = shfl(m, (int)((sy + 1) * ps + sx + 1));
            m = max4(t00, t01, t10, t11);
            lx = lane % side;
            ly = lane / side;
        }
        const uint32_t gx = (x0 >> k) + lx, gy = (y0 >> k) + ly;
        const bool live = lane < side * side && gx < Wl && gy < Hl;
        m = live ? m : 0.0f;
        if (live) pyr[a.level_off[l] + (uint64_t)gy * Wl + gx] = m;
        side /= 2;
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
        a.rx = (Wi + kRegion - 1) / kRegion;
        a.ry = (Hi + kRegion - 1) / kRegion;
        const uint64_t n = (uint64_t)a.rx * a.ry * images;
        if (n >= (1ull << 31)) return vren_fail(ctx, VREN_EINVAL, "depth_pyramid: too many 64x64 regions");
        a.n_regions = (uint32_t)n;
        pyramid_kernel<<<a.n_regions, kThreads, 0, s>>>(a);
        VREN_LAUNCH_CHECK(ctx);
        if (lin + kLevelsPerLaunch >= levels) break;
    }
    return VREN_OK;
}

} // extern "C"
