// bounce.hip -- point-light animation ahead of the light BVH build: bounce_point_lights
// (vren_demo/resources/shaders/bounce_point_lights.comp:33-73, dispatched by point_light_bouncer.cpp:17-57).
//
// Every light first has its position clamped into the scene AABB, then moves speed * dt along its direction,
// reflecting each direction component whose wall it reaches first (a segment stops 1e-5 short of the wall), for
// at most 32 segments. One thread per light with 16-byte loads and stores; the float operations and their order
// are the shader's (IEEE division, no contraction), and GLSL min(x, y) = y < x ? y : x and
// max(x, y) = x < y ? y : x keep their comparison forms, so the results are bit-exact against the CPU
// restatement, infinities from zero direction components included. Like the shader, a light clamped onto a wall
// and moving outward is never reflected (both wall times of that axis are <= 0) and leaves the box, and a light
// on a wall with a zero direction component there gets 0 / 0 = NaN.
#include "common.hpp"

namespace {

constexpr float kInf = 1e35f;
constexpr float kEps = 1e-5f;
constexpr int kMaxSegments = 32;

__device__ __forceinline__ float glsl_min(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float glsl_max(float x, float y) { return x < y ? y : x; }

// time to the wall at `wall` along a direction component, INF when the wall is behind or reached
__device__ __forceinline__ float wall_time(float wall, float p, float d)
{
    const float t = (wall - p) / d;
    return t <= 0.0f ? kInf : t;
}

__global__ __launch_bounds__(256) void bounce_kernel(float4* __restrict__ positions, float4* __restrict__ directions,
                                                     uint32_t n, float lx, float ly, float lz, float hx, float hy,
                                                     float hz, float speed, float dt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p4 = positions[i];
    const float4 d4 = directions[i];
    float px = glsl_min(glsl_max(p4.x, lx), hx);
    float py = glsl_min(glsl_max(p4.y, ly), hy);
    float pz = glsl_min(glsl_max(p4.z, lz), hz);
    float dx = d4.x, dy = d4.y, dz = d4.z;
    float rem = speed * dt;
    for (int j = 0; j < kMaxSegments && rem > 0.0f; j++) {
        const float tx1 = wall_time(lx, px, dx), tx2 = wall_time(hx, px, dx);
        const float ty1 = wall_time(ly, py, dy), ty2 = wall_time(hy, py, dy);
        const float tz1 = wall_time(lz, pz, dz), tz2 = wall_time(hz, pz, dz);
        const float min_t = glsl_min(tx1, glsl_min(tx2, glsl_min(ty1, glsl_min(ty2, glsl_min(tz1, tz2)))));
        const float step = glsl_min(min_t - kEps, rem);
        px = px + step * dx;
        py = py + step * dy;
        pz = pz + step * dz;
        if (min_t < rem) {
            dx = dx * ((tx1 == min_t || tx2 == min_t) ? -1.0f : 1.0f);
            dy = dy * ((ty1 == min_t || ty2 == min_t) ? -1.0f : 1.0f);
            dz = dz * ((tz1 == min_t || tz2 == min_t) ? -1.0f : 1.0f);
        }
        rem = rem - step;
    }
    positions[i] = make_float4(px, py, pz, 1.0f);
    directions[i] = make_float4(dx, dy, dz, 0.0f);
}

} // namespace

extern "C" VREN_EXPORT int vren_bounce_point_lights(vren_ctx* ctx, vren_stream_t stream, float* positions,
                                                    float* directions, uint32_t n, const float* aabb_min,
                                                    const float* aabb_max, float speed, float dt)
{
    if (!ctx) return VREN_EINVAL;
    if (!aabb_min || !aabb_max || (n > 0 && (!positions || !directions)))
        return vren_fail(ctx, VREN_EINVAL, "bounce_point_lights: null buffer or AABB");
    if (((uintptr_t)positions | (uintptr_t)directions) & 15)
        return vren_fail(ctx, VREN_EINVAL, "bounce_point_lights: vec4 buffers must be 16-byte aligned");
    if (ctx->device < 0) return vren_fail(ctx, VREN_EINVAL, "bounce_point_lights: context has no device");
    if (n == 0) return VREN_OK;
    bounce_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(directions), n, aabb_min[0], aabb_min[1],
        aabb_min[2], aabb_max[0], aabb_max[1], aabb_max[2], speed, dt);
    VREN_LAUNCH_CHECK(ctx);
    return VREN_OK;
}
