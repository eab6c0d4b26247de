/*
 * vren_hip/synth.h -- definition of the synthetic inputs the clustered-shading path is measured and tested on.
 *
 * The reference takes its inputs from a rasterised glTF scene (vren_demo) and tests its primitives on unseeded
 * rand() data (vren_test/vren_test/gpu_test_bench.hpp:14-141). This build replaces both with seeded, position-
 * addressable generators so that a GPU kernel and a CPU loop produce the same bytes:
 *   - PCG32 (PCG-XSH-RR 64/32, O'Neill 2014) with the standard srandom seeding, plus LCG jump-ahead
 *     (Brown 1994, "Random number generation with arbitrary strides") so element i can be generated
 *     without generating elements 0..i-1;
 *   - unit floats (u >> 8) * 2^-24, log-uniform via exp(lerp(ln a, ln b, U)) with vren_expf;
 *   - point lights (positions U[-50,50]^3, radius U[1,8], colour U[0,1]^3, intensity logU[0.1,100]
 *     pre-multiplied into the colour, radius stored in the intensity slot as in gpu_repr.hpp:61-64);
 *   - a procedural scene of axis-aligned boxes plus the ground plane y = -10 ray-cast per pixel in double
 *     precision, giving the D32 depth (camera.cpp:40-50 projection) and RGBA16F face normals
 *     (gbuffer.cpp:7-13) the cluster-key stage reads;
 *   - per-pixel resolved materials (albedo U[0,1]^3, roughness U[0.05,1], metallic Bernoulli(0.3)).
 * Everything here uses only +, -, *, / (correctly rounded) and comparisons, so the GPU generator in the product
 * library and the CPU generator in the oracle emit identical bytes when both compile without contraction.
 */
#ifndef VREN_HIP_SYNTH_H_
#define VREN_HIP_SYNTH_H_

#include <stdint.h>
#include "vmath.h"

#define VREN_PCG_MULT 6364136223846793005ULL

/* stream ids (PCG initseq) of the generators; initstate = 42 + 1000 * config (+ frame) */
#define VREN_SYNTH_SEQ_KEYS 54u
#define VREN_SYNTH_SEQ_LIGHTS 54u
#define VREN_SYNTH_SEQ_BOXES 55u
#define VREN_SYNTH_SEQ_MATERIALS 56u
#define VREN_SYNTH_SEQ_PERTURB 57u

#define VREN_SYNTH_DRAWS_PER_LIGHT 8u
#define VREN_SYNTH_DRAWS_PER_BOX 6u
#define VREN_SYNTH_DRAWS_PER_PIXEL 5u
#define VREN_SYNTH_MAX_BOXES 256u

typedef struct vren_pcg32 {
    uint64_t state;
    uint64_t inc;
} vren_pcg32;

VREN_HD uint32_t vren_pcg32_next(vren_pcg32* r)
{
    uint64_t old = r->state;
    r->state = old * VREN_PCG_MULT + r->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

VREN_HD vren_pcg32 vren_pcg32_seed(uint64_t initstate, uint64_t initseq)
{
    vren_pcg32 r;
    r.state = 0u;
    r.inc = (initseq << 1u) | 1u;
    vren_pcg32_next(&r);
    r.state += initstate;
    vren_pcg32_next(&r);
    return r;
}

/* advance the generator by `delta` steps in O(log delta) */
VREN_HD void vren_pcg32_advance(vren_pcg32* r, uint64_t delta)
{
    uint64_t cur_mult = VREN_PCG_MULT;
    uint64_t cur_plus = r->inc;
    uint64_t acc_mult = 1u;
    uint64_t acc_plus = 0u;
    while (delta > 0u) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    r->state = acc_mult * r->state + acc_plus;
}

/* generator positioned at draw `index` of stream (initstate, initseq) */
VREN_HD vren_pcg32 vren_pcg32_at(uint64_t initstate, uint64_t initseq, uint64_t index)
{
    vren_pcg32 r = vren_pcg32_seed(initstate, initseq);
    vren_pcg32_advance(&r, index);
    return r;
}

VREN_HD float vren_unit_f32(uint32_t u)
{
    return (float)(u >> 8) * 5.9604644775390625e-08f; /* 2^-24 */
}

VREN_HD float vren_uniform_f32(uint32_t u, float lo, float hi)
{
    return lo + (hi - lo) * vren_unit_f32(u);
}

VREN_HD float vren_loguniform_f32(uint32_t u, float lo, float hi)
{
    float a = vren_logf(lo);
    float b = vren_logf(hi);
    return vren_expf(a + (b - a) * vren_unit_f32(u));
}

/* Point light `i` of the light generator: pos4 = (x, y, z, 1); light4 = (colour * intensity, radius). */
VREN_HD void vren_synth_light(uint64_t initstate, uint32_t i, float pos4[4], float light4[4])
{
    vren_pcg32 r = vren_pcg32_at(initstate, VREN_SYNTH_SEQ_LIGHTS, (uint64_t)i * VREN_SYNTH_DRAWS_PER_LIGHT);
    uint32_t u[VREN_SYNTH_DRAWS_PER_LIGHT];
    for (uint32_t k = 0; k < VREN_SYNTH_DRAWS_PER_LIGHT; k++) {
        u[k] = vren_pcg32_next(&r);
    }
    pos4[0] = vren_uniform_f32(u[0], -50.0f, 50.0f);
    pos4[1] = vren_uniform_f32(u[1], -50.0f, 50.0f);
    pos4[2] = vren_uniform_f32(u[2], -50.0f, 50.0f);
    pos4[3] = 1.0f;
    float radius = vren_uniform_f32(u[3], 1.0f, 8.0f);
    float intensity = vren_loguniform_f32(u[7], 0.1f, 100.0f);
    light4[0] = vren_unit_f32(u[4]) * intensity;
    light4[1] = vren_unit_f32(u[5]) * intensity;
    light4[2] = vren_unit_f32(u[6]) * intensity;
    light4[3] = radius;
}

/* Scene box `b` in world space, from the camera basis. Centres are placed inside the view frustum at view
 * depth U[10,95]; half-extents U[1,8] per axis. `inv_view` is the double inverse of the view matrix
 * (column-major), `tan_half_fov` and `aspect` are the camera's. Output: center[3], half[3] (double). */
VREN_HD void vren_synth_box(uint64_t initstate, uint32_t b, const double inv_view[16], double tan_half_fov,
                            double aspect, double center[3], double half[3])
{
    vren_pcg32 r = vren_pcg32_at(initstate, VREN_SYNTH_SEQ_BOXES, (uint64_t)b * VREN_SYNTH_DRAWS_PER_BOX);
    uint32_t u[VREN_SYNTH_DRAWS_PER_BOX];
    for (uint32_t k = 0; k < VREN_SYNTH_DRAWS_PER_BOX; k++) {
        u[k] = vren_pcg32_next(&r);
    }
    double z = 10.0 + 85.0 * (double)vren_unit_f32(u[0]);
    double x = (2.0 * (double)vren_unit_f32(u[1]) - 1.0) * z * tan_half_fov * aspect * 0.9;
    double y = (2.0 * (double)vren_unit_f32(u[2]) - 1.0) * z * tan_half_fov * 0.9;
    double v[4] = {x, y, z, 1.0};
    for (int i = 0; i < 3; i++) {
        center[i] = ((inv_view[0 * 4 + i] * v[0] + inv_view[1 * 4 + i] * v[1]) + inv_view[2 * 4 + i] * v[2]) +
                    inv_view[3 * 4 + i] * v[3];
    }
    half[0] = 1.0 + 7.0 * (double)vren_unit_f32(u[3]);
    half[1] = 1.0 + 7.0 * (double)vren_unit_f32(u[4]);
    half[2] = 1.0 + 7.0 * (double)vren_unit_f32(u[5]);
}

VREN_HD uint16_t vren_float_to_half_axis(double v)
{
    /* only exact values occur: -1, 0, +1 */
    if (v > 0.5) {
        return 0x3C00u;
    }
    if (v < -0.5) {
        return 0xBC00u;
    }
    return 0u;
}

/*
 * Ray-cast pixel (x, y) of a W x H frame. The view ray goes through the same clip coordinates the reference
 * reconstructs positions from (shade.comp:75-80: u = x / W, v = y / H, no half-pixel offset), so the surface
 * it reports is the one the shade stage un-projects. Boxes: centers/halves (world space, n_boxes); ground
 * plane y = -10. Outputs the D32 NDC depth (1.0 = miss / beyond far) and the face normal as RGBA16F bits.
 * `cam[0..2]` = camera position, `view_rot_inv` = inverse view matrix (double, column-major),
 * `proj_c`, `proj_d` = projection m[2][2] and m[3][2] (double copies of the float matrix entries).
 */
VREN_HD void vren_synth_pixel(uint32_t x, uint32_t y, uint32_t W, uint32_t H, const double inv_view[16],
                              double tan_half_fov, double aspect, double near_plane, double far_plane,
                              double proj_c, double proj_d, const double* centers, const double* halves,
                              uint32_t n_boxes, float* depth_out, uint16_t normal_out[4])
{
    double u = (double)x / (double)W;
    double v = (double)y / (double)H;
    double cx = u * 2.0 - 1.0;
    double cy = (1.0 - v) * 2.0 - 1.0;
    double dv[3] = {cx * tan_half_fov * aspect, cy * tan_half_fov, 1.0};
    double o[3] = {inv_view[12], inv_view[13], inv_view[14]};
    double d[3];
    for (int i = 0; i < 3; i++) {
        d[i] = (inv_view[0 * 4 + i] * dv[0] + inv_view[1 * 4 + i] * dv[1]) + inv_view[2 * 4 + i] * dv[2];
    }
    double best_t = 1e300;
    int best_axis = -1;
    double best_sign = 0.0;
    for (uint32_t b = 0; b < n_boxes; b++) {
        double t_enter = -1e300, t_exit = 1e300;
        int enter_axis = -1;
        double enter_sign = 0.0;
        int hit = 1;
        for (int a = 0; a < 3; a++) {
            double lo = centers[3 * b + a] - halves[3 * b + a];
            double hi = centers[3 * b + a] + halves[3 * b + a];
            if (d[a] == 0.0) {
                if (o[a] < lo || o[a] > hi) {
                    hit = 0;
                }
                continue;
            }
            double t0 = (lo - o[a]) / d[a];
            double t1 = (hi - o[a]) / d[a];
            double s = -1.0; /* entering through the lo face: normal points to -axis */
            if (t0 > t1) {
                double tmp = t0;
                t0 = t1;
                t1 = tmp;
                s = 1.0;
            }
            if (t0 > t_enter) {
                t_enter = t0;
                enter_axis = a;
                enter_sign = s;
            }
            if (t1 < t_exit) {
                t_exit = t1;
            }
        }
        if (hit && enter_axis >= 0 && t_enter <= t_exit && t_enter > 0.0 && t_enter < best_t) {
            best_t = t_enter;
            best_axis = enter_axis;
            best_sign = enter_sign;
        }
    }
    if (d[1] < 0.0) {
        double t = (-10.0 - o[1]) / d[1];
        if (t > 0.0 && t < best_t) {
            best_t = t;
            best_axis = 1;
            best_sign = 1.0;
        }
    }
    /* view depth equals t because the view-space ray direction has z = 1 */
    normal_out[0] = normal_out[1] = normal_out[2] = normal_out[3] = 0u;
    if (best_axis < 0 || best_t < near_plane || best_t > far_plane) {
        *depth_out = 1.0f;
        return;
    }
    double ndc = proj_c + proj_d / best_t;
    float df = (float)ndc;
    if (df > 1.0f) {
        df = 1.0f;
    }
    if (df < 0.0f) {
        df = 0.0f;
    }
    *depth_out = df;
    double n[3] = {0.0, 0.0, 0.0};
    n[best_axis] = best_sign;
    normal_out[0] = vren_float_to_half_axis(n[0]);
    normal_out[1] = vren_float_to_half_axis(n[1]);
    normal_out[2] = vren_float_to_half_axis(n[2]);
}

/* Per-pixel resolved material p: albedo_roughness = (r, g, b, roughness), metallic in {0, 1}. */
VREN_HD void vren_synth_material(uint64_t initstate, uint64_t p, float albedo_roughness[4], float* metallic)
{
    vren_pcg32 r = vren_pcg32_at(initstate, VREN_SYNTH_SEQ_MATERIALS, p * VREN_SYNTH_DRAWS_PER_PIXEL);
    uint32_t u[VREN_SYNTH_DRAWS_PER_PIXEL];
    for (uint32_t k = 0; k < VREN_SYNTH_DRAWS_PER_PIXEL; k++) {
        u[k] = vren_pcg32_next(&r);
    }
    albedo_roughness[0] = vren_unit_f32(u[0]);
    albedo_roughness[1] = vren_unit_f32(u[1]);
    albedo_roughness[2] = vren_unit_f32(u[2]);
    albedo_roughness[3] = vren_uniform_f32(u[3], 0.05f, 1.0f);
    *metallic = vren_unit_f32(u[4]) < 0.3f ? 1.0f : 0.0f;
}

#endif /* VREN_HIP_SYNTH_H_ */
