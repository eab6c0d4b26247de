/*
 * oracle.cpp -- CPU restatement of the reference's clustered-shading compute path and its primitives.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or calls this file; only tests/,
 * the smoke check and bench.py's cpu_baseline leg load the shared object built from it, and only as the
 * checker (or the timed CPU baseline), never as the thing measured on the GPU.
 *
 * Each function names the reference file:line it restates. Build: oracle/build.py (g++ -O2
 * -ffp-contract=off, no fast-math), output oracle/_build/liboracle.so.
 *
 * Pins (what checks each function independently of this file; see tests/):
 *   o_pcg32_*            published PCG32 demo vector (seed 42, seq 54) + jump-ahead == sequential
 *   o_logf_first_decrease  a planted glitch is reported at exactly its bit pattern
 *   o_rcp_f16_newton_mismatches  the planted variant (no Newton step) is reported; exact rational arithmetic
 *                        agrees (the FMA and the division here are correctly rounded by the C library)
 *   o_third_index_mismatches  a planted threshold one float off is reported
 *   o_reduce_tree_*      closed form tree[i] = op over the aligned block ending at i; totals vs numpy
 *   o_exclusive_scan_u32 numpy cumsum (uint32 wrap-around)
 *   o_stable_sort_kv     numpy argsort(kind='stable')
 *   o_bucket_sort        numpy argsort(kind='stable') on the bucket id
 *   o_calc_bvh_*         build_bvh.utils known answers (vren_test/vren_test/primitives/build_bvh.cpp:223-254)
 *   o_build_bvh          numpy min/max over each group of 32 valid children; point-query property test
 *   o_build_lbvh         radix-tree property of every internal node (its leaves are a contiguous key range split
 *                        at the first differing bit), exact child unions, sphere queries vs brute force
 *   o_camera_matrices    closed forms of camera.cpp:32-50 for the fixture cameras
 *   o_find_unique_clusters   numpy sort/unique per tile, refs invariant keys[ref[p]] == key(p), key(p) of
 *                        sampled pixels recomputed from the formula
 *   o_assign_lights      an independent ancestor-predicate brute force (every leaf whose ancestors all pass
 *                        the node test and which passes the leaf test); plain brute force in TIGHT mode
 *   o_shade_f64          closed forms on hand-built single-light scenes; agreement with the float
 *                        restatement; finite differences of the smooth-attenuation forward
 *   o_shade_backward_f64 FP64 central differences of o_shade_loss_f64
 *   o_bounce_point_lights  closed forms: unobstructed motion p + speed dt d, clamping of outside positions, and
 *                        the unfolded billiard path (1-D and 2-D, several reflections) to within the 1e-5 wall
 *                        offsets; |d| per component preserved, final positions inside the AABB
 *   o_blelloch_downsweep_u32  of an up-sweep tree: numpy exclusive cumsum (root kept when clear_last is false);
 *                        on arbitrary input: the closed form P + sum over the index's set bits of the left
 *                        sibling's last element
 *   o_depth_pyramid      numpy 2x2 max over the in-bounds texels of each level, brute force on odd sizes
 *   o_resolve_materials_f64, o_shade_f64_materials  hand-made textures with known texels for each filter and
 *                        address mode; the material path against o_shade_f64 on the resolved values
 *   o_show_clusters, o_visualize_bvh  closed forms of each view mode; the 24 vertices of a node are its AABB's
 *                        12 edges
 *   o_cpu_*              timing helpers of bench.py's CPU baseline (std::stable_sort / std::exclusive_scan /
 *                        std::accumulate on 1 or N std::threads); their results are checked against the serial ones
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <chrono>
#include <numeric>
#include <thread>
#include <vector>

#include "../include/vren_hip/capi.h"
#include "../include/vren_hip/sampler.h"
#include "../include/vren_hip/synth.h"
#include "../include/vren_hip/vmath.h"

#define OEXPORT extern "C" __attribute__((visibility("default")))

namespace {

const float k_pi_ref = 3.14f;       /* common.glsl:7 */
const float k_epsilon_ref = 0.001f; /* common.glsl:8 */

uint32_t next_pow2(uint32_t v)
{
    /* base.hpp:37-47 */
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    v++;
    return v;
}

uint32_t ilog2(uint64_t v)
{
    uint32_t r = 0;
    while (v > 1) {
        v >>= 1;
        r++;
    }
    return r;
}

/* GLSL min / max (x, y) */
inline float glsl_min(float x, float y) { return y < x ? y : x; }
inline float glsl_max(float x, float y) { return x < y ? y : x; }

/* min / max in the total order -0 < +0 (order independent; used for BVH AABBs) */
inline float tmin(float a, float b)
{
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
}
inline float tmax(float a, float b)
{
    if (a > b) return a;
    if (b > a) return b;
    return std::signbit(a) ? b : a;
}

struct vec4f {
    float v[4];
};

} // namespace

/* ------------------------------------------------------------------------------------------------------------
 * RNG
 * ---------------------------------------------------------------------------------------------------------- */

/* sequential PCG32 draws [start, start + n) of stream (initstate, initseq) -- no jump-ahead */
OEXPORT void o_pcg32_fill(uint64_t initstate, uint64_t initseq, uint64_t start, uint32_t* out, uint64_t n)
{
    vren_pcg32 r = vren_pcg32_seed(initstate, initseq);
    for (uint64_t i = 0; i < start; i++) vren_pcg32_next(&r);
    for (uint64_t i = 0; i < n; i++) out[i] = vren_pcg32_next(&r);
}

/* same draws through the jump-ahead used by the GPU generators */
OEXPORT void o_pcg32_fill_at(uint64_t initstate, uint64_t initseq, uint64_t start, uint32_t* out, uint64_t n)
{
    for (uint64_t i = 0; i < n; i++) {
        vren_pcg32 r = vren_pcg32_at(initstate, initseq, start + i);
        out[i] = vren_pcg32_next(&r);
    }
}

OEXPORT float o_logf(float x) { return vren_logf(x); }

/* First float bit pattern b in [lo, hi) with vren_logf(b + 1) < vren_logf(b) (0xFFFFFFFF when there is none):
 * the GPU depth-slice threshold table relies on vren_logf being non-decreasing over the non-negative floats.
 * planted != 0 raises the value at bit pattern `planted` by 1e-4 relative (far more than one step of the log
 * between neighbouring floats), a decrease the scan must report there. */
static float logf_scanned(uint32_t b, uint32_t planted)
{
    const float v = vren_logf(vren_bits_f32(b));
    return (planted && b == planted) ? v + std::fabs(v) * 1e-4f + 1e-30f : v;
}

OEXPORT uint32_t o_logf_first_decrease(uint32_t lo, uint32_t hi, uint32_t planted)
{
    float prev = logf_scanned(lo, planted);
    for (uint32_t b = lo; b < hi; b++) {
        const float next = logf_scanned(b + 1, planted);
        if (next < prev) return b;
        prev = next;
    }
    return 0xFFFFFFFFu;
}

/* The cluster kernel's 1 / a for a (> 0) a finite f16 normal component widened to float: a reciprocal within
 * 1 ulp of 1 / a (every candidate: the correctly rounded value and its two neighbours) refined by one FMA Newton
 * step. Returns how many (a, candidate) pairs do not land on the correctly rounded 1 / a (IEEE division here);
 * `planted` skips the Newton step, which must then be caught. */
OEXPORT uint32_t o_rcp_f16_newton_mismatches(uint32_t planted)
{
    uint32_t bad = 0;
    for (uint32_t h = 1; h < 0x7C00u; h++) {
        const float a = vren_half_to_float((uint16_t)h);
        const float r = 1.0f / a;
        const float cand[3] = {std::nextafter(r, 0.0f), r, std::nextafter(r, INFINITY)};
        for (float y : cand) {
            const float y1 = planted ? y : std::fma(y, std::fma(-a, y, 1.0f), y);
            if (y1 != r) bad++;
        }
    }
    return bad;
}

/* Bit patterns b in [lo, hi) whose float u (< 1.5 or NaN) gets a different cell index from the comparison form
 * (vren_third_index) than from the formula; `planted` moves the first threshold up one float in the compared
 * copy, which must then be reported. */
OEXPORT uint32_t o_third_index_mismatches(uint64_t lo, uint64_t hi, uint32_t planted)
{
    const float t1 = planted ? std::nextafter(-0x1.555554p-2f, 1.0f) : -0x1.555554p-2f;
    uint32_t bad = 0;
    for (uint64_t b = lo; b < hi; b++) {
        const float u = vren_bits_f32((uint32_t)b);
        if (u >= 1.5f) continue;
        const uint32_t formula = vren_f2u_sat(vren_floorf((u + 1.0f) / 2.0f * 3.0f));
        const uint32_t cmp = planted ? (uint32_t)(u >= t1) + (uint32_t)(u >= 0x1.555556p-2f) + (uint32_t)(u >= 0x1.fffffep-1f)
                                     : vren_third_index(u);
        if (formula != cmp) bad++;
    }
    return bad;
}

OEXPORT float o_expf(float x) { return vren_expf(x); }
OEXPORT float o_powi(float a, uint32_t k) { return vren_powi(a, k); }

/* ------------------------------------------------------------------------------------------------------------
 * Reduce (reduce.cpp:32-114 dispatch loop + reduce.comp:50-87 workgroup up-sweep), restated dispatch by
 * dispatch: each dispatch at base level b reads elements offset + g * stride (offset = 2^b - 1, stride = 2^b),
 * up-sweeps 10 levels inside each group of 1024, and writes the group back to the same positions.
 * ---------------------------------------------------------------------------------------------------------- */

template <typename T, typename Op>
static void reduce_tree(const T* in, uint32_t n, T* out, Op op, T identity, uint32_t blocks)
{
    const uint32_t wg = 1024;
    uint32_t p = next_pow2(n);
    uint32_t levels = std::max<uint32_t>(ilog2(p), 1u);
    std::vector<T> buf(p);
    std::vector<T> s(wg);
    for (uint32_t b = 0; b < blocks; b++) {
        for (uint32_t level = 0; level < levels; level += 10) {
            uint64_t offset = (1ull << level) - 1;
            uint64_t stride = 1ull << level;
            uint64_t threads = 1ull << (levels - level);
            uint64_t groups = (threads + wg - 1) / wg;
            uint64_t in_len = level == 0 ? n : p;
            for (uint64_t g = 0; g < groups; g++) {
                for (uint32_t t = 0; t < wg; t++) {
                    uint64_t idx = offset + (g * wg + t) * stride;
                    if (idx < in_len) {
                        s[t] = level == 0 ? in[(uint64_t)b * n + idx] : buf[idx];
                    } else {
                        s[t] = identity;
                    }
                }
                for (uint32_t l = 0; l < 10; l++) {
                    uint32_t mask = (1u << (l + 1)) - 1;
                    for (uint32_t t = 0; t < wg; t++) {
                        if ((t & mask) == mask) {
                            s[t] = op(s[t], s[t - (1u << l)]); /* OP(s[b], s[a]), reduce.comp:76 */
                        }
                    }
                }
                for (uint32_t t = 0; t < wg; t++) {
                    uint64_t idx = offset + (g * wg + t) * stride;
                    if (idx < p) buf[idx] = s[t];
                }
            }
        }
        std::copy(buf.begin(), buf.end(), out + (uint64_t)b * p);
    }
}

OEXPORT void o_reduce_tree_u32(const uint32_t* in, uint32_t n, uint32_t* out, int op, uint32_t blocks)
{
    switch (op) {
    case VREN_REDUCE_ADD:
        reduce_tree<uint32_t>(in, n, out, [](uint32_t a, uint32_t b) { return a + b; }, 0u, blocks);
        break;
    case VREN_REDUCE_MIN:
        reduce_tree<uint32_t>(in, n, out, [](uint32_t a, uint32_t b) { return b < a ? b : a; }, 0xFFFFFFFFu, blocks);
        break;
    default:
        reduce_tree<uint32_t>(in, n, out, [](uint32_t a, uint32_t b) { return a < b ? b : a; }, 0u, blocks);
        break;
    }
}

OEXPORT void o_reduce_tree_vec4(const float* in, uint32_t n, float* out, int op, uint32_t blocks)
{
    const vec4f* vin = reinterpret_cast<const vec4f*>(in);
    vec4f* vout = reinterpret_cast<vec4f*>(out);
    switch (op) {
    case VREN_REDUCE_ADD:
        reduce_tree<vec4f>(vin, n, vout, [](vec4f a, vec4f b) {
            vec4f r;
            for (int i = 0; i < 4; i++) r.v[i] = a.v[i] + b.v[i];
            return r;
        }, vec4f{{0.f, 0.f, 0.f, 0.f}}, blocks);
        break;
    case VREN_REDUCE_MIN:
        reduce_tree<vec4f>(vin, n, vout, [](vec4f a, vec4f b) {
            vec4f r;
            for (int i = 0; i < 4; i++) r.v[i] = glsl_min(a.v[i], b.v[i]);
            return r;
        }, vec4f{{1e35f, 1e35f, 1e35f, 1e35f}}, blocks);
        break;
    default:
        reduce_tree<vec4f>(vin, n, vout, [](vec4f a, vec4f b) {
            vec4f r;
            for (int i = 0; i < 4; i++) r.v[i] = glsl_max(a.v[i], b.v[i]);
            return r;
        }, vec4f{{-1e35f, -1e35f, -1e35f, -1e35f}}, blocks);
        break;
    }
}

/* std::accumulate / min_element / max_element (the reference test's totals, reduce.cpp:72-98) */
OEXPORT uint32_t o_reduce_total_u32(const uint32_t* in, uint64_t n, int op)
{
    if (op == VREN_REDUCE_ADD) return std::accumulate(in, in + n, 0u);
    if (op == VREN_REDUCE_MIN) return n ? *std::min_element(in, in + n) : 0xFFFFFFFFu;
    return n ? *std::max_element(in, in + n) : 0u;
}

/* ------------------------------------------------------------------------------------------------------------
 * Scan / sorts
 * ---------------------------------------------------------------------------------------------------------- */

/* std::exclusive_scan (blelloch_scan.cpp:70-73, 135), per segment */
OEXPORT void o_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t blocks)
{
    for (uint32_t b = 0; b < blocks; b++) {
        std::exclusive_scan(in + (uint64_t)b * n, in + (uint64_t)(b + 1) * n, out + (uint64_t)b * n, 0u);
    }
}

/* blelloch_scan::downsweep (blelloch_scan.cpp:57-139, blelloch_scan_downsweep.comp:33-125), restated level by
 * level as the reference dispatches it, in place on `blocks` segments of `length` (a power of 2) elements:
 * global levels log2(length) - 1 down to 10 (the first one clears the segment's last element when clear_last),
 * then the 1024-thread LDS pass over each 1024-element chunk, whose shared array is zero-padded past the segment
 * and whose last element is not loaded (reads 0) when clear_last is still set. */
OEXPORT void o_blelloch_downsweep_u32(uint32_t* data, uint32_t length, uint32_t blocks, int clear_last_in)
{
    const uint32_t wg = 1024;
    for (uint32_t b = 0; b < blocks; b++) {
        uint32_t* d = data + (uint64_t)b * length;
        int clear_last = clear_last_in;
        int log_len = (int)ilog2(length);
        for (int level = log_len - 1; level > 9; level--) {
            uint32_t stride = 1u << level, offset = stride - 1;
            for (uint64_t x = 0;; x++) {
                uint64_t a = offset + x * stride * 2ull, bi = a + stride;
                if (bi >= length) break;
                uint32_t bv = (clear_last && bi == length - 1) ? 0u : d[bi];
                d[bi] = d[a] + bv;
                d[a] = bv;
            }
            clear_last = 0;
        }
        uint32_t groups = (length + wg - 1) / wg;
        std::vector<uint32_t> sh(wg);
        for (uint32_t g = 0; g < groups; g++) {
            for (uint32_t t = 0; t < wg; t++) {
                uint64_t idx = (uint64_t)g * wg + t;
                sh[t] = idx < (uint64_t)length - (clear_last ? 1u : 0u) ? d[idx] : 0u;
            }
            for (int level = 9; level >= 0; level--) {
                uint32_t mask = (1u << (level + 1)) - 1;
                for (uint32_t t = 0; t < wg; t++) {
                    if ((t & mask) != mask) continue;
                    uint32_t bi = t, a = bi - (1u << level);
                    uint32_t tmp = sh[bi];
                    sh[bi] = sh[a] + tmp;
                    sh[a] = tmp;
                }
            }
            for (uint32_t t = 0; t < wg; t++) {
                uint64_t idx = (uint64_t)g * wg + t;
                if (idx < length) d[idx] = sh[t];
            }
        }
    }
}

/* std::stable_sort by key bits [begin_bit, end_bit) (radix_sort.cpp:88 uses std::sort on keys only) */
OEXPORT void o_stable_sort_kv(uint32_t* keys, uint32_t* vals, uint64_t n, uint32_t begin_bit, uint32_t end_bit)
{
    uint32_t width = end_bit - begin_bit;
    uint32_t mask = width >= 32 ? 0xFFFFFFFFu : ((1u << width) - 1u);
    std::vector<uint64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0ull);
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
        return ((keys[a] >> begin_bit) & mask) < ((keys[b] >> begin_bit) & mask);
    });
    std::vector<uint32_t> k2(n), v2(vals ? n : 0);
    for (uint64_t i = 0; i < n; i++) {
        k2[i] = keys[idx[i]];
        if (vals) v2[i] = vals[idx[i]];
    }
    std::copy(k2.begin(), k2.end(), keys);
    if (vals) std::copy(v2.begin(), v2.end(), vals);
}

/* bucket_sort (bucket_sort.cpp:72-161): counting sort by (key >> shift) & (n_buckets - 1), stable */
OEXPORT void o_bucket_sort(const uint32_t* in_pairs, uint32_t n, uint32_t* out_pairs, uint32_t* offsets,
                           uint32_t n_buckets, uint32_t shift)
{
    std::vector<uint32_t> count(n_buckets, 0u);
    for (uint32_t i = 0; i < n; i++) count[(in_pairs[2 * i] >> shift) & (n_buckets - 1)]++;
    std::exclusive_scan(count.begin(), count.end(), offsets, 0u);
    std::vector<uint32_t> cursor(offsets, offsets + n_buckets);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t b = (in_pairs[2 * i] >> shift) & (n_buckets - 1);
        uint32_t dst = cursor[b]++;
        out_pairs[2 * dst] = in_pairs[2 * i];
        out_pairs[2 * dst + 1] = in_pairs[2 * i + 1];
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * BVH (build_bvh.cpp:101-136 sizes, build_bvh.comp:32-55 levels)
 * ---------------------------------------------------------------------------------------------------------- */

OEXPORT uint32_t o_calc_bvh_padded_leaf_count(uint32_t leaf_count)
{
    if (leaf_count <= 1) return 32u;
    uint64_t p = 32;
    while (p < leaf_count) p *= 32;
    return (uint32_t)p;
}

OEXPORT uint32_t o_calc_bvh_buffer_length(uint32_t leaf_count)
{
    uint32_t p = o_calc_bvh_padded_leaf_count(leaf_count);
    uint32_t len = 0;
    while (p != 0) {
        len += p;
        p >>= 5;
    }
    return len;
}

OEXPORT uint32_t o_calc_bvh_root_index(uint32_t leaf_count) { return o_calc_bvh_buffer_length(leaf_count) - 1; }

OEXPORT uint32_t o_calc_bvh_level_count(uint32_t leaf_count)
{
    return ilog2(o_calc_bvh_padded_leaf_count(leaf_count)) / 5;
}

static void set_empty(vren_bvh_node& n)
{
    n.min[0] = n.min[1] = n.min[2] = INFINITY;
    n.max[0] = n.max[1] = n.max[2] = -INFINITY;
}

OEXPORT void o_build_bvh(vren_bvh_node* nodes, uint32_t padded_leaf_count)
{
    uint32_t src = 0, dst = 0, level_nodes = padded_leaf_count;
    while (level_nodes > 1) {
        src = dst;
        dst = src + level_nodes;
        for (uint32_t g = 0; g < level_nodes / 32; g++) {
            vren_bvh_node parent;
            set_empty(parent);
            parent.pad = 0;
            bool any = false;
            for (uint32_t c = 0; c < 32; c++) {
                const vren_bvh_node& ch = nodes[src + g * 32 + c];
                if (ch.next == VREN_BVH_INVALID_NODE) continue;
                any = true;
                for (int a = 0; a < 3; a++) {
                    parent.min[a] = tmin(parent.min[a], ch.min[a]);
                    parent.max[a] = tmax(parent.max[a], ch.max[a]);
                }
            }
            parent.next = any ? src + g * 32 : VREN_BVH_INVALID_NODE;
            nodes[dst + g] = parent;
        }
        level_nodes >>= 5;
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * Camera (camera.cpp:9-50). Rotations are composed in double and rounded once; for yaw = pitch = 0 the view
 * matrix is the exact translation and the projection is the float expression of camera.cpp:40-50.
 * ---------------------------------------------------------------------------------------------------------- */

static void mat4_mul_d(const double* a, const double* b, double* r)
{
    for (int c = 0; c < 4; c++)
        for (int i = 0; i < 4; i++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += a[k * 4 + i] * b[c * 4 + k];
            r[c * 4 + i] = s;
        }
}

OEXPORT void o_camera_matrices(const vren_camera* cam, float* view, float* proj)
{
    double cy = std::cos((double)cam->yaw), sy = std::sin((double)cam->yaw);
    double cp = std::cos(-(double)cam->pitch), sp = std::sin(-(double)cam->pitch);
    double ry[16] = {cy, 0, -sy, 0, 0, 1, 0, 0, sy, 0, cy, 0, 0, 0, 0, 1};
    double rx[16] = {1, 0, 0, 0, 0, cp, sp, 0, 0, -sp, cp, 0, 0, 0, 0, 1};
    double orient[16];
    mat4_mul_d(ry, rx, orient);
    double inv_r[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) inv_r[c * 4 + r] = orient[r * 4 + c]; /* rotation inverse = transpose */
    double t[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -(double)cam->position[0], -(double)cam->position[1],
                    -(double)cam->position[2], 1};
    double v[16];
    mat4_mul_d(inv_r, t, v);
    for (int i = 0; i < 16; i++) view[i] = (float)v[i];
    float tan_half = (float)std::tan((double)(cam->fov_y / 2.0f));
    for (int i = 0; i < 16; i++) proj[i] = 0.0f;
    proj[0] = 1.0f / (tan_half * cam->aspect_ratio);
    proj[5] = 1.0f / tan_half;
    proj[10] = cam->far_plane / (cam->far_plane - cam->near_plane);
    proj[11] = 1.0f;
    proj[14] = -(cam->far_plane * cam->near_plane) / (cam->far_plane - cam->near_plane);
}

/* general 4x4 inverse in double (cofactor expansion), column-major */
static void mat4_inverse_d(const double* m, double* inv)
{
    double a[16];
    a[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    a[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    a[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    a[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    a[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    a[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    a[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    a[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    a[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    a[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    a[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    a[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    a[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    a[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    a[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    a[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * a[0] + m[1] * a[4] + m[2] * a[8] + m[3] * a[12];
    for (int i = 0; i < 16; i++) inv[i] = a[i] / det;
}

namespace {

/* Per-frame constants every stage derives from the camera (clustered_shading.cpp push constants). */
struct frame_consts {
    uint32_t W, H, Sx, Sy;
    float view[16], proj[16], inv_proj[16];
    double view_d[16], proj_d[16], inv_proj_d[16], inv_view_d[16];
    float near_plane, far_plane;
    float tan_half;   /* tan(m_fov_y / 2) */
    float slice_a;    /* 2 tan / Sy + 1 (clustered_shading.glsl:220) */
    float log_a;      /* log(1 + 2 tan / Sy) (find_unique_clusters.comp:65) */
    float log_fn;     /* log(far / near) (sliced keys) */
    float bg_threshold; /* far - EPSILON (shade.comp:86) */
    double eye[3];
};

frame_consts make_consts(uint32_t W, uint32_t H, const vren_camera* cam)
{
    frame_consts c;
    c.W = W;
    c.H = H;
    c.Sx = (W + 31) / 32;
    c.Sy = (H + 31) / 32;
    o_camera_matrices(cam, c.view, c.proj);
    for (int i = 0; i < 16; i++) {
        c.view_d[i] = c.view[i];
        c.proj_d[i] = c.proj[i];
    }
    mat4_inverse_d(c.proj_d, c.inv_proj_d);
    mat4_inverse_d(c.view_d, c.inv_view_d);
    for (int i = 0; i < 16; i++) c.inv_proj[i] = (float)c.inv_proj_d[i];
    c.near_plane = cam->near_plane;
    c.far_plane = cam->far_plane;
    c.tan_half = (float)std::tan((double)(cam->fov_y / 2.0f));
    c.slice_a = (2.0f * c.tan_half) / (float)c.Sy + 1.0f;
    c.log_a = vren_logf(1.0f + (2.0f * c.tan_half) / (float)c.Sy);
    c.log_fn = vren_logf(cam->far_plane / cam->near_plane);
    c.bg_threshold = cam->far_plane - k_epsilon_ref;
    for (int i = 0; i < 3; i++) c.eye[i] = (double)cam->position[i];
    return c;
}

} // namespace

/* the derived constants, exported so tests can compare them with the library's */
OEXPORT void o_frame_consts(uint32_t W, uint32_t H, const vren_camera* cam, float* inv_proj16, float* scalars4)
{
    frame_consts c = make_consts(W, H, cam);
    std::memcpy(inv_proj16, c.inv_proj, sizeof(c.inv_proj));
    scalars4[0] = c.tan_half;
    scalars4[1] = c.slice_a;
    scalars4[2] = c.log_a;
    scalars4[3] = c.log_fn;
}

/* ------------------------------------------------------------------------------------------------------------
 * Synthetic inputs (the same definitions as the GPU generator: include/vren_hip/synth.h)
 * ---------------------------------------------------------------------------------------------------------- */

OEXPORT void o_synth_lights(uint64_t initstate, uint32_t L, float* pos4, float* light4)
{
    for (uint32_t i = 0; i < L; i++) vren_synth_light(initstate, i, pos4 + 4 * i, light4 + 4 * i);
}

OEXPORT void o_synth_frame(uint64_t initstate, uint32_t W, uint32_t H, const vren_camera* cam, uint32_t n_boxes,
                           float* depth, uint16_t* normals)
{
    frame_consts c = make_consts(W, H, cam);
    std::vector<double> centers(3 * n_boxes), halves(3 * n_boxes);
    double tan_half = (double)c.tan_half;
    for (uint32_t b = 0; b < n_boxes; b++)
        vren_synth_box(initstate, b, c.inv_view_d, tan_half, (double)cam->aspect_ratio, &centers[3 * b],
                       &halves[3 * b]);
    for (uint32_t y = 0; y < H; y++)
        for (uint32_t x = 0; x < W; x++) {
            uint64_t p = (uint64_t)y * W + x;
            vren_synth_pixel(x, y, W, H, c.inv_view_d, tan_half, (double)cam->aspect_ratio, (double)cam->near_plane,
                             (double)cam->far_plane, c.proj_d[10], c.proj_d[14], centers.data(), halves.data(),
                             n_boxes, depth + p, normals + 4 * p);
        }
}

OEXPORT void o_synth_materials(uint64_t initstate, uint64_t n_pixels, float* albedo_roughness, float* metallic)
{
    for (uint64_t p = 0; p < n_pixels; p++) vren_synth_material(initstate, p, albedo_roughness + 4 * p, metallic + p);
}

/* ------------------------------------------------------------------------------------------------------------
 * construct_point_light_bvh (clustered_shading.cpp:60-346)
 * ---------------------------------------------------------------------------------------------------------- */

static uint32_t morton_interleave(uint32_t qx, uint32_t qy, uint32_t qz, uint32_t bits)
{
    /* discretize_point_light_positions.comp:33-39 (5 bits in the reference) */
    uint32_t code = 0;
    for (uint32_t i = 0; i < bits; i++) {
        code |= ((qx >> i) & 1u) << (i * 3);
        code |= ((qy >> i) & 1u) << (i * 3 + 1);
        code |= ((qz >> i) & 1u) << (i * 3 + 2);
    }
    return code;
}

OEXPORT void o_construct_point_light_bvh(const float* pos4, const float* light4, uint32_t L, const vren_camera* cam,
                                         const vren_cluster_cfg* cfg, float* view_pos4, uint32_t* codes,
                                         vren_bvh_node* bvh)
{
    frame_consts c = make_consts(32, 32, cam);
    /* 1. point_light_position_to_view_space.comp:26-32 */
    for (uint32_t i = 0; i < L; i++) {
        float p[4] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], 1.0f};
        vren_mat4_mul_vec4(c.view, p, view_pos4 + 4 * i);
    }
    /* 2-4. AABB of the view positions (reduce_vec4 min/max, clustered_shading.cpp:137-224) */
    float mn[3], mx[3];
    if (cfg->morton_aabb_mode == VREN_MORTON_GIVEN_AABB) {
        for (int a = 0; a < 3; a++) {
            mn[a] = cfg->morton_aabb_min[a];
            mx[a] = cfg->morton_aabb_max[a];
        }
    } else {
        for (int a = 0; a < 3; a++) {
            mn[a] = 1e35f;
            mx[a] = -1e35f;
        }
        for (uint32_t i = 0; i < L; i++)
            for (int a = 0; a < 3; a++) {
                mn[a] = tmin(mn[a], view_pos4[4 * i + a]);
                mx[a] = tmax(mx[a], view_pos4[4 * i + a]);
            }
    }
    /* 5. Morton codes (discretize_point_light_positions.comp:26-43) */
    uint32_t bits = cfg->morton_bits;
    float scale = (float)(1u << bits);
    std::vector<uint32_t> keys(L), vals(L);
    for (uint32_t i = 0; i < L; i++) {
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            float t = (view_pos4[4 * i + a] - mn[a]) / (mx[a] - mn[a]) * scale;
            q[a] = vren_f2u_sat(vren_floorf(t));
        }
        keys[i] = morton_interleave(q[0], q[1], q[2], bits);
        vals[i] = i;
    }
    /* 6. sort by code, ties by light index */
    o_stable_sort_kv(keys.data(), vals.data(), L, 0, 32);
    for (uint32_t i = 0; i < L; i++) {
        codes[2 * i] = keys[i];
        codes[2 * i + 1] = vals[i];
    }
    /* 7. leaves (init_light_array_bvh.comp:42-65) */
    uint32_t padded = o_calc_bvh_padded_leaf_count(L);
    for (uint32_t i = 0; i < padded; i++) {
        vren_bvh_node& n = bvh[i];
        n.pad = 0;
        if (i < L) {
            uint32_t li = vals[i];
            float r = light4[4 * li + 3];
            for (int a = 0; a < 3; a++) {
                n.min[a] = view_pos4[4 * li + a] - r;
                n.max[a] = view_pos4[4 * li + a] + r;
            }
            n.next = VREN_BVH_LEAF_NODE;
        } else {
            set_empty(n);
            n.next = VREN_BVH_INVALID_NODE;
        }
    }
    /* 8. build_bvh */
    o_build_bvh(bvh, padded);
}

/* ------------------------------------------------------------------------------------------------------------
 * Karras LBVH over the sorted (code, index) pairs (Karras, HPG 2012, "Maximizing Parallelism in the Construction
 * of BVHs, Octrees, and k-d Trees", section 4 and figure 4), with a sequential bottom-up refit.
 * ---------------------------------------------------------------------------------------------------------- */

static int lbvh_delta(const uint64_t* k, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n) return -1;
    uint64_t x = k[i] ^ k[j];
    return x ? __builtin_clzll(x) : 64;
}

static void lbvh_refit(vren_lbvh_node* nodes, uint32_t node, uint32_t n)
{
    if (node >= n - 1) return; /* leaf */
    vren_lbvh_node& p = nodes[node];
    lbvh_refit(nodes, p.left, n);
    lbvh_refit(nodes, p.right, n);
    const vren_lbvh_node& a = nodes[p.left];
    const vren_lbvh_node& b = nodes[p.right];
    for (int ax = 0; ax < 3; ax++) {
        p.min[ax] = tmin(a.min[ax], b.min[ax]);
        p.max[ax] = tmax(a.max[ax], b.max[ax]);
    }
}

OEXPORT void o_build_lbvh(const uint32_t* codes, uint32_t n, const float* view_pos4, const float* light4,
                          vren_lbvh_node* nodes, uint32_t* parents)
{
    std::vector<uint64_t> k(n);
    for (uint32_t i = 0; i < n; i++) k[i] = ((uint64_t)codes[2 * i] << 32) | codes[2 * i + 1];
    for (uint32_t i = 0; i < 2 * n - 1; i++) parents[i] = 0xFFFFFFFFu;
    /* leaves */
    for (uint32_t i = 0; i < n; i++) {
        vren_lbvh_node& l = nodes[n - 1 + i];
        uint32_t li = codes[2 * i + 1];
        float r = light4[4 * li + 3];
        for (int ax = 0; ax < 3; ax++) {
            l.min[ax] = view_pos4[4 * li + ax] - r;
            l.max[ax] = view_pos4[4 * li + ax] + r;
        }
        l.left = VREN_BVH_LEAF_NODE;
        l.right = li;
    }
    /* internal nodes: direction, range end, split (figure 4) */
    const int64_t N = n;
    for (int64_t i = 0; i + 1 < N; i++) {
        int d = lbvh_delta(k.data(), N, i, i + 1) - lbvh_delta(k.data(), N, i, i - 1) > 0 ? 1 : -1;
        int dmin = lbvh_delta(k.data(), N, i, i - d);
        int64_t lmax = 2;
        while (lbvh_delta(k.data(), N, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (lbvh_delta(k.data(), N, i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = lbvh_delta(k.data(), N, i, j);
        int64_t s = 0;
        int64_t t = l;
        do {
            t = (t + 1) / 2;
            if (lbvh_delta(k.data(), N, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
        uint32_t left = (uint32_t)(std::min(i, j) == gamma ? (N - 1 + gamma) : gamma);
        uint32_t right = (uint32_t)(std::max(i, j) == gamma + 1 ? (N - 1 + gamma + 1) : gamma + 1);
        nodes[i].left = left;
        nodes[i].right = right;
        parents[left] = (uint32_t)i;
        parents[right] = (uint32_t)i;
    }
    if (n > 1) lbvh_refit(nodes, 0, n);
}

/* ------------------------------------------------------------------------------------------------------------
 * find_unique_cluster_list (find_unique_clusters.comp:46-121, clustered_shading.glsl:7-43)
 * ---------------------------------------------------------------------------------------------------------- */

static uint32_t discretize_normal(const float n[3])
{
    if (n[0] == 0.0f && n[1] == 0.0f && n[2] == 0.0f) return 0xFFFFFFFFu;
    float min_t = 1e35f;
    uint32_t axis = 0, face_idx = 0;
    for (uint32_t i = 0; i < 3; i++) {
        float sgn = n[i] > 0.0f ? 1.0f : (n[i] < 0.0f ? -1.0f : 0.0f);
        float t = sgn / n[i];
        if (t < min_t) {
            min_t = t;
            axis = i;
            face_idx = (n[i] > 0.0f ? 1u : 0u) * 3u + i;
        }
    }
    float p[3] = {n[0] * min_t, n[1] * min_t, n[2] * min_t};
    float u = p[(axis + 1) % 3];
    float v = p[(axis + 2) % 3];
    uint32_t du = vren_f2u_sat(vren_floorf((u + 1.0f) / 2.0f * 3.0f));
    uint32_t dv = vren_f2u_sat(vren_floorf((v + 1.0f) / 2.0f * 3.0f));
    return (face_idx * 9u + du * 3u + dv) & 0x3Fu;
}

static uint32_t face_of_normal(const float n[3])
{
    if (n[0] == 0.0f && n[1] == 0.0f && n[2] == 0.0f) return 7u;
    float min_t = 1e35f;
    uint32_t face_idx = 0;
    for (uint32_t i = 0; i < 3; i++) {
        float sgn = n[i] > 0.0f ? 1.0f : (n[i] < 0.0f ? -1.0f : 0.0f);
        float t = sgn / n[i];
        if (t < min_t) {
            min_t = t;
            face_idx = (n[i] > 0.0f ? 1u : 0u) * 3u + i;
        }
    }
    return face_idx;
}

static uint32_t cluster_key(const frame_consts& c, const vren_cluster_cfg* cfg, uint32_t i, uint32_t j, uint32_t gx,
                            uint32_t gy, float depth, const uint16_t* nh)
{
    float fx = (float)gx / (float)c.W;
    float fy = (float)gy / (float)c.H;
    float v[4] = {fx, fy, depth, 1.0f};
    v[1] = 1.0f - v[1];
    v[0] = v[0] * 2.0f - 1.0f;
    v[1] = v[1] * 2.0f - 1.0f;
    float r[4];
    vren_mat4_mul_vec4(c.inv_proj, v, r);
    float z = r[2] / r[3];
    float n[3] = {vren_half_to_float(nh[0]), vren_half_to_float(nh[1]), vren_half_to_float(nh[2])};
    if (cfg->key_mode == VREN_KEYS_SLICED) {
        float s = vren_floorf(vren_logf(z / c.near_plane) / c.log_fn * (float)cfg->depth_slices);
        uint32_t k = vren_f2u_sat(s);
        if (k >= cfg->depth_slices) k = cfg->depth_slices - 1;
        return (i & 0xFFu) | ((j & 0xFFu) << 8) | ((k & 0xFFu) << 16) | (face_of_normal(n) << 24);
    }
    uint32_t k = vren_f2u_sat(vren_floorf(vren_logf(z / c.near_plane) / c.log_a));
    return (i & 0xFFu) | ((j & 0xFFu) << 8) | ((k & 0x3FFu) << 16) | (discretize_normal(n) << 26);
}

/* per-pixel keys of every tile pixel (including the wrapped rows/cols of partial tiles), tile-major */
OEXPORT void o_cluster_keys_raw(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                const float* depth, const uint16_t* normals, uint32_t* tile_keys)
{
    frame_consts c = make_consts(W, H, cam);
    for (uint32_t j = 0; j < c.Sy; j++)
        for (uint32_t i = 0; i < c.Sx; i++)
            for (uint32_t l = 0; l < 1024; l++) {
                uint32_t gx = i * 32 + (l & 31), gy = j * 32 + (l >> 5);
                uint64_t p = (uint64_t)(gy % H) * W + (gx % W);
                tile_keys[((uint64_t)j * c.Sx + i) * 1024 + l] =
                    cluster_key(c, cfg, i, j, gx, gy, depth[p], normals + 4 * p);
            }
}

/* returns the unique count U; keys beyond key_capacity are not written */
OEXPORT uint32_t o_find_unique_clusters(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                        const float* depth, const uint16_t* normals, uint32_t* keys,
                                        uint32_t key_capacity, uint32_t* refs)
{
    frame_consts c = make_consts(W, H, cam);
    uint32_t total = 0;
    uint32_t tk[1024];
    for (uint32_t j = 0; j < c.Sy; j++)
        for (uint32_t i = 0; i < c.Sx; i++) {
            for (uint32_t l = 0; l < 1024; l++) {
                uint32_t gx = i * 32 + (l & 31), gy = j * 32 + (l >> 5);
                uint64_t p = (uint64_t)(gy % H) * W + (gx % W);
                tk[l] = cluster_key(c, cfg, i, j, gx, gy, depth[p], normals + 4 * p);
            }
            std::vector<uint32_t> sorted(tk, tk + 1024);
            std::sort(sorted.begin(), sorted.end());
            sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
            for (size_t u = 0; u < sorted.size(); u++)
                if (total + u < key_capacity) keys[total + u] = sorted[u];
            for (uint32_t l = 0; l < 1024; l++) {
                uint32_t gx = i * 32 + (l & 31), gy = j * 32 + (l >> 5);
                if (gx >= W || gy >= H) continue; /* out-of-bounds imageStore is dropped */
                uint32_t rank = (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), tk[l]) - sorted.begin());
                refs[(uint64_t)gy * W + gx] = total + rank;
            }
            total += (uint32_t)sorted.size();
        }
    return total;
}

/* ------------------------------------------------------------------------------------------------------------
 * assign_lights (assign_lights.comp:77-241, clustered_shading.glsl:58-110)
 * ---------------------------------------------------------------------------------------------------------- */

namespace {

struct aabb {
    float mn[3], mx[3];
};

aabb cluster_aabb(const frame_consts& c, const vren_cluster_cfg* cfg, uint32_t key)
{
    uint32_t ci = key & 0xFFu, cj = (key >> 8) & 0xFFu, ck;
    float near_k, far_k;
    if (cfg->key_mode == VREN_KEYS_SLICED) {
        ck = (key >> 16) & 0xFFu;
        float S = (float)cfg->depth_slices;
        near_k = c.near_plane * vren_expf((float)ck / S * c.log_fn);
        far_k = c.near_plane * vren_expf((float)(ck + 1) / S * c.log_fn);
    } else {
        ck = (key >> 16) & 0x3FFu;
        near_k = c.near_plane * vren_powi(c.slice_a, ck);
        far_k = near_k * c.slice_a;
    }
    float fSx = (float)c.Sx, fSy = (float)c.Sy;
    /* tile corners in clip space */
    float x0 = (float)ci / fSx, y0 = (float)cj / fSy;
    float x1 = (float)(ci + 1) / fSx, y1 = (float)(cj + 1) / fSy;
    aabb box;
    if (cfg->cluster_aabb_mode == VREN_CLUSTER_AABB_TIGHT) {
        for (int a = 0; a < 3; a++) {
            box.mn[a] = INFINITY;
            box.mx[a] = -INFINITY;
        }
        for (int corner = 0; corner < 4; corner++) {
            float cx = (corner & 1) ? x1 : x0, cy = (corner & 2) ? y1 : y0;
            float v[4] = {cx * 2.0f - 1.0f, (1.0f - cy) * 2.0f - 1.0f, 0.0f, 1.0f};
            float r[4];
            vren_mat4_mul_vec4(c.inv_proj, v, r);
            float p[3] = {r[0] / r[3], r[1] / r[3], r[2] / r[3]};
            float d[3];
            vren_normalize3(p, d);
            for (int plane = 0; plane < 2; plane++) {
                float zp = plane ? far_k : near_k;
                float s = zp / d[2];
                for (int a = 0; a < 3; a++) {
                    float q = s * d[a];
                    box.mn[a] = tmin(box.mn[a], q);
                    box.mx[a] = tmax(box.mx[a], q);
                }
            }
        }
        return box;
    }
    /* reference: two corners, as written (clustered_shading.glsl:203-228) */
    float vmin[4] = {x0, 1.0f - y0, 0.0f, 1.0f};
    vmin[0] = vmin[0] * 2.0f - 1.0f;
    vmin[1] = vmin[1] * 2.0f - 1.0f;
    float vmax[4] = {x1, 1.0f - y1, 0.0f, 1.0f};
    vmax[0] = vmax[0] * 2.0f - 1.0f;
    vmax[1] = vmax[1] * 2.0f - 1.0f;
    float rmin[4], rmax[4];
    vren_mat4_mul_vec4(c.inv_proj, vmin, rmin);
    vren_mat4_mul_vec4(c.inv_proj, vmax, rmax);
    float pmin[3] = {rmin[0] / rmin[3], rmin[1] / rmin[3], rmin[2] / rmin[3]};
    float pmax[3] = {rmax[0] / rmax[3], rmax[1] / rmax[3], rmax[2] / rmax[3]};
    float d1[3], d2[3];
    vren_normalize3(pmin, d1);
    vren_normalize3(pmax, d2);
    float s1 = near_k / d1[2], s2 = far_k / d2[2];
    for (int a = 0; a < 3; a++) {
        box.mn[a] = s1 * d1[a];
        box.mx[a] = s2 * d2[a];
    }
    return box;
}

/* test_aabb_aabb (assign_lights.comp:84-94), cluster first */
inline bool test_aabb_aabb(const aabb& cl, const vren_bvh_node& n)
{
    return cl.mx[0] >= n.min[0] && cl.mn[0] <= n.max[0] && cl.mx[1] >= n.min[1] && cl.mn[1] <= n.max[1] &&
           cl.mx[2] >= n.min[2] && cl.mn[2] <= n.max[2];
}

/* leaf test (assign_lights.comp:97-102, 214): sphere(view pos, (max.x - min.x) / 2) vs the cluster AABB */
inline bool test_leaf(const aabb& cl, const float* o, float leaf_min_x, float leaf_max_x)
{
    float r = (leaf_max_x - leaf_min_x) / 2.0f;
    float d[3];
    for (int a = 0; a < 3; a++) {
        float p = glsl_max(cl.mn[a], glsl_min(o[a], cl.mx[a]));
        d[a] = p - o[a];
    }
    float dist = vren_sqrtf(vren_dot3(d, d));
    return dist < r;
}

/* DFS of assign_lights.comp:142-240 with 32-lane ballots, restated; returns hits in traversal order */
void traverse(const aabb& cl, const vren_bvh_node* bvh, uint32_t root, uint32_t level_count, uint32_t L,
              const uint32_t* codes, const float* view_pos4, std::vector<uint32_t>& hits)
{
    uint32_t overlaps[8] = {0};
    uint32_t level = 0;
    int state = 2; /* 0 advance, 1 pop, 2 entering */
    for (;;) {
        if (state == 0) {
            overlaps[level] &= ~(1u << __builtin_ctz(overlaps[level]));
            if (overlaps[level] == 0) {
                state = 1;
            } else {
                level++;
                state = 2;
            }
        } else if (state == 1) {
            if (level > 0) {
                level--;
                state = 0;
            } else {
                break;
            }
        } else {
            /* get_node_address (assign_lights.comp:108-119), integer form */
            uint64_t below = 0, pw = 32;
            for (uint32_t l = 0; l <= level; l++) {
                below += pw;
                pw *= 32;
            }
            uint64_t addr = (uint64_t)root - below;
            for (uint32_t l = 0; l < level; l++) {
                uint64_t w = 1;
                for (uint32_t k = 0; k < level - l; k++) w *= 32;
                addr += w * (uint64_t)__builtin_ctz(overlaps[l]);
            }
            if (level < level_count - 1) {
                uint32_t mask = 0;
                for (uint32_t lane = 0; lane < 32; lane++)
                    if (test_aabb_aabb(cl, bvh[addr + lane])) mask |= 1u << lane;
                overlaps[level] = mask;
                if (mask == 0) {
                    state = 1;
                } else {
                    level++;
                }
            } else {
                for (uint32_t lane = 0; lane < 32; lane++) {
                    const vren_bvh_node& n = bvh[addr + lane];
                    if (n.next != VREN_BVH_LEAF_NODE) continue; /* padding leaves never match */
                    uint32_t li = codes[2 * (addr + lane) + 1];
                    if (li >= L) continue;
                    if (test_leaf(cl, view_pos4 + 4 * li, n.min[0], n.max[0])) hits.push_back(li);
                }
                state = 1;
            }
        }
    }
}

} // namespace

/* counts[min(U, key_capacity)], offsets (exclusive scan of counts), indices (each list ascending);
 * returns the total number of assignments (indices beyond index_capacity are not written) */
OEXPORT uint64_t o_assign_lights(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                 const uint32_t* keys, uint32_t U, uint32_t key_capacity, const vren_bvh_node* bvh,
                                 uint32_t L, const uint32_t* codes, const float* view_pos4, uint32_t* counts,
                                 uint32_t* offsets, uint32_t* indices, uint32_t index_capacity)
{
    frame_consts c = make_consts(W, H, cam);
    uint32_t root = o_calc_bvh_root_index(L);
    uint32_t levels = o_calc_bvh_level_count(L);
    uint32_t n = std::min(U, key_capacity);
    uint64_t total = 0;
    std::vector<uint32_t> hits;
    for (uint32_t k = 0; k < n; k++) {
        aabb cl = cluster_aabb(c, cfg, keys[k]);
        hits.clear();
        if (L > 0) traverse(cl, bvh, root, levels, L, codes, view_pos4, hits);
        std::sort(hits.begin(), hits.end());
        counts[k] = (uint32_t)hits.size();
        offsets[k] = (uint32_t)total;
        for (size_t h = 0; h < hits.size(); h++)
            if (total + h < index_capacity) indices[total + h] = hits[h];
        total += hits.size();
    }
    return total;
}

/* Independent pin of the traversal: every valid leaf whose ancestors all pass test_aabb_aabb (the root is
 * never tested) and which passes the leaf test. Writes, per cluster, the ascending list; same layout as
 * o_assign_lights. mode 0 = ancestor predicate, 1 = plain brute force over all lights (no BVH). */
OEXPORT uint64_t o_assign_lights_brute(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                       const uint32_t* keys, uint32_t n_clusters, const vren_bvh_node* bvh,
                                       uint32_t L, const uint32_t* codes, const float* view_pos4,
                                       const float* light4, int mode, uint32_t* counts, uint32_t* indices,
                                       uint64_t index_capacity)
{
    frame_consts c = make_consts(W, H, cam);
    uint32_t padded = o_calc_bvh_padded_leaf_count(L);
    uint32_t levels = o_calc_bvh_level_count(L);
    /* level start offsets: leaves at 0, then each parent level */
    std::vector<uint64_t> level_start, level_size;
    {
        uint64_t s = 0, sz = padded;
        while (sz >= 1) {
            level_start.push_back(s);
            level_size.push_back(sz);
            s += sz;
            if (sz == 1) break;
            sz >>= 5;
        }
    }
    uint64_t total = 0;
    std::vector<uint32_t> hits;
    for (uint32_t k = 0; k < n_clusters; k++) {
        aabb cl = cluster_aabb(c, cfg, keys[k]);
        hits.clear();
        if (mode == 1) {
            for (uint32_t li = 0; li < L; li++) {
                float r = light4[4 * li + 3];
                float mnx = view_pos4[4 * li] - r, mxx = view_pos4[4 * li] + r;
                if (test_leaf(cl, view_pos4 + 4 * li, mnx, mxx)) hits.push_back(li);
            }
        } else {
            for (uint32_t leaf = 0; leaf < L; leaf++) {
                const vren_bvh_node& n = bvh[leaf];
                uint32_t li = codes[2 * leaf + 1];
                bool ok = test_leaf(cl, view_pos4 + 4 * li, n.min[0], n.max[0]);
                /* ancestors: levels 1 .. levels-1 above the leaves (level `levels` is the root, untested) */
                uint64_t pos = leaf;
                for (uint32_t l = 1; ok && l < levels; l++) {
                    pos >>= 5;
                    ok = test_aabb_aabb(cl, bvh[level_start[l] + pos]);
                }
                if (ok) hits.push_back(li);
            }
        }
        std::sort(hits.begin(), hits.end());
        counts[k] = (uint32_t)hits.size();
        for (size_t h = 0; h < hits.size(); h++)
            if (total + h < index_capacity) indices[total + h] = hits[h];
        total += hits.size();
    }
    return total;
}

OEXPORT void o_cluster_aabb(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                            uint32_t key, float* mn3, float* mx3)
{
    frame_consts c = make_consts(W, H, cam);
    aabb b = cluster_aabb(c, cfg, key);
    for (int a = 0; a < 3; a++) {
        mn3[a] = b.mn[a];
        mx3[a] = b.mx[a];
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * shade (shade.comp:69-141, pbr.glsl:6-112) restated over a scalar type T. T = double is the accuracy oracle;
 * T = float is the reference's own single-precision formula, kept to show how far it sits from FP64.
 * Directional lights: {dir[3], pad, color[3], pad}.
 * ---------------------------------------------------------------------------------------------------------- */

namespace {

template <typename T>
struct v3 {
    T x, y, z;
};
template <typename T>
inline v3<T> mk(T a, T b, T c) { return v3<T>{a, b, c}; }
template <typename T>
inline v3<T> operator+(v3<T> a, v3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename T>
inline v3<T> operator-(v3<T> a, v3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename T>
inline v3<T> operator*(v3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename T>
inline v3<T> operator*(v3<T> a, v3<T> b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
template <typename T>
inline T dot(v3<T> a, v3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <typename T>
inline v3<T> normalize(v3<T> a)
{
    T l = std::sqrt(dot(a, a));
    return mk(a.x / l, a.y / l, a.z / l);
}
template <typename T>
inline T mx0(T a) { return a < T(0) ? T(0) : a; }

template <typename T>
v3<T> pbr_apply_light(v3<T> V, v3<T> N, v3<T> L, v3<T> radiance, v3<T> albedo, T metallic, T roughness)
{
    const T PI = T(k_pi_ref);
    v3<T> H = normalize(V - L);
    v3<T> F0 = mk(T(0.04f), T(0.04f), T(0.04f));
    F0 = F0 * (T(1) - metallic) + albedo * metallic; /* mix */
    T a2 = roughness * roughness;
    T NdotH = mx0(dot(N, H));
    T NdotH2 = NdotH * NdotH;
    T dn = NdotH2 * (a2 - T(1)) + T(1);
    dn = PI * dn * dn;
    T NDF = a2 / dn;
    T NdotV = mx0(dot(N, V));
    v3<T> mL = mk(-L.x, -L.y, -L.z);
    T NdotL = mx0(dot(N, mL));
    T k = roughness;
    T G = (NdotV / (NdotV * (T(1) - k) + k)) * (NdotL / (NdotL * (T(1) - k) + k));
    T HdotV = mx0(dot(H, V));
    T f5 = std::pow(T(1) - HdotV, T(5));
    v3<T> F = F0 + (mk(T(1), T(1), T(1)) - F0) * f5;
    v3<T> kD = (mk(T(1), T(1), T(1)) - F) * (T(1) - metallic);
    T num = NDF * G;
    T den = T(4) * NdotV * NdotL + T(0.0001f);
    T spec = num / den;
    v3<T> r;
    r.x = mx0((kD.x * albedo.x / PI + F.x * spec) * radiance.x * NdotL);
    r.y = mx0((kD.y * albedo.y / PI + F.y * spec) * radiance.y * NdotL);
    r.z = mx0((kD.z * albedo.z / PI + F.z * spec) * radiance.z * NdotL);
    return r;
}

template <typename T>
T attenuation(T dist, T radius, uint32_t mode)
{
    if (mode == VREN_ATTEN_SMOOTH) {
        T q = dist / radius;
        T q2 = q * q;
        T w = mx0(T(1) - q2 * q2);
        return w * w;
    }
    T x = dist / radius;
    if (x > T(1)) x = T(1);
    return T(1) - std::pow(x, T(1.5));
}

struct pixel_in {
    bool shaded;
    double pos[3]; /* world position (double reconstruction) */
};

template <typename T>
bool reconstruct(const frame_consts& c, uint32_t gx, uint32_t gy, float depth, v3<T>& P)
{
    T fx = T(gx) / T(c.W), fy = T(gy) / T(c.H);
    T v[4] = {fx, fy, T(depth), T(1)};
    v[1] = T(1) - v[1];
    v[0] = v[0] * T(2) - T(1);
    v[1] = v[1] * T(2) - T(1);
    T r[4];
    for (int i = 0; i < 4; i++) {
        T m0 = sizeof(T) == 8 ? (T)c.inv_proj_d[0 * 4 + i] : (T)c.inv_proj[0 * 4 + i];
        T m1 = sizeof(T) == 8 ? (T)c.inv_proj_d[1 * 4 + i] : (T)c.inv_proj[1 * 4 + i];
        T m2 = sizeof(T) == 8 ? (T)c.inv_proj_d[2 * 4 + i] : (T)c.inv_proj[2 * 4 + i];
        T m3 = sizeof(T) == 8 ? (T)c.inv_proj_d[3 * 4 + i] : (T)c.inv_proj[3 * 4 + i];
        r[i] = ((m0 * v[0] + m1 * v[1]) + m2 * v[2]) + m3 * v[3];
    }
    T w = r[3];
    for (int i = 0; i < 4; i++) r[i] = r[i] / w;
    if (!((double)r[2] < (double)c.bg_threshold)) return false;
    T q[3];
    for (int i = 0; i < 3; i++) {
        T m0 = (T)c.inv_view_d[0 * 4 + i], m1 = (T)c.inv_view_d[1 * 4 + i];
        T m2 = (T)c.inv_view_d[2 * 4 + i], m3 = (T)c.inv_view_d[3 * 4 + i];
        q[i] = ((m0 * r[0] + m1 * r[1]) + m2 * r[2]) + m3 * r[3];
    }
    P = mk(q[0], q[1], q[2]);
    return true;
}

template <typename T>
void shade_frame(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg, const float* depth,
                 const uint16_t* normals, const float* albedo_rough, const float* metallic, const uint32_t* refs,
                 uint32_t key_capacity, const uint32_t* counts, const uint32_t* offsets, const uint32_t* indices,
                 uint64_t index_capacity, const float* pos4, const float* light4, const float* dir_lights,
                 uint32_t n_dir, T* out, uint8_t* shaded_mask)
{
    frame_consts c = make_consts(W, H, cam);
    v3<T> eye = mk((T)cam->position[0], (T)cam->position[1], (T)cam->position[2]);
    const T PI = T(k_pi_ref);
    for (uint32_t gy = 0; gy < H; gy++)
        for (uint32_t gx = 0; gx < W; gx++) {
            uint64_t p = (uint64_t)gy * W + gx;
            v3<T> P;
            bool sh = reconstruct<T>(c, gx, gy, depth[p], P);
            if (shaded_mask) shaded_mask[p] = sh ? 1 : 0;
            if (!sh) continue;
            v3<T> N = mk((T)vren_half_to_float(normals[4 * p]), (T)vren_half_to_float(normals[4 * p + 1]),
                         (T)vren_half_to_float(normals[4 * p + 2]));
            v3<T> albedo = mk((T)albedo_rough[4 * p], (T)albedo_rough[4 * p + 1], (T)albedo_rough[4 * p + 2]);
            T rough = (T)albedo_rough[4 * p + 3];
            T metal = (T)metallic[p];
            v3<T> V = normalize(eye - P);
            v3<T> Lo = mk(T(0), T(0), T(0));
            uint32_t cl = refs[p];
            if (cl < key_capacity) {
                uint32_t cnt = counts[cl], off = offsets[cl];
                for (uint32_t e = 0; e < cnt; e++) {
                    if ((uint64_t)off + e >= index_capacity) break;
                    uint32_t li = indices[off + e];
                    v3<T> lp = mk((T)pos4[4 * li], (T)pos4[4 * li + 1], (T)pos4[4 * li + 2]);
                    v3<T> d = P - lp;
                    T dist = std::sqrt(dot(d, d));
                    v3<T> L = mk(d.x / dist, d.y / dist, d.z / dist);
                    T att = attenuation<T>(dist, (T)light4[4 * li + 3], cfg->attenuation);
                    v3<T> rad = mk((T)light4[4 * li] * att, (T)light4[4 * li + 1] * att, (T)light4[4 * li + 2] * att);
                    Lo = Lo + pbr_apply_light<T>(V, N, L, rad, albedo, metal, rough);
                }
            }
            for (uint32_t dl = 0; dl < n_dir; dl++) {
                const float* D = dir_lights + 8 * dl;
                v3<T> L = mk((T)D[0], (T)D[1], (T)D[2]);
                v3<T> rad = mk((T)D[4], (T)D[5], (T)D[6]);
                Lo = Lo + pbr_apply_light<T>(V, N, L, rad, albedo, metal, rough);
            }
            v3<T> col = albedo * T(0.03f) + Lo;
            T g = T(1) / T(2.2f);
            out[4 * p] = std::pow(col.x, g);
            out[4 * p + 1] = std::pow(col.y, g);
            out[4 * p + 2] = std::pow(col.z, g);
            out[4 * p + 3] = T(1);
            (void)PI;
        }
}


/* one pixel of shade_frame<double>: returns false for background, writes rgb */
bool shade_pixel_f64(const frame_consts& c, const vren_camera* cam, const vren_cluster_cfg* cfg, uint32_t gx,
                     uint32_t gy, const float* depth, const uint16_t* normals, const float* albedo_rough,
                     const float* metallic, const uint32_t* refs, uint32_t key_capacity, const uint32_t* counts,
                     const uint32_t* offsets, const uint32_t* indices, uint64_t index_capacity, const double* pos4,
                     const double* light4, double out[3], double* albedo_override, double* rough_override,
                     const double* metal_override = nullptr)
{
    typedef double T;
    uint64_t p = (uint64_t)gy * c.W + gx;
    v3<T> P;
    if (!reconstruct<T>(c, gx, gy, depth[p], P)) return false;
    v3<T> N = mk((T)vren_half_to_float(normals[4 * p]), (T)vren_half_to_float(normals[4 * p + 1]),
                 (T)vren_half_to_float(normals[4 * p + 2]));
    v3<T> albedo = albedo_override ? mk(albedo_override[0], albedo_override[1], albedo_override[2])
                                   : mk((T)albedo_rough[4 * p], (T)albedo_rough[4 * p + 1], (T)albedo_rough[4 * p + 2]);
    T rough = rough_override ? *rough_override : (T)albedo_rough[4 * p + 3];
    T metal = metal_override ? *metal_override : (T)metallic[p];
    v3<T> eye = mk((T)cam->position[0], (T)cam->position[1], (T)cam->position[2]);
    v3<T> V = normalize(eye - P);
    v3<T> Lo = mk(T(0), T(0), T(0));
    uint32_t cl = refs[p];
    if (cl < key_capacity) {
        uint32_t cnt = counts[cl], off = offsets[cl];
        for (uint32_t e = 0; e < cnt; e++) {
            if ((uint64_t)off + e >= index_capacity) break;
            uint32_t li = indices[off + e];
            v3<T> lp = mk(pos4[4 * li], pos4[4 * li + 1], pos4[4 * li + 2]);
            v3<T> d = P - lp;
            T dist = std::sqrt(dot(d, d));
            v3<T> L = mk(d.x / dist, d.y / dist, d.z / dist);
            T att = attenuation<T>(dist, light4[4 * li + 3], cfg->attenuation);
            v3<T> rad = mk(light4[4 * li] * att, light4[4 * li + 1] * att, light4[4 * li + 2] * att);
            Lo = Lo + pbr_apply_light<T>(V, N, L, rad, albedo, metal, rough);
        }
    }
    v3<T> col = albedo * T(0.03f) + Lo;
    T g = T(1) / T(2.2f);
    out[0] = std::pow(col.x, g);
    out[1] = std::pow(col.y, g);
    out[2] = std::pow(col.z, g);
    return true;
}

} // namespace

OEXPORT void o_shade_f64(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                         const float* depth, const uint16_t* normals, const float* albedo_rough, const float* metallic,
                         const uint32_t* refs, uint32_t key_capacity, const uint32_t* counts, const uint32_t* offsets,
                         const uint32_t* indices, uint64_t index_capacity, const float* pos4, const float* light4,
                         const float* dir_lights, uint32_t n_dir, double* out, uint8_t* shaded_mask)
{
    shade_frame<double>(W, H, cam, cfg, depth, normals, albedo_rough, metallic, refs, key_capacity, counts, offsets,
                        indices, index_capacity, pos4, light4, dir_lights, n_dir, out, shaded_mask);
}

/* FP64 shade of selected pixels only (a sample of a large frame): out3[3 i ..] for pixels[i], shaded[i] = 0 for
 * background pixels (left unwritten) */
OEXPORT void o_shade_f64_pixels(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                const float* depth, const uint16_t* normals, const float* albedo_rough,
                                const float* metallic, const uint32_t* refs, uint32_t key_capacity,
                                const uint32_t* counts, const uint32_t* offsets, const uint32_t* indices,
                                uint64_t index_capacity, const float* pos4, const float* light4, uint32_t L,
                                const uint64_t* pixels, uint64_t n_pixels, double* out3, uint8_t* shaded)
{
    frame_consts c = make_consts(W, H, cam);
    std::vector<double> p4(pos4, pos4 + 4 * (size_t)L), l4(light4, light4 + 4 * (size_t)L);
    for (uint64_t i = 0; i < n_pixels; i++) {
        uint64_t p = pixels[i];
        shaded[i] = shade_pixel_f64(c, cam, cfg, (uint32_t)(p % W), (uint32_t)(p / W), depth, normals, albedo_rough,
                                    metallic, refs, key_capacity, counts, offsets, indices, index_capacity,
                                    p4.data(), l4.data(), out3 + 3 * i, nullptr, nullptr)
                        ? 1
                        : 0;
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * Depth-buffer pyramid (depth_buffer_pyramid.cpp:14-23, 177-299; depth_buffer_copy.comp;
 * depth_buffer_reduce.comp), level by level as the reference dispatches it: level 0 = the depth buffer; level l + 1
 * (max(W >> (l+1), 1) x max(H >> (l+1), 1)) texel (x, y) = GLSL max over the in-bounds level-l texels
 * (2x + i, 2y + j), i outer, j inner, starting from 0.0. Levels are packed one after another per image.
 * ---------------------------------------------------------------------------------------------------------- */

OEXPORT uint32_t o_depth_pyramid_level_count(uint32_t W, uint32_t H)
{
    return ilog2(std::max(W, H)) + 1;
}

OEXPORT void o_depth_pyramid(const float* depth, uint32_t W, uint32_t H, float* out)
{
    const uint32_t levels = o_depth_pyramid_level_count(W, H);
    std::memcpy(out, depth, sizeof(float) * (size_t)W * H);
    float* prev = out;
    uint32_t pw = W, ph = H;
    for (uint32_t l = 1; l < levels; l++) {
        const uint32_t w = std::max(W >> l, 1u), h = std::max(H >> l, 1u);
        float* cur = prev + (size_t)pw * ph;
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                float m = 0.0f;
                for (uint32_t i = 0; i < 2; i++)
                    for (uint32_t j = 0; j < 2; j++) {
                        uint32_t fx = 2 * x + i, fy = 2 * y + j;
                        if (fx < pw && fy < ph) m = glsl_max(m, prev[(size_t)fy * pw + fx]);
                    }
                cur[(size_t)y * w + x] = m;
            }
        prev = cur;
        pw = w;
        ph = h;
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * Debug consumers: show_clusters.comp:27-126 (per-pixel cluster views) and show_bvh.comp (node boxes as lines)
 * ---------------------------------------------------------------------------------------------------------- */

static uint32_t triple32(uint32_t x)
{
    x ^= x >> 17;
    x *= 0xed5ad4bbu;
    x ^= x >> 11;
    x *= 0xac4c1b51u;
    x ^= x >> 15;
    x *= 0x31848babu;
    x ^= x >> 14;
    return x;
}

static float hash01(uint32_t x) { return (float)triple32(x) / (float)0xffffffffu; }

OEXPORT uint32_t o_triple32(uint32_t x) { return triple32(x); }

OEXPORT void o_show_clusters(uint32_t W, uint32_t H, int32_t mode, const uint32_t* refs, const uint32_t* keys,
                             uint32_t key_cap, const uint32_t* counts, const uint32_t* offsets,
                             const uint32_t* indices, uint64_t index_cap, float* out)
{
    const uint32_t tx = (W + 31) / 32, ty = (H + 31) / 32;
    for (uint64_t p = 0; p < (uint64_t)W * H; p++) {
        float c[3] = {1.0f, 0.0f, 0.0f};
        const uint32_t ci = refs[p];
        if (ci < key_cap) {
            const uint32_t key = keys[ci];
            const uint32_t i = key & 0xFFu, j = (key >> 8) & 0xFFu, k = (key >> 16) & 0x3FFu, n = key >> 26;
            if (mode == 0) {
                c[0] = c[1] = c[2] = hash01(key);
            } else if (mode == 1) {
                c[0] = (float)i / (float)tx;
                c[1] = (float)j / (float)ty;
                c[2] = (float)k / 1024.0f;
            } else if (mode == 2) {
                c[0] = (float)(n / 9) / 6.0f;
                c[1] = (float)((n / 3) % 3) / 3.0f;
                c[2] = (float)(n % 3) / 3.0f;
            } else if (mode == 3) {
                c[0] = c[1] = c[2] = glsl_min((float)counts[ci] / 64.0f, 1.0f);
            } else if (mode == 4) {
                c[0] = c[1] = c[2] = hash01(offsets[ci]);
            } else if (mode == 5) {
                uint32_t h = 0;
                for (uint32_t q = 0; q < counts[ci] && (uint64_t)offsets[ci] + q < index_cap; q++)
                    h ^= indices[offsets[ci] + q];
                c[0] = c[1] = c[2] = hash01(h);
            }
        }
        out[4 * p] = c[0];
        out[4 * p + 1] = c[1];
        out[4 * p + 2] = c[2];
        out[4 * p + 3] = 1.0f;
    }
}

/* visualize_bvh.cpp:54-86 + show_bvh.comp: levels level_count (leaves) .. 0, 24 vertices {x, y, z, colour} per node */
OEXPORT void o_visualize_bvh(const vren_bvh_node* bvh, uint32_t level_count, uint32_t* vertices)
{
    static const uint32_t colors[7] = {0xff0000u, 0xffff00u, 0x00ff00u, 0x0000ffu, 0x00ffffu, 0xff00ffu, 0xffffffu};
    uint64_t node = 0;
    for (int32_t level = (int32_t)level_count; level >= 0; level--) {
        const uint64_t count = 1ull << (5 * level);
        const uint32_t ci = level_count - (uint32_t)level + 2;
        const uint32_t color = ci < 7 ? colors[ci] : 0xffffffu;
        for (uint64_t q = 0; q < count; q++, node++) {
            const vren_bvh_node& n = bvh[node];
            float m[3] = {n.min[0], n.min[1], n.min[2]}, M[3] = {n.max[0], n.max[1], n.max[2]};
            uint32_t col = color;
            if (n.next == VREN_BVH_INVALID_NODE) {
                m[0] = m[1] = m[2] = M[0] = M[1] = M[2] = 0.0f;
                col = 0u;
            }
            /* VREN_WRITE_DEBUG_DRAW_BUFFER_AABB: bottom face loop, top face loop, then the four verticals */
            const float v[24][3] = {
                {m[0], m[1], m[2]}, {M[0], m[1], m[2]}, {M[0], m[1], m[2]}, {M[0], m[1], M[2]},
                {M[0], m[1], M[2]}, {m[0], m[1], M[2]}, {m[0], m[1], M[2]}, {m[0], m[1], m[2]},
                {m[0], M[1], m[2]}, {M[0], M[1], m[2]}, {M[0], M[1], m[2]}, {M[0], M[1], M[2]},
                {M[0], M[1], M[2]}, {m[0], M[1], M[2]}, {m[0], M[1], M[2]}, {m[0], M[1], m[2]},
                {m[0], m[1], m[2]}, {m[0], M[1], m[2]}, {M[0], m[1], m[2]}, {M[0], M[1], m[2]},
                {M[0], m[1], M[2]}, {M[0], M[1], M[2]}, {m[0], m[1], M[2]}, {m[0], M[1], M[2]}};
            for (int k = 0; k < 24; k++) {
                uint32_t* o = vertices + 4 * (24 * node + k);
                std::memcpy(o, v[k], 12);
                o[3] = col;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * Material resolve (shade.comp:89-96; material.hpp:19-26; tinygltf_parser.cpp:157 RGBA8 UNORM textures) in FP64:
 * the texels are chosen by include/vren_hip/sampler.h (the float coordinate transformation, as the kernels do),
 * their values c / 255 and the bilinear weights (1 - a)(1 - b), a (1 - b), (1 - a) b, a b are evaluated in double.
 * Textures here hold host texel pointers. A material index past the table, a texture index past the table and an
 * empty texture read zeros.
 * ---------------------------------------------------------------------------------------------------------- */

static void texel_f64(const vren_texture& t, uint32_t i, uint32_t j, double o[4])
{
    const uint8_t* c = t.texels + 4 * ((uint64_t)j * t.width + i);
    for (int k = 0; k < 4; k++) o[k] = (double)c[k] / 255.0;
}

static void sample_texture_f64(const vren_texture* texs, uint32_t n_tex, uint32_t ti, float s, float t, double o[4])
{
    for (int k = 0; k < 4; k++) o[k] = 0.0;
    if (ti >= n_tex) return;
    const vren_texture& tx = texs[ti];
    if (!tx.texels || tx.width == 0 || tx.height == 0) return;
    uint32_t i0, i1, j0, j1;
    float a, b;
    vren_texel_axis(s, tx.width, tx.filter, tx.address_u, &i0, &i1, &a);
    vren_texel_axis(t, tx.height, tx.filter, tx.address_v, &j0, &j1, &b);
    double c00[4];
    texel_f64(tx, i0, j0, c00);
    if (tx.filter != VREN_FILTER_LINEAR) {
        for (int k = 0; k < 4; k++) o[k] = c00[k];
        return;
    }
    double c10[4], c01[4], c11[4];
    texel_f64(tx, i1, j0, c10);
    texel_f64(tx, i0, j1, c01);
    texel_f64(tx, i1, j1, c11);
    const double A = a, B = b;
    for (int k = 0; k < 4; k++)
        o[k] = (1.0 - A) * (1.0 - B) * c00[k] + A * (1.0 - B) * c10[k] + (1.0 - A) * B * c01[k] + A * B * c11[k];
}

/* albedo_rough4[4 n] (albedo rgb, roughness) and metal[n] of n pixels */
OEXPORT void o_resolve_materials_f64(uint64_t n, const float* texcoords, const uint16_t* mat_idx,
                                     const vren_material* mats, uint32_t n_mat, const vren_texture* texs,
                                     uint32_t n_tex, double* albedo_rough4, double* metal)
{
    for (uint64_t p = 0; p < n; p++) {
        double* ar = albedo_rough4 + 4 * p;
        ar[0] = ar[1] = ar[2] = ar[3] = 0.0;
        metal[p] = 0.0;
        const uint32_t mi = mat_idx[p];
        if (mi >= n_mat) continue;
        const vren_material& m = mats[mi];
        const float s = texcoords[2 * p], t = texcoords[2 * p + 1];
        double base[4], mr[4];
        sample_texture_f64(texs, n_tex, m.base_color_texture_idx, s, t, base);
        sample_texture_f64(texs, n_tex, m.metallic_roughness_texture_idx, s, t, mr);
        for (int k = 0; k < 3; k++) ar[k] = base[k] * (double)m.base_color_factor[k];
        ar[3] = mr[1] * (double)m.roughness_factor;
        metal[p] = mr[2] * (double)m.metallic_factor;
    }
}

/* shade (FP64) with the material path: every pixel's albedo, roughness and metallic resolved in FP64 first */
OEXPORT void o_shade_f64_materials(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                   const float* depth, const uint16_t* normals, const float* texcoords,
                                   const uint16_t* mat_idx, const vren_material* mats, uint32_t n_mat,
                                   const vren_texture* texs, uint32_t n_tex, const uint32_t* refs,
                                   uint32_t key_capacity, const uint32_t* counts, const uint32_t* offsets,
                                   const uint32_t* indices, uint64_t index_capacity, const float* pos4,
                                   const float* light4, uint32_t L, double* out, uint8_t* shaded_mask)
{
    frame_consts c = make_consts(W, H, cam);
    const uint64_t n = (uint64_t)W * H;
    std::vector<double> ar(4 * n), me(n);
    o_resolve_materials_f64(n, texcoords, mat_idx, mats, n_mat, texs, n_tex, ar.data(), me.data());
    std::vector<double> p4(pos4, pos4 + 4 * (size_t)L), l4(light4, light4 + 4 * (size_t)L);
    for (uint64_t p = 0; p < n; p++) {
        double rgb[3];
        double rough = ar[4 * p + 3];
        /* no per-pixel material arrays: every value comes from the overrides */
        bool sh = shade_pixel_f64(c, cam, cfg, (uint32_t)(p % W), (uint32_t)(p / W), depth, normals, nullptr, nullptr,
                                  refs, key_capacity, counts, offsets, indices, index_capacity, p4.data(), l4.data(),
                                  rgb, &ar[4 * p], &rough, &me[p]);
        shaded_mask[p] = sh ? 1 : 0;
        for (int k = 0; k < 3; k++) out[4 * p + k] = sh ? rgb[k] : NAN;
        out[4 * p + 3] = sh ? 1.0 : NAN;
    }
}

OEXPORT void o_shade_f32(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                         const float* depth, const uint16_t* normals, const float* albedo_rough, const float* metallic,
                         const uint32_t* refs, uint32_t key_capacity, const uint32_t* counts, const uint32_t* offsets,
                         const uint32_t* indices, uint64_t index_capacity, const float* pos4, const float* light4,
                         const float* dir_lights, uint32_t n_dir, float* out, uint8_t* shaded_mask)
{
    shade_frame<float>(W, H, cam, cfg, depth, normals, albedo_rough, metallic, refs, key_capacity, counts, offsets,
                       indices, index_capacity, pos4, light4, dir_lights, n_dir, out, shaded_mask);
}


/* ------------------------------------------------------------------------------------------------------------
 * Differentiable shade (no reference counterpart): L2 loss and its FP64 adjoint with the light lists held fixed.
 * Parameters are the light positions and light records (rgb = colour * intensity, radius) in double, and the
 * per-pixel albedo / roughness. Use the smooth attenuation (VREN_ATTEN_SMOOTH).
 * ---------------------------------------------------------------------------------------------------------- */

/* sum over the listed pixels (background pixels contribute nothing) of |out - target|^2 over rgb */
OEXPORT double o_shade_loss_f64(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                const float* depth, const uint16_t* normals, const float* albedo_rough,
                                const float* metallic, const uint32_t* refs, uint32_t key_capacity,
                                const uint32_t* counts, const uint32_t* offsets, const uint32_t* indices,
                                uint64_t index_capacity, const double* pos4, const double* light4,
                                const float* target, const uint32_t* pixels, uint64_t n_pixels,
                                const double* pixel_albedo_rough)
{
    frame_consts c = make_consts(W, H, cam);
    double loss = 0.0;
    for (uint64_t i = 0; i < n_pixels; i++) {
        uint32_t p = pixels[i];
        double out[3];
        double alb[3], rgh;
        double* ao = nullptr;
        double* ro = nullptr;
        if (pixel_albedo_rough) {
            alb[0] = pixel_albedo_rough[4 * i];
            alb[1] = pixel_albedo_rough[4 * i + 1];
            alb[2] = pixel_albedo_rough[4 * i + 2];
            rgh = pixel_albedo_rough[4 * i + 3];
            ao = alb;
            ro = &rgh;
        }
        if (!shade_pixel_f64(c, cam, cfg, p % W, p / W, depth, normals, albedo_rough, metallic, refs, key_capacity,
                             counts, offsets, indices, index_capacity, pos4, light4, out, ao, ro))
            continue;
        for (int ch = 0; ch < 3; ch++) {
            double e = out[ch] - (double)target[4 * (uint64_t)p + ch];
            loss += e * e;
        }
    }
    return loss;
}

/* FP64 adjoint: given dLoss/d(output rgb) per pixel, ACCUMULATES into g_pos[L*4] (xyz), g_light[L*4]
 * (rgb, radius) and g_ar[W*H*4] (albedo rgb, roughness). Derived by hand from pbr.glsl:42-93 with the smooth window
 * att = (1 - (d/r)^4)^2; clamps max(., 0) contribute no gradient where they are active. */
OEXPORT void o_shade_backward_f64(uint32_t W, uint32_t H, const vren_camera* cam, const vren_cluster_cfg* cfg,
                                  const float* depth, const uint16_t* normals, const float* albedo_rough,
                                  const float* metallic, const uint32_t* refs, uint32_t key_capacity,
                                  const uint32_t* counts, const uint32_t* offsets, const uint32_t* indices,
                                  uint64_t index_capacity, const double* pos4, const double* light4,
                                  const float* grad_out, double* g_pos, double* g_light, double* g_ar)
{
    typedef double T;
    frame_consts c = make_consts(W, H, cam);
    const T PI = T(k_pi_ref);
    v3<T> eye = mk((T)cam->position[0], (T)cam->position[1], (T)cam->position[2]);
    for (uint32_t gy = 0; gy < H; gy++)
        for (uint32_t gx = 0; gx < W; gx++) {
            uint64_t p = (uint64_t)gy * W + gx;
            v3<T> P;
            if (!reconstruct<T>(c, gx, gy, depth[p], P)) continue;
            uint32_t cl = refs[p];
            v3<T> N = mk((T)vren_half_to_float(normals[4 * p]), (T)vren_half_to_float(normals[4 * p + 1]),
                         (T)vren_half_to_float(normals[4 * p + 2]));
            T alb[3] = {(T)albedo_rough[4 * p], (T)albedo_rough[4 * p + 1], (T)albedo_rough[4 * p + 2]};
            T rough = (T)albedo_rough[4 * p + 3], metal = (T)metallic[p];
            v3<T> V = normalize(eye - P);
            T NdotV = mx0(dot(N, V));
            T k = rough, a2 = rough * rough;
            T gvden = NdotV * (T(1) - k) + k;
            T ggxV = NdotV / gvden;
            T F0[3];
            for (int ch = 0; ch < 3; ch++) F0[ch] = T(0.04f) * (T(1) - metal) + alb[ch] * metal;
            uint32_t cnt = 0, off = 0;
            if (cl < key_capacity) {
                cnt = counts[cl];
                off = offsets[cl];
            }
            struct Pair {
                uint32_t li;
                v3<T> L;
                T dist, r, x, w, att;
            };
            std::vector<Pair> pairs;
            T col[3] = {T(0.03f) * alb[0], T(0.03f) * alb[1], T(0.03f) * alb[2]};
            for (uint32_t e = 0; e < cnt; e++) {
                if ((uint64_t)off + e >= index_capacity) break;
                uint32_t li = indices[off + e];
                v3<T> lp = mk(pos4[4 * li], pos4[4 * li + 1], pos4[4 * li + 2]);
                v3<T> d = P - lp;
                T dist = std::sqrt(dot(d, d));
                Pair pr;
                pr.li = li;
                pr.dist = dist;
                pr.r = light4[4 * li + 3];
                pr.x = dist / pr.r;
                pr.w = T(1) - pr.x * pr.x * pr.x * pr.x;
                pr.att = attenuation<T>(dist, pr.r, cfg->attenuation);
                pr.L = mk(d.x / dist, d.y / dist, d.z / dist);
                v3<T> rad = mk(light4[4 * li] * pr.att, light4[4 * li + 1] * pr.att, light4[4 * li + 2] * pr.att);
                v3<T> ctb = pbr_apply_light<T>(V, N, pr.L, rad, mk(alb[0], alb[1], alb[2]), metal, rough);
                col[0] += ctb.x;
                col[1] += ctb.y;
                col[2] += ctb.z;
                if (pr.att > T(0) && pr.w > T(0)) pairs.push_back(pr);
            }
            T g = T(1) / T(2.2f);
            T gT[3];
            for (int ch = 0; ch < 3; ch++) {
                gT[ch] = (T)grad_out[4 * p + ch] * g * std::pow(col[ch], g - T(1));
                g_ar[4 * p + ch] += T(0.03f) * gT[ch];
            }
            for (const Pair& pr : pairs) {
                const uint32_t li = pr.li;
                v3<T> mL = mk(-pr.L.x, -pr.L.y, -pr.L.z);
                T NdotL = dot(N, mL);
                if (!(NdotL > T(0))) continue;
                v3<T> h = V + mL;
                T hl = std::sqrt(dot(h, h));
                v3<T> Hn = mk(h.x / hl, h.y / hl, h.z / hl);
                T NHraw = dot(N, Hn);
                T NdotH = mx0(NHraw);
                T inner = NdotH * NdotH * (a2 - T(1)) + T(1);
                T NDF = a2 / (PI * inner * inner);
                T glden = NdotL * (T(1) - k) + k;
                T glv = NdotL / glden;
                T G = ggxV * glv;
                T HVraw = dot(Hn, V);
                T HdotV = mx0(HVraw);
                T om = T(1) - HdotV;
                T f5 = std::pow(om, T(5));
                T D = T(4) * NdotV * NdotL + T(0.0001f);
                T spec = NDF * G / D;
                T km = T(1) - metal;
                T lc[3] = {light4[4 * li], light4[4 * li + 1], light4[4 * li + 2]};
                T g_att = 0, g_NdotL = 0, g_spec = 0, g_f5 = 0;
                for (int ch = 0; ch < 3; ch++) {
                    T F = F0[ch] + (T(1) - F0[ch]) * f5;
                    T kD = (T(1) - F) * km;
                    T B = kD * alb[ch] / PI + F * spec;
                    g_light[4 * li + ch] += gT[ch] * B * pr.att * NdotL;
                    T gB = gT[ch] * lc[ch] * pr.att * NdotL;
                    g_att += gT[ch] * B * lc[ch] * NdotL;
                    g_NdotL += gT[ch] * B * lc[ch] * pr.att;
                    g_spec += gB * F;
                    T gF = gB * (spec - alb[ch] * km / PI);
                    g_ar[4 * p + ch] += gB * kD / PI + gF * (T(1) - f5) * metal;
                    g_f5 += gF * (T(1) - F0[ch]);
                }
                T g_HdotV = HVraw > T(0) ? g_f5 * (T(-5) * std::pow(om, T(4))) : T(0);
                T g_NDF = g_spec * G / D, g_G = g_spec * NDF / D;
                g_NdotL += -g_spec * spec / D * T(4) * NdotV;
                T g_gl = g_G * ggxV, g_ggxV = g_G * glv;
                g_NdotL += g_gl * k / (glden * glden);
                T g_k = g_gl * (-NdotL * (T(1) - NdotL) / (glden * glden)) +
                        g_ggxV * (-NdotV * (T(1) - NdotV) / (gvden * gvden));
                T g_inner = g_NDF * (T(-2) * NDF / inner);
                T g_a2 = g_NDF / (PI * inner * inner) + g_inner * NdotH * NdotH;
                T g_NdotH = NHraw > T(0) ? g_inner * T(2) * NdotH * (a2 - T(1)) : T(0);
                g_ar[4 * p + 3] += g_a2 * T(2) * rough + g_k;
                T gh[3], gL[3];
                T Ncomp[3] = {N.x, N.y, N.z}, Vcomp[3] = {V.x, V.y, V.z}, Hcomp[3] = {Hn.x, Hn.y, Hn.z};
                T Lcomp[3] = {pr.L.x, pr.L.y, pr.L.z};
                for (int i = 0; i < 3; i++) {
                    gh[i] = (g_NdotH * (Ncomp[i] - NdotH * Hcomp[i]) + g_HdotV * (Vcomp[i] - HdotV * Hcomp[i])) / hl;
                    gL[i] = -(gh[i] + g_NdotL * Ncomp[i]);
                }
                T datt_dx = T(-8) * pr.x * pr.x * pr.x * pr.w;
                T g_dist = g_att * datt_dx / pr.r;
                T gLL = gL[0] * Lcomp[0] + gL[1] * Lcomp[1] + gL[2] * Lcomp[2];
                for (int i = 0; i < 3; i++) g_pos[4 * li + i] -= (gL[i] - gLL * Lcomp[i]) / pr.dist + g_dist * Lcomp[i];
                g_light[4 * li + 3] += g_att * datt_dx * (-pr.x / pr.r);
            }
        }
}

/* bounce_point_lights.comp:33-73 (dispatched by point_light_bouncer.cpp:17-57), one light at a time: clamp the
 * position into [lo, hi] with GLSL min(max(p, lo), hi), then up to 32 segments: the six wall times
 * (wall - p) / d with non-positive ones replaced by 1e35, min_t = their GLSL min in the shader's nesting,
 * step = min(min_t - 1e-5, remaining), p += step * d, and when min_t < remaining the components whose wall time
 * equals min_t flip sign. Writes w = 1 (positions) and w = 0 (directions). */
static inline float o_glsl_min(float x, float y) { return y < x ? y : x; }
static inline float o_glsl_max(float x, float y) { return x < y ? y : x; }

OEXPORT void o_bounce_point_lights(float* pos, float* dir, uint32_t n, const float* lo, const float* hi, float speed,
                                   float dt)
{
    const float kInf = 1e35f, kEps = 1e-5f;
    for (uint32_t i = 0; i < n; i++) {
        float p[3], d[3];
        for (int c = 0; c < 3; c++) {
            p[c] = o_glsl_min(o_glsl_max(pos[4 * i + c], lo[c]), hi[c]);
            d[c] = dir[4 * i + c];
        }
        float rem = speed * dt;
        for (int j = 0; j < 32 && rem > 0.0f; j++) {
            float t1[3], t2[3];
            for (int c = 0; c < 3; c++) {
                t1[c] = (lo[c] - p[c]) / d[c];
                t1[c] = t1[c] <= 0.0f ? kInf : t1[c];
                t2[c] = (hi[c] - p[c]) / d[c];
                t2[c] = t2[c] <= 0.0f ? kInf : t2[c];
            }
            const float min_t =
                o_glsl_min(t1[0], o_glsl_min(t2[0], o_glsl_min(t1[1], o_glsl_min(t2[1], o_glsl_min(t1[2], t2[2])))));
            const float step = o_glsl_min(min_t - kEps, rem);
            for (int c = 0; c < 3; c++) p[c] = p[c] + step * d[c];
            if (min_t < rem)
                for (int c = 0; c < 3; c++) d[c] = d[c] * ((t1[c] == min_t || t2[c] == min_t) ? -1.0f : 1.0f);
            rem = rem - step;
        }
        for (int c = 0; c < 3; c++) {
            pos[4 * i + c] = p[c];
            dir[4 * i + c] = d[c];
        }
        pos[4 * i + 3] = 1.0f;
        dir[4 * i + 3] = 0.0f;
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * CPU baseline helpers (bench.py): the reference tests' CPU operations (radix_sort.cpp:88 std::sort,
 * blelloch_scan.cpp:100-116 std::exclusive_scan, reduce.cpp std::accumulate) on 1 or `threads` std::threads
 * (no TBB): chunks in parallel, then a combine (pairwise std::inplace_merge levels for the sort, a prefix over
 * chunk totals for the scan). Each returns its wall time in seconds (std::chrono::steady_clock), inputs in place.
 * ---------------------------------------------------------------------------------------------------------- */

namespace {

template <typename F>
void parallel_for(uint32_t threads, F f)
{
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; t++) pool.emplace_back(f, t);
    f(0u);
    for (auto& th : pool) th.join();
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

/* stable sort of (key, value) pairs by key: std::stable_sort per chunk, then log2(threads) levels of
 * std::inplace_merge (stable) */
OEXPORT double o_cpu_stable_sort_pairs(uint64_t* pairs, uint64_t n, uint32_t threads)
{
    auto by_key = [](uint64_t a, uint64_t b) { return (uint32_t)a < (uint32_t)b; };
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 1) {
        std::stable_sort(pairs, pairs + n, by_key);
        return seconds_since(t0);
    }
    std::vector<uint64_t> bounds(threads + 1);
    for (uint32_t t = 0; t <= threads; t++) bounds[t] = n * t / threads;
    parallel_for(threads, [&](uint32_t t) { std::stable_sort(pairs + bounds[t], pairs + bounds[t + 1], by_key); });
    for (uint32_t w = 1; w < threads; w *= 2) {
        uint32_t merges = (threads + 2 * w - 1) / (2 * w);
        parallel_for(merges, [&](uint32_t m) {
            uint32_t a = 2 * w * m, b = std::min(a + w, threads), c = std::min(a + 2 * w, threads);
            if (b < c) std::inplace_merge(pairs + bounds[a], pairs + bounds[b], pairs + bounds[c], by_key);
        });
    }
    return seconds_since(t0);
}

/* exclusive scan mod 2^32: chunk totals, their prefix, then std::exclusive_scan per chunk from its prefix */
OEXPORT double o_cpu_exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t threads)
{
    auto t0 = std::chrono::steady_clock::now();
    if (threads <= 1) {
        std::exclusive_scan(in, in + n, out, 0u);
        return seconds_since(t0);
    }
    std::vector<uint64_t> bounds(threads + 1);
    for (uint32_t t = 0; t <= threads; t++) bounds[t] = n * t / threads;
    std::vector<uint32_t> tot(threads);
    parallel_for(threads, [&](uint32_t t) { tot[t] = std::accumulate(in + bounds[t], in + bounds[t + 1], 0u); });
    std::vector<uint32_t> pre(threads);
    std::exclusive_scan(tot.begin(), tot.end(), pre.begin(), 0u);
    parallel_for(threads, [&](uint32_t t) {
        std::exclusive_scan(in + bounds[t], in + bounds[t + 1], out + bounds[t], pre[t]);
    });
    return seconds_since(t0);
}

/* sum mod 2^32 (std::accumulate per chunk, then over the chunk sums) */
OEXPORT double o_cpu_reduce_add(const uint32_t* in, uint64_t n, uint32_t threads, uint32_t* result)
{
    auto t0 = std::chrono::steady_clock::now();
    threads = std::max(threads, 1u);
    std::vector<uint64_t> bounds(threads + 1);
    for (uint32_t t = 0; t <= threads; t++) bounds[t] = n * t / threads;
    std::vector<uint32_t> tot(threads);
    parallel_for(threads, [&](uint32_t t) { tot[t] = std::accumulate(in + bounds[t], in + bounds[t + 1], 0u); });
    *result = std::accumulate(tot.begin(), tot.end(), 0u);
    return seconds_since(t0);
}
