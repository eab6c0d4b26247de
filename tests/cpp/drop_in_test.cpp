// drop_in_test.cpp -- exercises the C++ drop-in classes (include/vren) the way the reference's tests exercise
// its operators (vren_test/vren_test/primitives/*.cpp): seeded inputs, the GPU result against std:: on the host.
// The clustered stages run through vren::cluster_and_shade and must give the same bytes as the batch chain,
// whose parity with the CPU oracle is tested separately.
//
//   drop_in_test            all checks on HIP device 0
//   drop_in_test --host     host-only context: precondition errors map to the reference's exceptions
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "primitives/blelloch_scan.hpp"
#include "primitives/bucket_sort.hpp"
#include "primitives/build_bvh.hpp"
#include "primitives/radix_sort.hpp"
#include "primitives/reduce.hpp"
#include "point_light_bouncer.hpp"
#include "debug_consumers.hpp"
#include "pipeline/clustered_shading.hpp"
#include "pipeline/depth_buffer_pyramid.hpp"
#include "../../include/vren_hip/synth.h"

static int g_failures = 0;
#define EXPECT(cond, ...)                                                                                              \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                           \
            std::printf(__VA_ARGS__);                                                                                  \
            std::printf("\n");                                                                                         \
            g_failures++;                                                                                              \
        }                                                                                                              \
    } while (0)

#define HIPCHECK(x)                                                                                                    \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) {                                                                                        \
            std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__);                         \
            std::exit(2);                                                                                              \
        }                                                                                                              \
    } while (0)

template <typename F>
static bool throws_invalid_argument(F f)
{
    try {
        f();
    } catch (std::invalid_argument const&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static std::vector<uint32_t> pcg_keys(uint64_t initstate, uint64_t initseq, size_t n)
{
    vren_pcg32 r = vren_pcg32_seed(initstate, initseq);
    std::vector<uint32_t> k(n);
    for (auto& x : k) x = vren_pcg32_next(&r);
    return k;
}

template <typename T>
static vren::owned_device_buffer upload(std::vector<T> const& h)
{
    vren::owned_device_buffer b(h.size() * sizeof(T));
    HIPCHECK(hipMemcpy(b.m_ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return b;
}

template <typename T>
static std::vector<T> download(vren::device_buffer const& b, size_t n, size_t byte_offset = 0)
{
    std::vector<T> h(n);
    HIPCHECK(hipMemcpy(h.data(), b.at<char>(byte_offset), n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

static void host_only()
{
    vren::context ctx(-1);
    vren::resource_container rc;
    vren::device_buffer fake;
    fake.m_ptr = reinterpret_cast<void*>(4096);
    fake.m_size = 1 << 20;
    vren::radix_sort sort(ctx);
    // radix_sort.cpp:158-161: length below 1024 or not a power of 2
    EXPECT(throws_invalid_argument([&] { sort({}, rc, fake, 1000, fake, fake); }), "radix length 1000");
    EXPECT(throws_invalid_argument([&] { sort({}, rc, fake, 3000, fake, fake); }), "radix length 3000");
    // valid arguments, but the host-only context has no device: reported, not run
    EXPECT(throws_invalid_argument([&] { sort({}, rc, fake, 2048, fake, fake); }), "no device");
    vren::build_bvh bvh(ctx);
    EXPECT(throws_invalid_argument([&] { bvh({}, rc, fake, 33); }), "bvh leaf count 33");
    vren::reduce<uint32_t, vren::ReduceOperationAdd> red(ctx);
    EXPECT(throws_invalid_argument([&] { red({}, rc, fake, 0, 0, 1); }), "reduce n = 0");
    vren::blelloch_scan scan(ctx);
    // blelloch_scan.cpp:68: the down-sweep's length must be a power of 2
    EXPECT(throws_invalid_argument([&] { scan.downsweep({}, rc, fake, 1000, 0, 1, true); }), "downsweep length 1000");
    vren_demo::point_light_bouncer bouncer(ctx);
    vren::device_buffer small = fake;
    small.m_size = 16 * 10;
    EXPECT(throws_invalid_argument([&] {
               bouncer.bounce(0, {}, rc, small, fake, 11, vren::vec3{-1.0f, -1.0f, -1.0f}, vren::vec3{1.0f, 1.0f, 1.0f},
                              1.0f, 0.1f);
           }),
           "bouncer buffers smaller than 16 * count bytes");
    vren::clustered_shading::find_unique_cluster_list fuc(ctx);
    vren::camera cam;
    cam.m_near_plane = 0.0f;
    vren::gbuffer gb;
    vren::depth_buffer_t db;
    db.m_buffer = fake;
    gb.m_normal_buffer = fake;
    vren::device_buffer cnt = fake;
    EXPECT(throws_invalid_argument([&] { fuc(0, {}, rc, {64, 64}, cam, gb, db, fake, cnt, fake); }), "near = 0");
    EXPECT(vren::calc_bvh_padded_leaf_count(4096) == 32768 && vren::calc_bvh_root_index(4096) == 33824,
           "calc_bvh helpers");
    EXPECT(vren::calc_reduce_output_buffer_length(1000) == 1024, "calc_reduce_output_buffer_length");
    EXPECT(vren::bucket_sort::get_required_output_buffer_size(4096) == 4096 * 8 + 65536 * 4, "bucket output size");
}

static void primitives(vren::context& ctx)
{
    vren::resource_container rc;
    vren::stream s;
    // radix_sort, keys only (radix_sort.cpp:68-128 test shape, seeded)
    {
        const uint32_t n = 1u << 20;
        auto keys = pcg_keys(42, 54, n);
        auto buf = upload(keys);
        vren::radix_sort sort(ctx);
        auto s1 = sort.create_scratch_buffer_1(n);
        auto s2 = sort.create_scratch_buffer_2(n);
        sort(s, rc, buf, n, s1, s2);
        ctx.sync(s);
        auto got = download<uint32_t>(buf, n);
        std::sort(keys.begin(), keys.end());
        EXPECT(got == keys, "radix_sort keys-only 2^20");
    }
    // radix_sort, key-value, ragged length and a partial bit range: equals std::stable_sort by (key & 0xFFFFF)
    {
        const uint32_t n = 123457;
        auto keys = pcg_keys(43, 54, n);
        std::vector<uint32_t> vals(n);
        std::iota(vals.begin(), vals.end(), 0u);
        auto kb = upload(keys);
        auto vb = upload(vals);
        vren::owned_device_buffer scratch(vren::radix_sort::get_required_scratch_buffer_size(n, true));
        vren::radix_sort sort(ctx);
        sort(s, rc, kb, vb, n, scratch, 0, 20);
        ctx.sync(s);
        auto gk = download<uint32_t>(kb, n);
        auto gv = download<uint32_t>(vb, n);
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t b) { return (keys[a] & 0xFFFFF) < (keys[b] & 0xFFFFF); });
        bool ok = true;
        for (uint32_t i = 0; i < n && ok; i++) ok = gv[i] == idx[i] && gk[i] == keys[idx[i]];
        EXPECT(ok, "radix_sort key-value, bits [0, 20), n = %u", n);
    }
    // reduce<uint32, add> and <uint32, max>: tree totals vs std::accumulate / max_element
    {
        const uint32_t n = 1000003, p = vren::calc_reduce_output_buffer_length(n);
        auto keys = pcg_keys(44, 54, n);
        auto in = upload(keys);
        vren::owned_device_buffer out((size_t) p * 4);
        vren::reduce<uint32_t, vren::ReduceOperationAdd> add(ctx);
        add(s, rc, in, n, 0, out, 0, 1);
        ctx.sync(s);
        auto tree = download<uint32_t>(out, p);
        EXPECT(tree[p - 1] == std::accumulate(keys.begin(), keys.end(), 0u), "reduce add total");
        vren::reduce<uint32_t, vren::ReduceOperationMax> mx(ctx);
        mx(s, rc, in, n, 0, out, 0, 1);
        ctx.sync(s);
        tree = download<uint32_t>(out, p);
        EXPECT(tree[p - 1] == *std::max_element(keys.begin(), keys.end()), "reduce max total");
        // leaves of the tree at even indices are the inputs
        EXPECT(tree[0] == keys[0] && tree[2] == keys[2], "reduce tree leaves");
    }
    // reduce<vec4, min> in place over 4096 vec4
    {
        const uint32_t n = 4096;
        auto r = pcg_keys(45, 54, n * 4);
        std::vector<vren::vec4> v(n);
        for (uint32_t i = 0; i < n; i++)
            v[i] = {vren_unit_f32(r[4 * i]), vren_unit_f32(r[4 * i + 1]), vren_unit_f32(r[4 * i + 2]),
                    vren_unit_f32(r[4 * i + 3])};
        auto b = upload(v);
        vren::reduce<vren::vec4, vren::ReduceOperationMin> mn(ctx);
        mn(s, rc, b, n, 0, 1);
        ctx.sync(s);
        auto t = download<vren::vec4>(b, n);
        float m = 2.0f;
        for (auto& x : v) m = std::min(m, x.y);
        EXPECT(t[n - 1].y == m, "reduce vec4 min");
    }
    // blelloch_scan in place, 3 segments of 2^18
    {
        const uint32_t n = 1u << 18;
        auto keys = pcg_keys(46, 54, 3 * n);
        auto b = upload(keys);
        vren::blelloch_scan scan(ctx);
        scan(s, rc, b, n, 0, 3);
        ctx.sync(s);
        auto got = download<uint32_t>(b, 3 * n);
        std::vector<uint32_t> want(3 * n);
        for (int seg = 0; seg < 3; seg++) {
            uint32_t run = 0;
            for (uint32_t i = 0; i < n; i++) {
                want[seg * n + i] = run;
                run += keys[seg * n + i];
            }
        }
        EXPECT(got == want, "blelloch_scan 3 x 2^18");
    }
    // bucket_sort: uvec2 pairs by key & 0xFFFF, offsets appended after the 256-byte aligned pairs
    {
        const uint32_t n = 4097;
        auto r = pcg_keys(47, 54, n);
        std::vector<uint32_t> pairs(2 * n);
        for (uint32_t i = 0; i < n; i++) {
            pairs[2 * i] = r[i];
            pairs[2 * i + 1] = i;
        }
        auto in = upload(pairs);
        vren::owned_device_buffer out(vren::bucket_sort::get_required_output_buffer_size(n));
        vren::bucket_sort bs(ctx);
        bs(s, rc, in, n, 0, out, 0);
        ctx.sync(s);
        auto got = download<uint32_t>(out, 2 * n);
        auto off = download<uint32_t>(out, 65536, ((size_t) n * 8 + 255) / 256 * 256);
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return (r[a] & 0xFFFF) < (r[b] & 0xFFFF); });
        bool ok = true;
        for (uint32_t i = 0; i < n && ok; i++) ok = got[2 * i + 1] == idx[i] && got[2 * i] == r[idx[i]];
        EXPECT(ok, "bucket_sort order");
        uint32_t key0 = r[idx[0]] & 0xFFFF, key_last = r[idx[n - 1]] & 0xFFFF;
        EXPECT(off[key0] == 0 && off[key_last] <= n - 1 && off[65535] <= n, "bucket offsets");
    }
    // blelloch_scan::downsweep after reduce, the radix sort's use (radix_sort.cpp:230-289): 16 segments of 2^14
    {
        const uint32_t len = 1u << 14, blocks = 16;
        auto keys = pcg_keys(49, 54, len * blocks);
        auto buf = upload(keys);
        vren::reduce<uint32_t, vren::ReduceOperationAdd> red(ctx);
        vren::blelloch_scan scan(ctx);
        red(s, rc, buf, len, 0, blocks);
        scan.downsweep(s, rc, buf, len, 0, blocks, true);
        ctx.sync(s);
        auto got = download<uint32_t>(buf, len * blocks);
        std::vector<uint32_t> want(len * blocks);
        for (uint32_t b = 0; b < blocks; b++)
            std::exclusive_scan(keys.begin() + b * len, keys.begin() + (b + 1) * len, want.begin() + b * len, 0u);
        EXPECT(got == want, "reduce + downsweep == std::exclusive_scan");
    }
    // point_light_bouncer: a light with no wall within reach moves by exactly speed * dt * d (w = 1 and 0)
    {
        std::vector<float> pos = {1.0f, 2.0f, 3.0f, 7.0f}, dir = {0.5f, -0.25f, 0.125f, 9.0f};
        auto dp = upload(pos);
        auto dd = upload(dir);
        vren_demo::point_light_bouncer bouncer(ctx);
        bouncer.bounce(0, s, rc, dp, dd, 1, vren::vec3{-50.0f, -50.0f, -50.0f}, vren::vec3{50.0f, 50.0f, 50.0f}, 4.0f,
                       0.5f);
        ctx.sync(s);
        auto gp = download<float>(dp, 4);
        auto gd = download<float>(dd, 4);
        EXPECT(gp == std::vector<float>({2.0f, 1.5f, 3.25f, 1.0f}) &&
                   gd == std::vector<float>({0.5f, -0.25f, 0.125f, 0.0f}),
               "point_light_bouncer");
    }
    // build_bvh over 1024 leaves: every parent = min/max over its valid children (build_bvh.cpp:175-210)
    {
        const uint32_t leaves = 1024, len = vren::calc_bvh_buffer_length(leaves);
        auto r = pcg_keys(48, 54, leaves * 6);
        std::vector<vren::bvh_node> nodes(len);
        for (uint32_t i = 0; i < leaves; i++) {
            float a = vren_uniform_f32(r[6 * i], -10, 10), b = vren_uniform_f32(r[6 * i + 1], -10, 10);
            float c = vren_uniform_f32(r[6 * i + 2], -10, 10);
            nodes[i].m_min = {a, b, c};
            nodes[i].m_max = {a + vren_unit_f32(r[6 * i + 3]), b + vren_unit_f32(r[6 * i + 4]),
                              c + vren_unit_f32(r[6 * i + 5])};
            nodes[i].m_next = i < 1000 ? vren::bvh_node::k_leaf_node : vren::bvh_node::k_invalid_node;
            nodes[i]._pad = 0;
        }
        auto b = upload(nodes);
        vren::build_bvh build(ctx);
        build(s, rc, b, leaves);
        ctx.sync(s);
        auto got = download<vren::bvh_node>(b, len);
        bool ok = true;
        uint32_t src = 0, count = leaves;
        while (count > 1 && ok) {
            uint32_t dst = src + count;
            for (uint32_t p = 0; p < count / 32 && ok; p++) {
                float mn[3] = {1e30f, 1e30f, 1e30f}, mxv[3] = {-1e30f, -1e30f, -1e30f};
                bool any = false;
                for (uint32_t c = 0; c < 32; c++) {
                    auto const& ch = got[src + 32 * p + c];
                    if (ch.is_invalid()) continue;
                    any = true;
                    mn[0] = std::min(mn[0], ch.m_min.x);
                    mn[1] = std::min(mn[1], ch.m_min.y);
                    mn[2] = std::min(mn[2], ch.m_min.z);
                    mxv[0] = std::max(mxv[0], ch.m_max.x);
                    mxv[1] = std::max(mxv[1], ch.m_max.y);
                    mxv[2] = std::max(mxv[2], ch.m_max.z);
                }
                auto const& par = got[dst + p];
                if (any)
                    ok = par.m_next == src + 32 * p && par.m_min.x == mn[0] && par.m_min.y == mn[1] &&
                         par.m_min.z == mn[2] && par.m_max.x == mxv[0] && par.m_max.y == mxv[1] &&
                         par.m_max.z == mxv[2];
                else
                    ok = par.is_invalid();
            }
            src = dst;
            count /= 32;
        }
        EXPECT(ok, "build_bvh parents");
    }
}

static void clustered(vren::context& ctx)
{
    const uint32_t W = 640, H = 360, L = 512;
    vren::stream s;
    vren_camera cc;
    cc.position[0] = 0.0f;
    cc.position[1] = 0.0f;
    cc.position[2] = -50.0f;
    cc.fov_y = 1.5707964f;
    cc.aspect_ratio = (float) W / (float) H;
    cc.near_plane = 0.1f;
    cc.far_plane = 100.0f;
    cc.yaw = 0.0f;
    cc.pitch = 0.0f;
    vren::owned_device_buffer depth((size_t) W * H * 4), normals((size_t) W * H * 8), albedo((size_t) W * H * 16),
        metal((size_t) W * H * 4), pos((size_t) L * 16), lights((size_t) L * 16), out_a((size_t) W * H * 16),
        out_b((size_t) W * H * 16);
    vren_ctx* h = ctx.handle();
    vren::check_status(h, vren_synth_frame(h, nullptr, 4042, W, H, &cc, 64, depth.at<float>(), normals.at<uint16_t>()),
                       "synth_frame");
    vren::check_status(h, vren_synth_materials(h, nullptr, 4042, (uint64_t) W * H, albedo.at<float>(), metal.at<float>()),
                       "synth_materials");
    vren::check_status(h, vren_synth_lights(h, nullptr, 4049, L, pos.at<float>(), lights.at<float>()), "synth_lights");
    HIPCHECK(hipMemset(out_a.m_ptr, 0, out_a.m_size));
    HIPCHECK(hipMemset(out_b.m_ptr, 0, out_b.m_size));
    HIPCHECK(hipDeviceSynchronize());

    vren::camera cam;
    cam.m_position = {cc.position[0], cc.position[1], cc.position[2]};
    cam.m_fov_y = cc.fov_y;
    cam.m_aspect_ratio = cc.aspect_ratio;
    cam.m_near_plane = cc.near_plane;
    cam.m_far_plane = cc.far_plane;
    vren::gbuffer gb;
    gb.m_normal_buffer = normals;
    gb.m_albedo_roughness_buffer = albedo;
    gb.m_metallic_buffer = metal;
    vren::depth_buffer_t db;
    db.m_buffer = depth;
    vren::light_array la;
    la.m_point_light_position_buffer = pos;
    la.m_point_light_buffer = lights;
    la.m_point_light_count = L;

    vren::cluster_and_shade::limits lim;
    lim.m_max_point_light_count = 4096;
    vren::cluster_and_shade cas(ctx, lim);
    cas(s, {W, H}, cam, gb, db, la, out_a);
    ctx.sync(s);

    // the same frame through the batch chain (one frame, plain mode)
    vren_frame_desc fd;
    std::memset(&fd, 0, sizeof(fd));
    fd.width = W;
    fd.height = H;
    fd.camera = cc;
    fd.depth = depth.at<float>();
    fd.normals = normals.at<uint16_t>();
    fd.albedo_roughness = albedo.at<float>();
    fd.metallic = metal.at<float>();
    fd.light_positions = pos.at<float>();
    fd.point_lights = lights.at<float>();
    fd.light_count = L;
    fd.output = out_b.at<float>();
    vren_cluster_cfg cfg = vren::clustered_shading::reference_config();
    vren_batch* batch = nullptr;
    vren::check_status(h, vren_batch_create(h, &fd, 1, &cfg, &batch), "batch_create");
    vren::check_status(h, vren_batch_run(batch, nullptr, VREN_RUN_PER_FRAME), "batch_run");
    ctx.sync(s);
    vren_frame_outputs fo;
    vren::check_status(h, vren_batch_frame_outputs(batch, 0, &fo), "frame_outputs");

    auto a = download<uint32_t>(out_a, (size_t) W * H * 4);
    auto b = download<uint32_t>(out_b, (size_t) W * H * 4);
    EXPECT(a == b, "cluster_and_shade output == batch chain output");
    uint32_t ua = download<uint32_t>(cas.m_cluster_key_dispatch_params_buffer, 1)[0];
    uint32_t ub = 0;
    HIPCHECK(hipMemcpy(&ub, fo.cluster_count, 4, hipMemcpyDeviceToHost));
    EXPECT(ua == ub && ua > 0, "unique cluster count %u vs %u", ua, ub);
    auto ka = download<uint32_t>(cas.m_cluster_key_buffer, ua);
    std::vector<uint32_t> kb(ua);
    HIPCHECK(hipMemcpy(kb.data(), fo.cluster_keys, (size_t) ua * 4, hipMemcpyDeviceToHost));
    EXPECT(ka == kb, "unique cluster keys");
    size_t lit = 0;
    for (size_t i = 0; i < a.size(); i += 4) lit += a[i] != 0;
    EXPECT(lit > (size_t) W * H / 4, "shaded pixels %zu", lit);
    vren_batch_destroy(batch);

    // the reference's material path (cluster_and_shade with a material_buffer, shade.comp:89-96): one material over
    // 1x1 textures of 255s gives albedo, roughness and metallic equal to its factors exactly (255 / 255 = 1), so
    // the output must equal the resolved-material path's over G-buffer planes filled with those factors
    {
        const float fac[4] = {0.25f, 0.5f, 0.75f, 0.6f}, metal_f = 0.3f;
        std::vector<float> ar((size_t) W * H * 4), me((size_t) W * H, metal_f);
        for (size_t i = 0; i < (size_t) W * H; i++)
            for (int c = 0; c < 4; c++) ar[4 * i + c] = fac[c];
        auto d_ar = upload(ar);
        auto d_me = upload(me);
        std::vector<uint8_t> white(4, 255);
        vren::texture t;
        auto d_white = upload(white);
        t.m_texels = d_white;
        t.m_width = t.m_height = 1;
        ctx.m_texture_manager.m_textures = {t, t};
        ctx.m_texture_manager.rewrite_descriptor_set();
        vren::material m{0u, 1u, metal_f, fac[3], vren::vec4{fac[0], fac[1], fac[2], 1.0f}};
        vren::material_buffer mb(ctx, 4);
        mb.write(&m, 1);
        std::vector<float> tc((size_t) W * H * 2, 0.5f);
        std::vector<uint16_t> mi((size_t) W * H, 0);
        auto d_tc = upload(tc);
        auto d_mi = upload(mi);
        vren::gbuffer gm;
        gm.m_normal_buffer = normals;
        gm.m_texcoord_buffer = d_tc;
        gm.m_material_index_buffer = d_mi;
        vren::gbuffer gr;
        gr.m_normal_buffer = normals;
        gr.m_albedo_roughness_buffer = d_ar;
        gr.m_metallic_buffer = d_me;
        cas(s, {W, H}, cam, gm, db, la, mb, out_a);
        cas(s, {W, H}, cam, gr, db, la, out_b);
        ctx.sync(s);
        EXPECT(download<uint32_t>(out_a, (size_t) W * H * 4) == download<uint32_t>(out_b, (size_t) W * H * 4),
               "material path == resolved path for constant materials");
    }

    // depth_buffer_reductor: level 0 is the depth copy; the last level is the max over the image's covered texels
    {
        vren::depth_buffer_pyramid pyr(ctx, W, H);
        vren::depth_buffer_reductor red(ctx);
        red.copy_and_reduce(s, db, pyr);
        ctx.sync(s);
        auto d = download<float>(depth, (size_t) W * H);
        auto p = download<float>(pyr.buffer(), (size_t) (pyr.buffer().m_size / 4));
        EXPECT(std::equal(d.begin(), d.end(), p.begin()), "pyramid level 0 == depth");
        const uint32_t last = pyr.get_level_count() - 1;
        EXPECT(pyr.get_image_width(last) == 1 && pyr.get_image_height(last) == 1, "pyramid top is 1x1");
        // level 1 texel (1, 2) = max(0, depth texels (2..3, 4..5)) (depth_buffer_reduce.comp)
        float m = 0.0f;
        for (uint32_t i = 0; i < 2; i++)
            for (uint32_t j = 0; j < 2; j++) {
                const float v = d[(size_t) (4 + j) * W + 2 + i];
                m = m < v ? v : m;
            }
        EXPECT(p[pyr.level_offset(1) + 2 * pyr.get_image_width(1) + 1] == m, "pyramid level 1 texel");
    }

    // show_clusters / visualize_bvh over the cluster_and_shade outputs
    {
        vren_demo::show_clusters sc(ctx);
        sc(s, {W, H}, vren_demo::show_clusters::k_light_assignment_counts, cas, out_a);
        vren_demo::visualize_bvh vb(ctx);
        const uint32_t levels = vren::calc_bvh_level_count(L);
        vren::owned_device_buffer lines(vren_demo::visualize_bvh::get_required_draw_buffer_size(levels));
        vren::resource_container rc2;
        vb.write(s, rc2, cas.m_point_light_bvh_buffer, levels, lines);
        ctx.sync(s);
        auto v = download<float>(lines, 8);
        EXPECT(v[0] <= v[4] && v[1] == v[5] && v[2] == v[6], "first edge of the first leaf runs along x");
    }

    // a camera with near >= far is rejected like the reference's assert
    vren::camera bad = cam;
    bad.m_far_plane = 0.05f;
    EXPECT(throws_invalid_argument([&] { cas(s, {W, H}, bad, gb, db, la, out_a); }), "near >= far");
    EXPECT(throws_invalid_argument([&] { cas(s, {4000, H}, cam, gb, db, la, out_a); }), "screen above limits");
}

int main(int argc, char** argv)
{
    bool host = argc > 1 && std::strcmp(argv[1], "--host") == 0;
    try {
        host_only();
        if (!host) {
            vren::context ctx(0);
            primitives(ctx);
            clustered(ctx);
        }
    } catch (std::exception const& e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 1;
    }
    if (g_failures) {
        std::printf("%d check(s) failed\n", g_failures);
        return 1;
    }
    std::printf("drop-in: all checks passed%s\n", host ? " (host only)" : "");
    return 0;
}
