// Sanitizer driver for the CPU side: the C++ host layer (scene preprocessing,
// SAH / spatial-split / HLBVH builds, 8-wide tree collapse, texture atlas and
// mips, light tree) through the flat rhx_* ABI, and the CPU oracle's
// generate / trace / sort / render / SH / adjoint entry points over the
// resulting scene views.  Built with -fsanitize=address,undefined by the
// Makefile beside it; every buffer handed to the library is sized exactly, so
// a read or write past its end is reported.  Standalone scenes only: nothing
// here touches a GPU.
//
// Exit status 0 when every call returned 0 and the outputs are finite.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "ray_oracle.h"
#include "rh_scene.h"

namespace {

int failures = 0;

void check(int rc, const char *what) {
    if (rc != 0) {
        std::fprintf(stderr, "FAILED %s: rc %d (%s)\n", what, rc, rhx_last_error());
        failures++;
    }
}

struct Lcg {
    uint32_t s;
    float next() {
        s = s * 1664525u + 1013904223u;
        return float(s >> 8) * (1.0f / 16777216.0f);
    }
};

// RGBA8 texture of w x h texels with a checker pattern
std::vector<uint8_t> checker(int w, int h, uint8_t a, uint8_t b) {
    std::vector<uint8_t> t(size_t(w) * h * 4);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            uint8_t v = ((x / 4 + y / 4) & 1) ? a : b;
            uint8_t *p = &t[(size_t(y) * w + x) * 4];
            p[0] = v, p[1] = uint8_t(255 - v), p[2] = uint8_t(v / 2), p[3] = 255;
        }
    return t;
}

// vertex layout PxyzNxyzTuv: 8 floats a vertex
void push_vertex(std::vector<float> &a, float x, float y, float z, float nx, float ny, float nz, float u, float v) {
    const float f[8] = {x, y, z, nx, ny, nz, u, v};
    a.insert(a.end(), f, f + 8);
}

struct MeshData {
    std::vector<float> attrs;
    std::vector<uint32_t> indices;
};

// a UV sphere of radius r
MeshData sphere(float r, int rings, int segs) {
    MeshData m;
    for (int i = 0; i <= rings; i++) {
        const float th = float(M_PI) * i / rings;
        for (int j = 0; j <= segs; j++) {
            const float ph = 2.0f * float(M_PI) * j / segs;
            const float nx = std::sin(th) * std::cos(ph), ny = std::cos(th), nz = std::sin(th) * std::sin(ph);
            push_vertex(m.attrs, r * nx, r * ny, r * nz, nx, ny, nz, float(j) / segs, float(i) / rings);
        }
    }
    for (int i = 0; i < rings; i++)
        for (int j = 0; j < segs; j++) {
            const uint32_t a = i * (segs + 1) + j, b = a + segs + 1;
            const uint32_t q[6] = {a, b, a + 1, a + 1, b, b + 1};
            m.indices.insert(m.indices.end(), q, q + 6);
        }
    return m;
}

// long thin triangles crossing one another (the case spatial splits are for)
MeshData slivers(int n, uint32_t seed) {
    MeshData m;
    Lcg g{seed};
    for (int i = 0; i < n; i++) {
        const float cx = g.next() * 4 - 2, cy = g.next() * 2, cz = g.next() * 4 - 2;
        const float dx = g.next() * 6 - 3, dz = g.next() * 6 - 3;
        const float w = 0.01f + 0.02f * g.next();
        push_vertex(m.attrs, cx - dx, cy, cz - dz, 0, 1, 0, 0, 0);
        push_vertex(m.attrs, cx + dx, cy + w, cz + dz, 0, 1, 0, 1, 0);
        push_vertex(m.attrs, cx + dx, cy - w, cz + dz + w, 0, 1, 0, 1, 1);
        for (int k = 0; k < 3; k++) m.indices.push_back(uint32_t(3 * i + k));
    }
    return m;
}

// an open box of side l centred on the origin, normals inwards
MeshData room(float l) {
    MeshData m;
    const float h = l / 2;
    const float faces[5][4][3] = {
        {{-h, -h, -h}, {h, -h, -h}, {h, -h, h}, {-h, -h, h}},  // floor
        {{-h, h, -h}, {-h, h, h}, {h, h, h}, {h, h, -h}},      // ceiling
        {{-h, -h, -h}, {-h, -h, h}, {-h, h, h}, {-h, h, -h}},  // left
        {{h, -h, -h}, {h, h, -h}, {h, h, h}, {h, -h, h}},      // right
        {{-h, -h, h}, {h, -h, h}, {h, h, h}, {-h, h, h}},      // back
    };
    const float normals[5][3] = {{0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, -1}};
    for (int f = 0; f < 5; f++) {
        const uint32_t base = uint32_t(m.attrs.size() / 8);
        for (int k = 0; k < 4; k++)
            push_vertex(m.attrs, faces[f][k][0], faces[f][k][1], faces[f][k][2], normals[f][0], normals[f][1],
                        normals[f][2], float(k & 1), float(k >> 1));
        const uint32_t q[6] = {base, base + 1, base + 2, base, base + 2, base + 3};
        m.indices.insert(m.indices.end(), q, q + 6);
    }
    return m;
}

rhx_mat_desc material(int type, float r, float g, float b, uint32_t tex = 0xffffffffu) {
    rhx_mat_desc d;
    std::memset(&d, 0, sizeof(d));
    d.type = type;
    d.main_color[0] = r, d.main_color[1] = g, d.main_color[2] = b;
    d.main_texture = tex;
    d.normal_map = 0xffffffffu;
    d.mix_materials[0] = d.mix_materials[1] = 0xffffffffu;
    d.roughness = 0.1f;
    d.strength = 1.0f;
    d.fresnel = 1.0f;
    d.int_ior = 1.5f;
    d.ext_ior = 1.0f;
    return d;
}

rhx_cam_desc camera(int type, int filter, int output_sh) {
    rhx_cam_desc c;
    std::memset(&c, 0, sizeof(c));
    c.type = type, c.filter = filter, c.dtype = RH_DEVICE_SRGB;
    const float o[3] = {0.3f, 0.5f, -4.5f}, f[3] = {0, -0.05f, 1}, u[3] = {0, 1, 0};
    std::memcpy(c.origin, o, sizeof(o));
    std::memcpy(c.fwd, f, sizeof(f));
    std::memcpy(c.up, u, sizeof(u));
    c.fov = 45.0f, c.gamma = 2.2f, c.focus_distance = 1.0f, c.focus_factor = 0.0f;
    c.mi_index = c.uv_index = 0;
    c.output_sh = uint8_t(output_sh);
    c.max_diff_depth = 4, c.max_glossy_depth = 4, c.max_refr_depth = 6, c.max_transp_depth = 6;
    c.max_total_depth = 8, c.termination_start_depth = 2;
    return c;
}

uint32_t add_mesh(rhx_scene *s, const MeshData &m, const uint32_t *mats, int n_shapes, int flags) {
    std::vector<rhx_shape_desc> shapes(n_shapes);
    const size_t per = (m.indices.size() / 3 / n_shapes) * 3;
    for (int i = 0; i < n_shapes; i++) {
        shapes[i].mat_index = mats[i];
        shapes[i].back_mat_index = mats[i];
        shapes[i].vtx_start = per * i;
        shapes[i].vtx_count = i + 1 < n_shapes ? per : m.indices.size() - per * i;
    }
    return rhx_add_mesh(s, 0, m.attrs.data(), m.attrs.size() / 8, m.indices.data(), m.indices.size(), 0, shapes.data(),
                        shapes.size(), flags);
}

double checksum(const float *p, size_t n) {
    double s = 0;
    for (size_t i = 0; i < n; i++) {
        if (!std::isfinite(p[i])) {
            std::fprintf(stderr, "non-finite value at %zu\n", i);
            failures++;
            return 0;
        }
        s += p[i];
    }
    return s;
}

// Builds one scene and runs every oracle stage over it.
void run_scene(int wide, int variant) {
    rhx_scene *s = nullptr;
    check(rhx_create_scene_standalone(wide, &s), "rhx_create_scene_standalone");
    if (!s) return;

    const float env[3] = {0.2f, 0.25f, 0.3f};
    check(rhx_set_environment(s, env, 10.0f, 0xffffffffu), "rhx_set_environment");

    // textures: a power-of-two sRGB one with mips, an odd-sized linear one
    // with mips, a 1x1 one without
    const auto t0 = checker(64, 64, 200, 40), t1 = checker(37, 23, 180, 90), t2 = checker(1, 1, 255, 255);
    const uint32_t tex0 = rhx_add_texture(s, t0.data(), 64, 64, 1, 1);
    const uint32_t tex1 = rhx_add_texture(s, t1.data(), 37, 23, 0, 1);
    const uint32_t tex2 = rhx_add_texture(s, t2.data(), 1, 1, 0, 0);

    std::vector<uint32_t> mats;
    rhx_mat_desc d = material(RH_MAT_DIFFUSE, 0.8f, 0.7f, 0.6f, tex0);
    mats.push_back(rhx_add_material(s, &d));
    d = material(RH_MAT_GLOSSY, 0.9f, 0.9f, 0.9f, tex1);
    mats.push_back(rhx_add_material(s, &d));
    d = material(RH_MAT_REFRACTIVE, 1.0f, 1.0f, 1.0f, tex2);
    mats.push_back(rhx_add_material(s, &d));
    d = material(RH_MAT_EMISSIVE, 1.0f, 0.9f, 0.8f, tex2);
    d.strength = 4.0f;
    const uint32_t emissive = rhx_add_material(s, &d);
    d = material(RH_MAT_TRANSPARENT, 1.0f, 1.0f, 1.0f, tex2);
    mats.push_back(rhx_add_material(s, &d));
    d = material(RH_MAT_MIX, 1.0f, 1.0f, 1.0f, tex1);
    d.strength = 0.4f;
    d.mix_materials[0] = mats[0], d.mix_materials[1] = mats[1];
    mats.push_back(rhx_add_material(s, &d));
    d = material(RH_MAT_DIFFUSE, 0.5f, 0.5f, 0.5f, tex0);
    d.normal_map = tex1;
    mats.push_back(rhx_add_material(s, &d));
    for (uint32_t m : mats)
        if (m == 0xffffffffu) failures++;

    // indices that name nothing are refused, and the scene is left as it was
    d = material(RH_MAT_DIFFUSE, 1, 1, 1);
    if (rhx_add_material(s, &d) != 0xffffffffu) failures++;
    d = material(RH_MAT_MIX, 1, 1, 1, tex0);
    d.mix_materials[0] = mats[0], d.mix_materials[1] = 1000;
    if (rhx_add_material(s, &d) != 0xffffffffu) failures++;
    {
        const MeshData tri = slivers(1, 3u);
        const uint32_t bad_index[3] = {0, 1, 3};
        rhx_shape_desc sh = {mats[0], 0xffffffffu, 0, 3};
        if (rhx_add_mesh(s, 0, tri.attrs.data(), 3, bad_index, 3, 0, &sh, 1, 0) != 0xffffffffu) failures++;
        sh.mat_index = 77;
        if (rhx_add_mesh(s, 0, tri.attrs.data(), 3, tri.indices.data(), 3, 0, &sh, 1, 0) != 0xffffffffu) failures++;
        sh.mat_index = mats[0], sh.vtx_count = 6;
        if (rhx_add_mesh(s, 0, tri.attrs.data(), 3, tri.indices.data(), 3, 0, &sh, 1, 0) != 0xffffffffu) failures++;
    }
    if (rhx_set_environment(s, env, 10.0f, 12345u) == 0) failures++;

    const MeshData walls = room(4.0f), ball = sphere(0.6f, 12 + variant, 18), sl = slivers(300, 7u + variant);
    const uint32_t wall_mats[2] = {mats[0], mats[5]};
    const uint32_t m_room = add_mesh(s, walls, wall_mats, 2, 0);
    const uint32_t ball_mats[3] = {mats[1], mats[2], mats[3]};
    const uint32_t m_ball = add_mesh(s, ball, ball_mats, 3, RHX_MESH_FAST_BVH);
    const uint32_t sl_mats[1] = {mats[4]};
    const uint32_t m_sl = add_mesh(s, sl, sl_mats, 1, RHX_MESH_SPATIAL_SPLITS);
    const MeshData lamp = sphere(0.15f, 6, 8);
    const uint32_t lamp_mats[1] = {emissive};
    const uint32_t m_lamp = add_mesh(s, lamp, lamp_mats, 1, 0);

    float xf[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    rhx_add_mesh_instance(s, m_room, xf);
    xf[12] = -0.8f, xf[13] = -1.3f, xf[14] = 0.5f;
    const uint32_t ball_a = rhx_add_mesh_instance(s, m_ball, xf);
    xf[12] = 0.9f, xf[13] = -1.2f, xf[14] = -0.2f;
    xf[0] = 0.7f, xf[5] = 1.2f, xf[10] = 0.9f;
    rhx_add_mesh_instance(s, m_ball, xf);
    float xs[16] = {0.5f, 0, 0, 0, 0, 0.5f, 0, 0, 0, 0, 0.5f, 0, 0, -0.5f, 0.8f, 1};
    rhx_add_mesh_instance(s, m_sl, xs);
    float xl[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0.2f, 1.7f, 0.3f, 1};
    rhx_add_mesh_instance(s, m_lamp, xl);

    // scene edits: move one instance, add and remove another
    xf[12] = -0.7f;
    rhx_set_mesh_instance_transform(s, ball_a, xf);
    const uint32_t extra = rhx_add_mesh_instance(s, m_lamp, xl);
    rhx_remove_mesh_instance(s, extra);

    rhx_light_desc l;
    std::memset(&l, 0, sizeof(l));
    l.type = RH_LIGHT_POINT;
    l.position[0] = 0.5f, l.position[1] = 1.5f, l.position[2] = -1.0f, l.radius = 0.1f;
    l.color[0] = l.color[1] = l.color[2] = 3.0f;
    rhx_add_light(s, &l);
    l.type = RH_LIGHT_SPOT;
    l.position[0] = -1.0f, l.direction[1] = -1.0f, l.angle = 40.0f;
    rhx_add_light(s, &l);
    l.type = RH_LIGHT_DIRECTIONAL;
    l.direction[0] = 0.3f, l.direction[1] = -1.0f, l.direction[2] = 0.2f, l.angle = 5.0f;
    rhx_add_light(s, &l);

    const rhx_cam_desc c0 = camera(variant ? RH_CAM_ORTHO : RH_CAM_PERSP, RH_FILTER_TENT, 0);
    const rhx_cam_desc c1 = camera(RH_CAM_PERSP, RH_FILTER_BOX, 1);
    const uint32_t cam0 = rhx_add_camera(s, &c0), cam1 = rhx_add_camera(s, &c1);
    rhx_set_current_cam(s, cam0);

    rh_scene_view v;
    check(rhx_scene_view(s, &v), "rhx_scene_view");
    rh_camera cam;
    check(rhx_get_camera_record(s, cam0, &cam), "rhx_get_camera_record");
    std::printf("scene wide=%d variant=%d: %u triangles, %u nodes, %u mnodes, %u texels\n", wide, variant,
                rhx_triangle_count(s), v.nodes_count, v.mnodes_count, v.texels_count);

    // primary rays of a sub-rectangle, traced closest-hit and any-hit
    const int w = 40, h = 30;
    const int32_t rect[4] = {3, 2, 31, 25};
    std::vector<float> halton(RH_HALTON_COUNT * RH_HALTON_SEQ_LEN);
    check(orc_halton_table(5, halton.data()), "orc_halton_table");
    const size_t n = size_t(rect[2]) * rect[3];
    std::vector<rh_ray_packet> rays(n);
    check(orc_gen_primary(&cam, rect, w, h, 5, halton.data(), rays.data()) == int(n) ? 0 : -1, "orc_gen_primary");
    std::vector<rh_hit> hits(n);
    orc_counters cnt{};
    check(orc_trace(&v, rays.data(), n, 0, nullptr, hits.data(), &cnt), "orc_trace");
    std::vector<float> tmax(n, 2.0f);
    check(orc_trace(&v, rays.data(), n, 1, tmax.data(), hits.data(), nullptr), "orc_trace any-hit");
    std::vector<uint32_t> hashes(n);
    const float root_min[3] = {-2, -2, -2}, cell[3] = {0.25f, 0.25f, 0.25f};
    check(orc_ray_hashes(rays.data(), n, root_min, cell, hashes.data()), "orc_ray_hashes");
    check(orc_sort_rays(rays.data(), n, root_min, cell), "orc_sort_rays");
    check(orc_ray_hashes(rays.data(), n, root_min, cell, hashes.data()), "orc_ray_hashes sorted");
    for (size_t i = 1; i < n; i++)
        if (hashes[i - 1] > hashes[i]) {
            std::fprintf(stderr, "rays not sorted by hash at %zu\n", i);
            failures++;
            break;
        }

    // full passes, then OutputSH passes, then the adjoint with texel gradients
    std::vector<float> clean(size_t(w) * h * 4), final_rgba(size_t(w) * h * 4);
    unsigned long long counts[3] = {0, 0, 0};
    int iteration = 0;
    check(orc_set_max_bounces(variant ? 12 : 10), "orc_set_max_bounces");
    check(orc_render(&v, &cam, w, h, rect, 3, &iteration, clean.data(), final_rgba.data(), counts), "orc_render");
    const double sum = checksum(clean.data(), clean.size());

    rh_camera cam_sh;
    check(rhx_get_camera_record(s, cam1, &cam_sh), "rhx_get_camera_record sh");
    std::vector<float> clean_sh(size_t(w) * h * 4), final_sh(size_t(w) * h * 4);
    std::vector<float> sh_temp(size_t(w) * h * 12), sh_clean(size_t(w) * h * 12);
    int it_sh = 0;
    check(orc_render_sh(&v, &cam_sh, w, h, rect, 2, &it_sh, clean_sh.data(), final_sh.data(), counts, sh_temp.data(),
                        sh_clean.data()),
          "orc_render_sh");
    checksum(sh_clean.data(), sh_clean.size());

    std::vector<float> adj(size_t(w) * h * 4, 1.0f);
    std::vector<double> grad(size_t(v.materials_count) * 3), tgrad(size_t(v.texels_count) * 4);
    check(orc_adjoint(&v, &cam, w, h, rect, 2, adj.data(), 0.5, grad.data(), tgrad.data()), "orc_adjoint");
    check(orc_set_max_bounces(10), "orc_set_max_bounces");
    double g = 0;
    for (double x : grad) g += x;
    std::printf("  rays %zu: %llu node visits, %llu tri tests; image sum %.4f; rays %llu/%llu/%llu; grad sum %.4f\n", n,
                cnt.nodes, cnt.tris, sum, counts[0], counts[1], counts[2], g);

    // removing a mesh together with its instances, and a rebuild of the view
    rhx_remove_mesh(s, m_sl);
    check(rhx_scene_view(s, &v), "rhx_scene_view after removal");
    check(orc_trace(&v, rays.data(), n, 0, nullptr, hits.data(), nullptr), "orc_trace after removal");
    rhx_destroy_scene(s);
}

// the host HLBVH builder of rhx_build_fast_bvh (device < 0)
void run_fast_bvh() {
    Lcg g{99u};
    const size_t n = 1000;
    std::vector<float> boxes(6 * n);
    for (size_t i = 0; i < n; i++) {
        const float x = g.next() * 10, y = g.next() * 10, z = g.next() * 10, e = 0.01f + g.next() * 0.3f;
        const float b[6] = {x, y, z, x + e, y + e, z + e};
        std::memcpy(&boxes[6 * i], b, sizeof(b));
    }
    std::vector<rh_bvh_node> nodes(2 * n);
    std::vector<uint32_t> indices(n);
    uint32_t count = 0;
    check(rhx_build_fast_bvh(boxes.data(), n, -1, nodes.data(), nodes.size(), &count, indices.data()),
          "rhx_build_fast_bvh");
    std::printf("fast bvh: %zu boxes, %u nodes\n", n, count);
    if (count == 0 || count > 2 * n) failures++;
}

}  // namespace

int main() {
    for (int wide = 0; wide < 2; wide++)
        for (int variant = 0; variant < 2; variant++) run_scene(wide, variant);
    run_fast_bvh();
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("sanitized host layer and oracle: OK\n");
    return 0;
}
