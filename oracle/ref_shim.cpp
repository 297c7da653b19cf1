/*
 * ref_shim.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * Builds the reference Ray library (its CPU "Ref" backend) from the sources
 * where they lie under /root/reference/src/Ray and exposes it through the same
 * flat C ABI as the HIP backend (include/rh_scene.h), plus a few entry points
 * that reach the reference's internal per-ray functions.  The result,
 * oracle/_ref/libray_ref.so, is the checker the CPU restatement
 * (oracle/ray_oracle.cpp) and the GPU path are compared against.  Nothing in
 * the product links or loads it.
 *
 * This translation unit includes the reference's unity source _Ray.cpp so the
 * file-local inline helpers (get_ray_hash, hash, construct_float) are visible.
 */
#include "_Ray.cpp"

#include <cstring>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "internal/Halton.h"

#include "rh_scene.h"

namespace {
thread_local std::string g_err;

struct SceneSpy : public Ray::Ref::Scene {
    explicit SceneSpy(bool wide) : Ray::Ref::Scene(wide) {}
    using Ray::Ref::Scene::env_;
    using Ray::Ref::Scene::light_nodes_root_;
    using Ray::Ref::Scene::li_indices_;
    using Ray::Ref::Scene::lights_;
    using Ray::Ref::Scene::macro_nodes_root_;
    using Ray::Ref::Scene::materials_;
    using Ray::Ref::Scene::mesh_instances_;
    using Ray::Ref::Scene::meshes_;
    using Ray::Ref::Scene::mi_indices_;
    using Ray::Ref::Scene::nodes_;
    using Ray::Ref::Scene::mnodes_;
    using Ray::Ref::Scene::texture_atlas_;
    using Ray::Ref::Scene::textures_;
    using Ray::Ref::Scene::transforms_;
    using Ray::Ref::Scene::tri_indices_;
    using Ray::Ref::Scene::tris_;
    using Ray::Ref::Scene::vertices_;
    using Ray::Ref::Scene::vtx_indices_;
    using Ray::Ref::Scene::cams_;
};

std::vector<uint16_t> make_perms() {
    auto rand_func = std::bind(std::uniform_int_distribution<int>(), std::mt19937(0));
    return Ray::ComputeRadicalInversePermutations(Ray::g_primes, Ray::PrimesCount, rand_func);
}
} // namespace

struct rhx_renderer {
    std::shared_ptr<Ray::RendererBase> r;
    bool wide = false;
    int w = 0, h = 0;
};

struct rhx_scene {
    std::shared_ptr<SceneSpy> s;
};

struct rhx_region {
    Ray::RegionContext ctx;
    explicit rhx_region(const Ray::rect_t &rc) : ctx(rc) {}
};

extern "C" {

const char *rhx_last_error(void) { return g_err.c_str(); }

int rhx_create_renderer(const rhx_settings *s, uint32_t flags, rhx_renderer **out) {
    Ray::settings_t st;
    st.w = s->w;
    st.h = s->h;
    st.use_wide_bvh = s->use_wide_bvh != 0;
    auto *r = new rhx_renderer;
    r->wide = st.use_wide_bvh;
    r->w = s->w;
    r->h = s->h;
    uint32_t f = flags & (Ray::RendererRef | Ray::RendererSSE2 | Ray::RendererAVX | Ray::RendererAVX2);
    if (!f) f = Ray::RendererRef;
    std::ostringstream log;
    r->r = Ray::CreateRenderer(st, f, log);
    *out = r;
    return 0;
}

void rhx_destroy_renderer(rhx_renderer *r) { delete r; }
uint32_t rhx_renderer_type(const rhx_renderer *r) { return (uint32_t)r->r->type(); }
int rhx_resize(rhx_renderer *r, int w, int h) {
    r->r->Resize(w, h);
    r->w = w;
    r->h = h;
    return 0;
}
int rhx_clear(rhx_renderer *r, const float rgba[4]) {
    Ray::pixel_color_t c = {rgba[0], rgba[1], rgba[2], rgba[3]};
    r->r->Clear(c);
    return 0;
}
int rhx_get_pixels(rhx_renderer *r, float *out, size_t count) {
    auto sz = r->r->size();
    size_t n = (size_t)sz.first * sz.second * 4;
    if (count < n) {
        g_err = "output buffer too small";
        return -1;
    }
    memcpy(out, r->r->get_pixels_ref(), n * sizeof(float));
    return 0;
}
int rhx_get_sh_data(rhx_renderer *r, float *out, size_t count) {
    auto sz = r->r->size();
    size_t n = (size_t)sz.first * sz.second * 12;
    if (count < n) {
        g_err = "output buffer too small";
        return -1;
    }
    // the Ref framebuffer allocates its SH data at the first pass with OutputSH
    const Ray::shl1_data_t *sh = r->r->get_sh_data_ref();
    if (!sh) return 1;
    memcpy(out, sh, n * sizeof(float));
    return 0;
}
int rhx_get_stats(rhx_renderer *r, rh_stats *out) {
    Ray::RendererBase::stats_t st;
    r->r->GetStats(st);
    static_assert(sizeof(st) == sizeof(rh_stats), "stats layout");
    memcpy(out, &st, sizeof(st));
    return 0;
}
int rhx_reset_stats(rhx_renderer *r) {
    r->r->ResetStats();
    return 0;
}
int rhx_get_ray_counts(rhx_renderer *, rh_ray_counts *out) {
    memset(out, 0, sizeof(*out));
    return -1;
}
int rhx_set_option(rhx_renderer *, const char *, int) { return -1; }
int rhx_synchronize(rhx_renderer *) { return 0; }
int rhx_get_kernel_profile(rhx_renderer *, rh_kernel_profile *out) {
    memset(out, 0, sizeof(*out));
    return -1;
}
int rhx_reset_kernel_profile(rhx_renderer *) { return 0; }

int rhx_create_scene(rhx_renderer *r, rhx_scene **out) {
    auto *s = new rhx_scene;
    s->s = std::make_shared<SceneSpy>(r ? r->wide : false);
    *out = s;
    return 0;
}
int rhx_create_scene_standalone(int use_wide_bvh, rhx_scene **out) {
    *out = new rhx_scene{std::make_shared<SceneSpy>(use_wide_bvh != 0)};
    return 0;
}
void rhx_destroy_scene(rhx_scene *s) { delete s; }

int rhx_region_create(int x, int y, int w, int h, rhx_region **out) {
    *out = new rhx_region(Ray::rect_t{x, y, w, h});
    return 0;
}
void rhx_region_destroy(rhx_region *reg) { delete reg; }
int rhx_region_iteration(const rhx_region *reg) { return reg->ctx.iteration; }
int rhx_region_clear(rhx_region *reg) {
    reg->ctx.Clear();
    return 0;
}
int rhx_region_halton(const rhx_region *reg, float *out, size_t count) {
    if (!reg->ctx.halton_seq || count < Ray::HALTON_COUNT * Ray::HALTON_SEQ_LEN) return -1;
    memcpy(out, reg->ctx.halton_seq.get(), sizeof(float) * Ray::HALTON_COUNT * Ray::HALTON_SEQ_LEN);
    return 0;
}

int rhx_render_scene(rhx_renderer *r, rhx_scene *s, rhx_region *reg) {
    r->r->RenderScene(s->s, reg->ctx);
    return 0;
}

int rhx_set_environment(rhx_scene *s, const float env_col[3], float env_clamp, uint32_t env_map) {
    Ray::environment_desc_t e;
    memcpy(e.env_col, env_col, 3 * sizeof(float));
    e.env_clamp = env_clamp;
    e.env_map = env_map;
    s->s->SetEnvironment(e);
    return 0;
}

uint32_t rhx_add_texture(rhx_scene *s, const uint8_t *rgba, int w, int h, int is_srgb, int gen_mips) {
    Ray::tex_desc_t t;
    t.data = reinterpret_cast<const Ray::pixel_color8_t *>(rgba);
    t.w = w;
    t.h = h;
    t.is_srgb = is_srgb != 0;
    t.generate_mipmaps = gen_mips != 0;
    return s->s->AddTexture(t);
}

uint32_t rhx_add_material(rhx_scene *s, const rhx_mat_desc *m) {
    Ray::mat_desc_t d;
    d.type = (Ray::eMaterialType)m->type;
    memcpy(d.main_color, m->main_color, 3 * sizeof(float));
    d.main_texture = m->main_texture;
    d.normal_map = m->normal_map;
    d.mix_materials[0] = m->mix_materials[0];
    d.mix_materials[1] = m->mix_materials[1];
    d.roughness = m->roughness;
    d.strength = m->strength;
    d.fresnel = m->fresnel;
    d.int_ior = m->int_ior;
    d.ext_ior = m->ext_ior;
    return s->s->AddMaterial(d);
}

uint32_t rhx_add_mesh(rhx_scene *s, int layout, const float *attrs, size_t attrs_count, const uint32_t *indices,
                      size_t indices_count, int base_vertex, const rhx_shape_desc *shapes, size_t shapes_count,
                      int bvh_flags) {
    Ray::mesh_desc_t d;
    d.prim_type = Ray::TriangleList;
    d.layout = (Ray::eVertexLayout)layout;
    d.vtx_attrs = attrs;
    d.vtx_attrs_count = attrs_count;
    d.vtx_indices = indices;
    d.vtx_indices_count = indices_count;
    d.base_vertex = base_vertex;
    d.allow_spatial_splits = (bvh_flags & 1) != 0;
    d.use_fast_bvh_build = (bvh_flags & 2) != 0;
    for (size_t i = 0; i < shapes_count; i++) {
        d.shapes.emplace_back(shapes[i].mat_index, shapes[i].back_mat_index, (size_t)shapes[i].vtx_start,
                              (size_t)shapes[i].vtx_count);
    }
    return s->s->AddMesh(d);
}

void rhx_remove_mesh(rhx_scene *s, uint32_t i) { s->s->RemoveMesh(i); }

uint32_t rhx_add_light(rhx_scene *s, const rhx_light_desc *l) {
    Ray::light_desc_t d;
    d.type = (Ray::eLightType)l->type;
    memcpy(d.position, l->position, sizeof(d.position));
    d.radius = l->radius;
    memcpy(d.color, l->color, sizeof(d.color));
    memcpy(d.direction, l->direction, sizeof(d.direction));
    d.angle = l->angle;
    return s->s->AddLight(d);
}

uint32_t rhx_add_mesh_instance(rhx_scene *s, uint32_t mesh, const float xform[16]) {
    return s->s->AddMeshInstance(mesh, xform);
}
void rhx_set_mesh_instance_transform(rhx_scene *s, uint32_t mi, const float xform[16]) {
    s->s->SetMeshInstanceTransform(mi, xform);
}
void rhx_remove_mesh_instance(rhx_scene *s, uint32_t mi) { s->s->RemoveMeshInstance(mi); }

static Ray::camera_desc_t to_cam_desc(const rhx_cam_desc *c) {
    Ray::camera_desc_t d;
    d.type = (Ray::eCamType)c->type;
    d.filter = (Ray::eFilterType)c->filter;
    d.dtype = (Ray::eDeviceType)c->dtype;
    memcpy(d.origin, c->origin, sizeof(d.origin));
    memcpy(d.fwd, c->fwd, sizeof(d.fwd));
    memcpy(d.up, c->up, sizeof(d.up));
    d.fov = c->fov;
    d.gamma = c->gamma;
    d.focus_distance = c->focus_distance;
    d.focus_factor = c->focus_factor;
    d.mi_index = c->mi_index;
    d.uv_index = c->uv_index;
    d.lighting_only = c->lighting_only;
    d.skip_direct_lighting = c->skip_direct_lighting;
    d.skip_indirect_lighting = c->skip_indirect_lighting;
    d.no_background = c->no_background;
    d.clamp = c->clamp;
    d.output_sh = c->output_sh;
    d.use_coherent_sampling = c->use_coherent_sampling;
    d.max_diff_depth = c->max_diff_depth;
    d.max_glossy_depth = c->max_glossy_depth;
    d.max_refr_depth = c->max_refr_depth;
    d.max_transp_depth = c->max_transp_depth;
    d.max_total_depth = c->max_total_depth;
    d.termination_start_depth = c->termination_start_depth;
    return d;
}

uint32_t rhx_add_camera(rhx_scene *s, const rhx_cam_desc *c) { return s->s->AddCamera(to_cam_desc(c)); }
void rhx_set_camera(rhx_scene *s, uint32_t i, const rhx_cam_desc *c) { s->s->SetCamera(i, to_cam_desc(c)); }
void rhx_set_current_cam(rhx_scene *s, uint32_t i) { s->s->set_current_cam(i); }
int rhx_get_camera_record(rhx_scene *s, uint32_t i, rh_camera *out) {
    static_assert(sizeof(Ray::camera_t) == sizeof(rh_camera), "camera layout");
    memcpy(out, &s->s->cams_[i].cam, sizeof(rh_camera));
    return 0;
}
uint32_t rhx_triangle_count(rhx_scene *s) { return s->s->triangle_count(); }
uint32_t rhx_node_count(rhx_scene *s) { return s->s->node_count(); }

int rhx_scene_view(rhx_scene *s, rh_scene_view *v) {
    SceneSpy &sc = *s->s;
    memset(v, 0, sizeof(*v));
#define RH_VIEW(field, vec, T)                                                                                         \
    v->field = vec.empty() ? nullptr : reinterpret_cast<const T *>(vec.data());                                        \
    v->field##_count = (uint32_t)vec.size();
    v->env = reinterpret_cast<const rh_environment *>(&sc.env_);
    RH_VIEW(mesh_instances, sc.mesh_instances_, rh_mesh_instance)
    RH_VIEW(mi_indices, sc.mi_indices_, uint32_t)
    RH_VIEW(meshes, sc.meshes_, rh_mesh)
    RH_VIEW(transforms, sc.transforms_, rh_transform)
    RH_VIEW(vtx_indices, sc.vtx_indices_, uint32_t)
    RH_VIEW(vertices, sc.vertices_, rh_vertex)
    RH_VIEW(nodes, sc.nodes_, rh_bvh_node)
    RH_VIEW(mnodes, sc.mnodes_, rh_mbvh_node)
    RH_VIEW(tris, sc.tris_, rh_tri_accel)
    RH_VIEW(tri_indices, sc.tri_indices_, uint32_t)
    RH_VIEW(materials, sc.materials_, rh_material)
    RH_VIEW(textures, sc.textures_, rh_texture)
    RH_VIEW(lights, sc.lights_, rh_light)
    RH_VIEW(li_indices, sc.li_indices_, uint32_t)
#undef RH_VIEW
    v->macro_nodes_root = sc.macro_nodes_root_;
    v->light_nodes_root = sc.light_nodes_root_;
    return 0;
}

/* ---------------- reference internals, per ray ---------------- */

/* Halton table of RendererRef.cpp:327-358 for a given iteration. */
int ref_halton_table(int iteration, float *out) {
    static const std::vector<uint16_t> perms = make_perms();
    // UpdateHaltonSequence is private to the renderer; this drives the same
    // reference templates with the same permutation table and index layout.
    const uint16_t *p = perms.data();
    for (int i = 0; i < Ray::HALTON_SEQ_LEN; i++) {
        uint64_t a = (uint64_t)(iteration + i);
        float *row = out + i * Ray::HALTON_COUNT;
        row[0] = Ray::ScrambledRadicalInverse<2>(p + 0, a);
        row[1] = Ray::ScrambledRadicalInverse<3>(p + 2, a);
        row[2] = Ray::ScrambledRadicalInverse<5>(p + 5, a);
        row[3] = Ray::ScrambledRadicalInverse<7>(p + 10, a);
        row[4] = Ray::ScrambledRadicalInverse<11>(p + 17, a);
        row[5] = Ray::ScrambledRadicalInverse<13>(p + 28, a);
        row[6] = Ray::ScrambledRadicalInverse<17>(p + 41, a);
        row[7] = Ray::ScrambledRadicalInverse<19>(p + 58, a);
        row[8] = Ray::ScrambledRadicalInverse<23>(p + 77, a);
        row[9] = Ray::ScrambledRadicalInverse<29>(p + 100, a);
        row[10] = Ray::ScrambledRadicalInverse<31>(p + 129, a);
        row[11] = Ray::ScrambledRadicalInverse<37>(p + 160, a);
        row[12] = Ray::ScrambledRadicalInverse<41>(p + 197, a);
        row[13] = Ray::ScrambledRadicalInverse<43>(p + 238, a);
        row[14] = Ray::ScrambledRadicalInverse<47>(p + 281, a);
        row[15] = Ray::ScrambledRadicalInverse<53>(p + 328, a);
        row[16] = Ray::ScrambledRadicalInverse<59>(p + 381, a);
        row[17] = Ray::ScrambledRadicalInverse<61>(p + 440, a);
        row[18] = Ray::ScrambledRadicalInverse<67>(p + 501, a);
        row[19] = Ray::ScrambledRadicalInverse<71>(p + 568, a);
        row[20] = Ray::ScrambledRadicalInverse<73>(p + 639, a);
        row[21] = Ray::ScrambledRadicalInverse<79>(p + 712, a);
        row[22] = Ray::ScrambledRadicalInverse<83>(p + 791, a);
        row[23] = Ray::ScrambledRadicalInverse<89>(p + 874, a);
    }
    return 0;
}

/* The 963-entry permutation table (mt19937(0) + libstdc++ uniform_int_distribution). */
int ref_halton_perms(uint16_t *out, size_t count) {
    static const std::vector<uint16_t> perms = make_perms();
    if (count < perms.size()) return -1;
    memcpy(out, perms.data(), perms.size() * sizeof(uint16_t));
    return (int)perms.size();
}

int ref_gen_primary(const rh_camera *cam, int rx, int ry, int rw, int rh, int w, int h, int iteration,
                    const float *halton, rh_ray_packet *out) {
    Ray::aligned_vector<Ray::Ref::ray_packet_t> rays;
    Ray::camera_t c;
    memcpy(&c, cam, sizeof(c));
    Ray::Ref::GeneratePrimaryRays(iteration, c, Ray::rect_t{rx, ry, rw, rh}, w, h, halton, rays);
    static_assert(sizeof(Ray::Ref::ray_packet_t) == sizeof(rh_ray_packet), "ray layout");
    memcpy(out, rays.data(), rays.size() * sizeof(rh_ray_packet));
    return (int)rays.size();
}

static Ray::scene_data_t scene_data(SceneSpy &sc) {
    Ray::scene_data_t d;
    d.env = &sc.env_;
    d.mesh_instances = sc.mesh_instances_.empty() ? nullptr : sc.mesh_instances_.data();
    d.mi_indices = sc.mi_indices_.empty() ? nullptr : sc.mi_indices_.data();
    d.meshes = sc.meshes_.empty() ? nullptr : sc.meshes_.data();
    d.transforms = sc.transforms_.empty() ? nullptr : sc.transforms_.data();
    d.vtx_indices = sc.vtx_indices_.empty() ? nullptr : sc.vtx_indices_.data();
    d.vertices = sc.vertices_.empty() ? nullptr : sc.vertices_.data();
    d.nodes = sc.nodes_.empty() ? nullptr : sc.nodes_.data();
    d.mnodes = sc.mnodes_.empty() ? nullptr : sc.mnodes_.data();
    d.tris = sc.tris_.empty() ? nullptr : sc.tris_.data();
    d.tri_indices = sc.tri_indices_.empty() ? nullptr : sc.tri_indices_.data();
    d.materials = sc.materials_.empty() ? nullptr : sc.materials_.data();
    d.textures = sc.textures_.empty() ? nullptr : sc.textures_.data();
    d.lights = sc.lights_.empty() ? nullptr : sc.lights_.data();
    d.li_indices = sc.li_indices_.empty() ? nullptr : sc.li_indices_.data();
    return d;
}

/* Closest-hit (any_hit == 0) or any-hit traversal of the macro/micro trees
 * (binary, or 8-wide when the scene was built with use_wide_bvh). */
int ref_trace(rhx_scene *s, const rh_ray_packet *rays, size_t n, int any_hit, const float *tmax, rh_hit *out) {
    SceneSpy &sc = *s->s;
    const bool wide = !sc.mnodes_.empty();
    for (size_t i = 0; i < n; i++) {
        Ray::Ref::ray_packet_t r;
        memcpy(&r, &rays[i], sizeof(r));
        Ray::Ref::hit_data_t inter;
        inter.xy = r.xy;
        if (tmax) inter.t = tmax[i];
        if (sc.macro_nodes_root_ != 0xffffffff && wide) {
            // the 8-wide trees (CoreRef.cpp:1133-1222, 1270-1368)
            if (!any_hit) {
                Ray::Ref::Traverse_MacroTree_WithStack_ClosestHit(
                    r, sc.mnodes_.data(), sc.macro_nodes_root_, sc.mesh_instances_.data(), sc.mi_indices_.data(),
                    sc.meshes_.data(), sc.transforms_.data(), sc.tris_.data(), sc.tri_indices_.data(), inter);
            } else {
                Ray::Ref::Traverse_MacroTree_WithStack_AnyHit(
                    r, sc.mnodes_.data(), sc.macro_nodes_root_, sc.mesh_instances_.data(), sc.mi_indices_.data(),
                    sc.meshes_.data(), sc.transforms_.data(), sc.tris_.data(), sc.tri_indices_.data(), inter);
            }
        } else if (sc.macro_nodes_root_ != 0xffffffff) {
            if (!any_hit) {
                Ray::Ref::Traverse_MacroTree_WithStack_ClosestHit(
                    r, sc.nodes_.data(), sc.macro_nodes_root_, sc.mesh_instances_.data(), sc.mi_indices_.data(),
                    sc.meshes_.data(), sc.transforms_.data(), sc.tris_.data(), sc.tri_indices_.data(), inter);
            } else {
                Ray::Ref::Traverse_MacroTree_WithStack_AnyHit(
                    r, sc.nodes_.data(), sc.macro_nodes_root_, sc.mesh_instances_.data(), sc.mi_indices_.data(),
                    sc.meshes_.data(), sc.transforms_.data(), sc.tris_.data(), sc.tri_indices_.data(), inter);
            }
        }
        out[i].mask = inter.mask_values[0];
        out[i].obj = inter.obj_indices[0];
        out[i].prim = inter.prim_indices[0];
        out[i].t = inter.t;
        out[i].u = inter.u;
        out[i].v = inter.v;
        out[i].xy = inter.xy;
        out[i].pad = 0;
    }
    return 0;
}

} // extern "C"

/* The sort hash grid of RendererRef.cpp:48-77: the macro root's box (for the
 * wide tree, as that code reads it: slot 0 of a leaf, else the last used slot). */
static void root_grid(SceneSpy &sc, float root_min[3], float cell_size[3]) {
    float root_max[3];
    if (!sc.mnodes_.empty()) {
        const Ray::mbvh_node_t &root = sc.mnodes_[sc.macro_nodes_root_];
        for (int i = 0; i < 3; i++) {
            root_min[i] = Ray::MAX_DIST;
            root_max[i] = -Ray::MAX_DIST;
        }
        const bool leaf = (root.child[0] & Ray::LEAF_NODE_BIT) != 0;
        for (int j = 0; j < (leaf ? 1 : 8); j++) {
            if (!leaf && root.child[j] == 0x7fffffff) continue;
            for (int i = 0; i < 3; i++) {
                root_min[i] = root.bbox_min[i][j];
                root_max[i] = root.bbox_max[i][j];
            }
        }
    } else {
        const Ray::bvh_node_t &root = sc.nodes_[sc.macro_nodes_root_];
        for (int i = 0; i < 3; i++) {
            root_min[i] = root.bbox_min[i];
            root_max[i] = root.bbox_max[i];
        }
    }
    for (int i = 0; i < 3; i++) cell_size[i] = (root_max[i] - root_min[i]) / 255;
}

extern "C" {

/* Ray hash keys of CoreRef.cpp:316-333 over the macro-tree root box. */
int ref_ray_hashes(rhx_scene *s, const rh_ray_packet *rays, size_t n, uint32_t *out) {
    SceneSpy &sc = *s->s;
    float root_min[3], cell_size[3];
    root_grid(sc, root_min, cell_size);
    for (size_t i = 0; i < n; i++) {
        Ray::Ref::ray_packet_t r;
        memcpy(&r, &rays[i], sizeof(r));
        out[i] = Ray::Ref::get_ray_hash(r, root_min, cell_size);
    }
    return 0;
}

/* SortRays_CPU (CoreRef.cpp:615-657): reorders rays in place. */
int ref_sort_rays(rhx_scene *s, rh_ray_packet *rays, size_t n) {
    SceneSpy &sc = *s->s;
    float root_min[3], cell_size[3];
    root_grid(sc, root_min, cell_size);
    std::vector<uint32_t> hashes(n), scan(n);
    std::vector<Ray::ray_chunk_t> chunks(n), chunks_tmp(n);
    if (!n) return 0;
    Ray::Ref::SortRays_CPU(reinterpret_cast<Ray::Ref::ray_packet_t *>(rays), n, root_min, cell_size, hashes.data(),
                           scan.data(), chunks.data(), chunks_tmp.data());
    return 0;
}

/* ShadeSurface (CoreRef.cpp:2055-2391) for a batch of (ray, hit) pairs. */
int ref_shade(rhx_scene *s, const rh_pass_info *pi_in, const rh_ray_packet *rays, const rh_hit *hits, size_t n,
              const float *halton, float *out_rgba, rh_ray_packet *out_secondary, int *out_secondary_count,
              int *out_secondary_src) {
    SceneSpy &sc = *s->s;
    Ray::scene_data_t d = scene_data(sc);
    int count = 0;
    for (size_t i = 0; i < n; i++) {
        Ray::pass_info_t pi;
        static_assert(sizeof(pi) == sizeof(rh_pass_info), "pass info layout");
        memcpy(&pi, &pi_in[i], sizeof(pi));
        Ray::Ref::ray_packet_t r;
        memcpy(&r, &rays[i], sizeof(r));
        Ray::Ref::hit_data_t inter;
        inter.mask_values[0] = hits[i].mask;
        inter.obj_indices[0] = hits[i].obj;
        inter.prim_indices[0] = hits[i].prim;
        inter.t = hits[i].t;
        inter.u = hits[i].u;
        inter.v = hits[i].v;
        inter.xy = hits[i].xy;
        int before = count;
        Ray::pixel_color_t col =
            Ray::Ref::ShadeSurface(pi, inter, r, halton, d, sc.macro_nodes_root_, sc.light_nodes_root_,
                                   sc.texture_atlas_, reinterpret_cast<Ray::Ref::ray_packet_t *>(out_secondary), &count);
        if (out_secondary_src && count > before) out_secondary_src[before] = (int)i;
        out_rgba[i * 4 + 0] = col.r;
        out_rgba[i * 4 + 1] = col.g;
        out_rgba[i * 4 + 2] = col.b;
        out_rgba[i * 4 + 3] = col.a;
    }
    *out_secondary_count = count;
    return 0;
}

/* Texel fetch through the reference atlas (TextureAtlasRef.h:60-65). */
uint32_t ref_atlas_get(rhx_scene *s, int page, int x, int y) {
    Ray::pixel_color8_t p = s->s->texture_atlas_.Get(page, x, y);
    return (uint32_t)p.r | ((uint32_t)p.g << 8) | ((uint32_t)p.b << 16) | ((uint32_t)p.a << 24);
}

/* SampleBilinear (CoreRef.cpp:1652-1678) through the reference atlas. */
int ref_sample_bilinear(rhx_scene *s, uint32_t tex, const float *uvs, int lod, size_t n, float *out) {
    SceneSpy &sc = *s->s;
    for (size_t i = 0; i < n; i++) {
        Ray::Ref::simd_fvec2 uv = {uvs[2 * i], uvs[2 * i + 1]};
        Ray::Ref::simd_fvec4 c = Ray::Ref::SampleBilinear(sc.texture_atlas_, sc.textures_[tex], uv, lod);
        for (int k = 0; k < 4; k++) out[4 * i + k] = c[k];
    }
    return 0;
}

} // extern "C"
