// Flat C ABI (include/rh_scene.h) over Ray::Hip::Renderer / Ray::Hip::Scene.
// Each function is a thin forwarder to the C++ API method named in the header.
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>

#include "Ray/RendererFactory.h"
#include "hip_backend.h"
#include "rh_scene.h"

namespace {
thread_local std::string g_err;

template <typename F> int guarded(F &&f) {
    try {
        return f();
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}
} // namespace

struct rhx_renderer {
    std::shared_ptr<Ray::Hip::Renderer> r;
};
struct rhx_scene {
    std::shared_ptr<Ray::Hip::Scene> s;
};
struct rhx_batch {
    std::unique_ptr<Ray::Hip::BatchRenderer> b;
};
struct rhx_region {
    Ray::RegionContext ctx;
    explicit rhx_region(const Ray::rect_t &rc) : ctx(rc) {}
};

extern "C" {

const char *rhx_last_error(void) { return g_err.c_str(); }

int rhx_create_renderer(const rhx_settings *s, uint32_t flags, rhx_renderer **out) {
    return guarded([&] {
        if (!(flags & Ray::RendererHIP)) {
            g_err = "this library provides only the HIP backend";
            return -2;
        }
        Ray::settings_t st;
        st.w = s->w;
        st.h = s->h;
        Ray::set_settings_device(st, s->device_index, 0);
        st.use_wide_bvh = s->use_wide_bvh != 0;
        std::ostringstream log;
        auto r = std::dynamic_pointer_cast<Ray::Hip::Renderer>(Ray::CreateRenderer(st, Ray::RendererHIP, log));
        r->set_sort_rays(s->sort_rays != 0);
        if (s->lds_stack_depth) rh_set_option(r->context(), "lds_stack_depth", s->lds_stack_depth);
        *out = new rhx_renderer{r};
        return 0;
    });
}

void rhx_destroy_renderer(rhx_renderer *r) { delete r; }
uint32_t rhx_renderer_type(const rhx_renderer *r) { return (uint32_t)r->r->type(); }

int rhx_resize(rhx_renderer *r, int w, int h) {
    return guarded([&] {
        r->r->Resize(w, h);
        return 0;
    });
}

int rhx_clear(rhx_renderer *r, const float rgba[4]) {
    r->r->Clear(Ray::pixel_color_t{rgba[0], rgba[1], rgba[2], rgba[3]});
    return 0;
}

int rhx_get_pixels(rhx_renderer *r, float *out, size_t count) {
    const auto sz = r->r->size();
    const size_t n = (size_t)sz.first * sz.second * 4;
    if (count < n) {
        g_err = "output buffer too small";
        return -1;
    }
    const Ray::pixel_color_t *p = r->r->get_pixels_ref();
    memcpy(out, p, n * sizeof(float));
    return 0;
}
int rhx_get_sh_data(rhx_renderer *r, float *out, size_t count) {
    const auto sz = r->r->size();
    const size_t n = (size_t)sz.first * sz.second * 12;
    if (count < n) {
        g_err = "output buffer too small";
        return -1;
    }
    const Ray::shl1_data_t *p = r->r->get_sh_data_ref();
    if (!p) return 1;
    memcpy(out, p, n * sizeof(float));
    return 0;
}

int rhx_get_stats(rhx_renderer *r, rh_stats *out) {
    Ray::RendererBase::stats_t st;
    r->r->GetStats(st);
    memcpy(out, &st, sizeof(st));
    return 0;
}

int rhx_reset_stats(rhx_renderer *r) {
    r->r->ResetStats();
    return 0;
}

int rhx_get_ray_counts(rhx_renderer *r, rh_ray_counts *out) {
    const int rc = rh_get_ray_counts(r->r->context(), out);
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_set_option(rhx_renderer *r, const char *name, int value) {
    if (!strcmp(name, "sort_rays")) {
        r->r->set_sort_rays(value != 0);
        return 0;
    }
    const int rc = rh_set_option(r->r->context(), name, value);
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_get_option(rhx_renderer *r, const char *name, int *value) {
    const int rc = rh_get_option(r->r->context(), name, value);
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_synchronize(rhx_renderer *r) {
    const int rc = rh_device_synchronize(r->r->context());
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_get_kernel_profile(rhx_renderer *r, rh_kernel_profile *out) {
    const int rc = rh_get_kernel_profile(r->r->context(), out);
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_reset_kernel_profile(rhx_renderer *r) {
    const int rc = rh_reset_kernel_profile(r->r->context());
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_create_scene(rhx_renderer *r, rhx_scene **out) {
    return guarded([&] {
        // a scene without a renderer is allowed: host-side preprocessing only
        *out = new rhx_scene{r ? std::dynamic_pointer_cast<Ray::Hip::Scene>(r->r->CreateScene())
                               : std::make_shared<Ray::Hip::Scene>()};
        return 0;
    });
}
int rhx_create_scene_standalone(int use_wide_bvh, rhx_scene **out) {
    return guarded([&] {
        *out = new rhx_scene{std::make_shared<Ray::Hip::Scene>(use_wide_bvh != 0)};
        return 0;
    });
}
void rhx_destroy_scene(rhx_scene *s) { delete s; }

namespace {
// one builder per device, kept for the process (the standalone build entry points)
rh_builder *process_builder(int device) {
    static std::mutex mu;
    static std::map<int, rh_builder *> builders;
    std::lock_guard<std::mutex> lock(mu);
    rh_builder *&b = builders[device];
    if (!b && rh_builder_create(device, &b) != 0) {
        b = nullptr;
        throw std::runtime_error(std::string("BVH build: ") + rh_last_error());
    }
    return b;
}
} // namespace

int rhx_build_sah(const float *boxes, size_t n, int device, float oversplit, float cost, rh_bvh_node *out_nodes,
                  size_t nodes_cap, uint32_t *out_count, uint32_t *out_indices, int *on_device) {
    return guarded([&] {
        *out_count = 0;
        *on_device = 0;
        if (nodes_cap < 2 * n) throw std::runtime_error("node output must hold 2 n nodes");
        std::vector<RayHip::prim_t> prims(n);
        for (size_t i = 0; i < n; i++) {
            prims[i].i0 = prims[i].i1 = prims[i].i2 = 0;
            memcpy(prims[i].bbox_min.v, boxes + 6 * i, 3 * sizeof(float));
            memcpy(prims[i].bbox_max.v, boxes + 6 * i + 3, 3 * sizeof(float));
        }
        RayHip::bvh_settings_t st;
        st.oversplit_threshold = oversplit;
        st.node_traversal_cost = cost;
        std::vector<rh_bvh_node> nodes;
        std::vector<uint32_t> indices;
        RayHip::SahBoxes fn;
        if (device >= 0) {
            rh_builder *b = process_builder(device);
            fn = [b, on_device](const float *bx, size_t m, const RayHip::bvh_settings_t &s,
                                std::vector<rh_bvh_node> &nd, std::vector<uint32_t> &ix) {
                nd.resize(2 * m);
                ix.resize(m);
                uint32_t count = 0;
                int host_needed = 0;
                if (rh_build_sah(b, bx, m, s.oversplit_threshold, s.node_traversal_cost, nd.data(), nd.size(), &count,
                                 ix.data(), &host_needed) != 0)
                    throw std::runtime_error(std::string("SAH build: ") + rh_last_error());
                if (host_needed) return false;
                nd.resize(count);
                *on_device = 1;
                return true;
            };
        }
        *out_count =
            RayHip::PreprocessPrims_SAH_Boxes(prims.data(), n, st, nodes, indices, device >= 0 ? &fn : nullptr);
        memcpy(out_nodes, nodes.data(), nodes.size() * sizeof(rh_bvh_node));
        memcpy(out_indices, indices.data(), indices.size() * sizeof(uint32_t));
        return 0;
    });
}

int rhx_build_fast_bvh(const float *boxes, size_t n, int device, rh_bvh_node *out_nodes, size_t nodes_cap,
                       uint32_t *out_count, uint32_t *out_indices) {
    return guarded([&] {
        *out_count = 0;
        if (nodes_cap < 2 * n) throw std::runtime_error("node output must hold 2 n nodes");
        std::vector<rh_bvh_node> nodes;
        std::vector<uint32_t> indices;
        if (device < 0) {
            std::vector<RayHip::prim_t> prims(n);
            for (size_t i = 0; i < n; i++) {
                prims[i].i0 = prims[i].i1 = prims[i].i2 = 0;
                memcpy(prims[i].bbox_min.v, boxes + 6 * i, 3 * sizeof(float));
                memcpy(prims[i].bbox_max.v, boxes + 6 * i + 3, 3 * sizeof(float));
            }
            *out_count = RayHip::PreprocessPrims_HLBVH(prims.data(), n, nodes, indices);
        } else {
            rh_builder *b = process_builder(device);
            // the boxes go to the device as they are (the bottom half reads
            // nothing else of the primitives); the top tree is built there too
            const RayHip::LbvhBottom fn = [b, boxes](const RayHip::prim_t *, size_t m, uint32_t offset,
                                                     std::vector<uint32_t> &order, std::vector<rh_bvh_node> &bottom,
                                                     std::vector<RayHip::HlbvhTreelet> &treelets) {
                Ray::Hip::DeviceLbvhBoxes(b, boxes, m, offset, order, bottom, treelets);
            };
            const RayHip::SahBoxes top = [b](const float *bx, size_t m, const RayHip::bvh_settings_t &s,
                                             std::vector<rh_bvh_node> &nd, std::vector<uint32_t> &ix) {
                nd.resize(2 * m);
                ix.resize(m);
                uint32_t count = 0;
                int host_needed = 0;
                if (rh_build_sah(b, bx, m, s.oversplit_threshold, s.node_traversal_cost, nd.data(), nd.size(), &count,
                                 ix.data(), &host_needed) != 0)
                    throw std::runtime_error(std::string("SAH build: ") + rh_last_error());
                if (host_needed) return false;
                nd.resize(count);
                return true;
            };
            *out_count = RayHip::PreprocessPrims_HLBVH(nullptr, n, nodes, indices, &fn, &top);
        }
        memcpy(out_nodes, nodes.data(), nodes.size() * sizeof(rh_bvh_node));
        memcpy(out_indices, indices.data(), indices.size() * sizeof(uint32_t));
        return 0;
    });
}

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
    if (!reg->ctx.halton_seq || count < RH_HALTON_COUNT * RH_HALTON_SEQ_LEN) return -1;
    memcpy(out, reg->ctx.halton_seq.get(), sizeof(float) * RH_HALTON_COUNT * RH_HALTON_SEQ_LEN);
    return 0;
}

int rhx_render_scene(rhx_renderer *r, rhx_scene *s, rhx_region *reg) {
    return guarded([&] {
        r->r->last_error.clear();
        r->r->RenderScene(s->s, reg->ctx);
        if (!r->r->last_error.empty()) {
            g_err = r->r->last_error;
            return -1;
        }
        return 0;
    });
}

int rhx_adjoint_scene(rhx_renderer *r, rhx_scene *s, rhx_region *reg, const float *d_adj_rgba, float *d_grad,
                      float *d_texel_grad, float scale) {
    return guarded([&] {
        r->r->last_error.clear();
        if (!r->r->AdjointScene(s->s, reg->ctx, d_adj_rgba, d_grad, d_texel_grad, scale)) {
            g_err = r->r->last_error.empty() ? "adjoint pass not run" : r->r->last_error;
            return -1;
        }
        return 0;
    });
}

int rhx_set_material_colors(rhx_renderer *r, rhx_scene *s, uint32_t first, uint32_t n, const float *rgb) {
    return guarded([&] {
        if (!r->r->SetMaterialColors(s->s, first, n, rgb)) {
            g_err = r->r->last_error;
            return -1;
        }
        return 0;
    });
}

int rhx_set_gradient_textures(rhx_renderer *r, const uint32_t *ids, uint32_t n) {
    const int rc = rh_set_gradient_textures(r->r->context(), ids, n);
    if (rc) g_err = rh_last_error();
    return rc;
}

int rhx_set_texture_data(rhx_renderer *r, rhx_scene *s, uint32_t tex, const uint8_t *rgba) {
    return guarded([&] {
        if (!r->r->SetTextureData(s->s, tex, rgba)) {
            g_err = r->r->last_error;
            return -1;
        }
        return 0;
    });
}

int rhx_read_accum(rhx_renderer *r, float *dst, size_t count) {
    return guarded([&] {
        if (!r->r->ReadAccumulated(dst, count)) {
            g_err = r->r->last_error;
            return -1;
        }
        return 0;
    });
}

int rhx_set_environment(rhx_scene *s, const float env_col[3], float env_clamp, uint32_t env_map) {
    return guarded([&] {
        Ray::environment_desc_t e;
        memcpy(e.env_col, env_col, 3 * sizeof(float));
        e.env_clamp = env_clamp;
        e.env_map = env_map;
        s->s->SetEnvironment(e);
        return 0;
    });
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
    try {
        return s->s->AddMaterial(d);
    } catch (const std::exception &e) {
        g_err = e.what(); // an index that names no texture or material: nothing is added
        return 0xffffffffu;
    }
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
    try {
        return s->s->AddMesh(d);
    } catch (const std::exception &e) {
        g_err = e.what(); // a device build that failed: the mesh is not added
        return 0xffffffffu;
    }
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
    memcpy(out, &s->s->camera(i), sizeof(rh_camera));
    return 0;
}
uint32_t rhx_triangle_count(rhx_scene *s) { return s->s->triangle_count(); }
uint32_t rhx_node_count(rhx_scene *s) { return s->s->node_count(); }

int rhx_scene_view(rhx_scene *s, rh_scene_view *out) {
    *out = s->s->View();
    return 0;
}


int rhx_batch_create(int device_index, rhx_batch **out) {
    return guarded([&] {
        *out = new rhx_batch{std::unique_ptr<Ray::Hip::BatchRenderer>(new Ray::Hip::BatchRenderer(device_index))};
        return 0;
    });
}

void rhx_batch_destroy(rhx_batch *b) { delete b; }

int rhx_batch_add(rhx_batch *b, rhx_scene *s, int w, int h, int spp) {
    return guarded([&] {
        const int i = b->b->AddScene(s->s, w, h, spp);
        if (i < 0) g_err = b->b->last_error;
        return i;
    });
}

int rhx_batch_passes(const rhx_batch *b) { return b->b->passes(); }

int rhx_batch_render_pass(rhx_batch *b) {
    return guarded([&] {
        if (!b->b->RenderPass()) {
            g_err = b->b->last_error;
            return -1;
        }
        return 0;
    });
}

int rhx_batch_get_pixels(rhx_batch *b, int scene, float *out, size_t count) {
    return guarded([&] {
        if (scene < 0 || scene >= b->b->scenes()) {
            g_err = "scene index out of range";
            return -1;
        }
        const auto sz = b->b->size(scene);
        const size_t n = (size_t)sz.first * sz.second * 4;
        if (count < n) {
            g_err = "output buffer too small";
            return -1;
        }
        const Ray::pixel_color_t *p = b->b->get_pixels_ref(scene);
        if (!p) {
            g_err = b->b->last_error;
            return -1;
        }
        memcpy(out, p, n * sizeof(float));
        return 0;
    });
}

int rhx_batch_set_option(rhx_batch *b, const char *name, int value) {
    const int rc = rh_batch_set_option(b->b->handle(), name, value);
    if (rc != 0) g_err = rh_last_error();
    return rc;
}

int rhx_batch_get_ray_counts(rhx_batch *b, rh_ray_counts *out) {
    const int rc = rh_batch_get_ray_counts(b->b->handle(), out);
    if (rc != 0) g_err = rh_last_error();
    return rc;
}

int rhx_batch_reset_stats(rhx_batch *b) { return rh_batch_reset_stats(b->b->handle()); }

int rhx_batch_read_queue(rhx_batch *b, int scene, int stage, int buffer, rh_ray *out, size_t cap, int *count) {
    const int rc = rh_batch_read_queue(b->b->handle(), scene, stage, buffer, out, cap, count);
    if (rc != 0) g_err = rh_last_error();
    return rc;
}

int rhx_batch_synchronize(rhx_batch *b) { return rh_batch_synchronize(b->b->handle()); }

int rhx_batch_restart(rhx_batch *b) {
    if (!b->b->Restart()) {
        g_err = b->b->last_error;
        return -1;
    }
    return 0;
}

} // extern "C"
