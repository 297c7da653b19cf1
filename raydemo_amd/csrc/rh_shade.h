// Surface shading with next-event estimation, restating the Ref backend's
// ShadeSurface and its helpers (src/Ray/internal/CoreRef.cpp:1801-2391) for one
// lane.  Visibility rays for the light samples reuse the traversal of
// rh_device.h in any-hit mode on the lane's own stack.
#pragma once

#include "rh_device.h"

namespace rh {

struct RayFull {
    f3 o, d, c;
    float ior;
    f3 do_dx, dd_dx, do_dy, dd_dy;
    int xy, depth;
};

struct ShadeCounters {
    uint32_t shadow_rays;
};

// ComputeVisibility (CoreRef.cpp:1801-1896): steps through transparent
// occluders, resolving mix materials with the lane's random offset; tc (when
// given) counts the traversal's node visits, triangle and instance tests
// Solid: the scene's triangles are all solid (Scene::all_solid), so the
// first triangle found decides and no transparent surface is stepped through
// (the same visibility; the kernel holds none of that path's registers).
template <int Tree = kTreeAny, bool Solid = false>
__device__ __forceinline__ float compute_visibility(const Scene &sc, f3 p1, f3 p2, const float *halton, int hi,
                                                    int rand_hash2, Stack &st, ShadeCounters &cnt,
                                                    TraceCounters *tc = nullptr) {
    f3 dir = p2 - p1;
    float dist = length(dir);
    dir = dir / dist;
    f3 ro = p1;
    float visibility = 1.0f;
    if constexpr (Solid) {
        if (dist > RH_HIT_EPS) {
            Hit sh;
            sh.mask = 0;
            sh.obj = -1;
            sh.prim = -1;
            sh.t = dist;
            sh.u = sh.v = 0.0f;
            cnt.shadow_rays++;
            if (sc.macro_root != RH_INVALID) traverse<true, Tree>(sc, ro, dir, sh, st, tc, true);
            if (sh.mask) visibility = 0;
        }
        return visibility;
    }

    while (dist > RH_HIT_EPS) {
        Hit sh;
        sh.mask = 0;
        sh.obj = -1;
        sh.prim = -1;
        sh.t = dist;
        sh.u = sh.v = 0.0f;
        cnt.shadow_rays++;
        // all-solid scene: the first blocking triangle found decides (visibility 0)
        if (sc.macro_root != RH_INVALID) traverse<true, Tree>(sc, ro, dir, sh, st, tc, sc.all_solid != 0);
        if (!sh.mask) break;
        if (sc.all_solid) {
            // every triangle is solid: the one found blocks (no load to see it)
            visibility = 0;
            break;
        }

        const rh_tri_accel tri = load_tri(sc.tris, (uint32_t)sh.prim);
        if (tri.ci & RH_TRI_SOLID_BIT) {
            visibility = 0;
            break;
        }
        const rh_material *mat = &sc.materials[tri.mi];
        const float *inv = sc.transforms[sc.mesh_instances[sh.obj].tr_index].inv_xform;
        const rh_vertex &v1 = sc.vertices[sc.vtx_indices[sh.prim * 3 + 0]];
        const rh_vertex &v2 = sc.vertices[sc.vtx_indices[sh.prim * 3 + 1]];
        const rh_vertex &v3 = sc.vertices[sc.vtx_indices[sh.prim * 3 + 2]];
        const f3 I = dir;
        const float w = 1.0f - sh.u - sh.v;
        const f2 sh_uvs = f2{v1.t[0][0] * w + v2.t[0][0] * sh.u + v3.t[0][0] * sh.v,
                             v1.t[0][1] * w + v2.t[0][1] * sh.u + v3.t[0][1] * sh.v};
        f3 sh_plane_N = transform_normal(extract_plane_normal(tri), inv);

        bool skip = false;
        const bool is_backfacing = dot(sh_plane_N, I) < 0.0f;
        if (is_backfacing) {
            if (tri.back_mi == RH_INVALID) {
                skip = true;
            } else {
                mat = &sc.materials[tri.back_mi];
                sh_plane_N = -sh_plane_N;
            }
        }
        if (!skip) {
            const int sh_rand_hash = hash(rand_hash2);
            const float sh_rand_offset = construct_float((uint32_t)sh_rand_hash);
            float sh_r = fract_modf(halton[hi] + sh_rand_offset);
            while (mat->type == RH_MAT_MIX) {
                const float mix_val = sample_bilinear(sc, mat->textures[RH_MAIN_TEXTURE], sh_uvs, 0).x * mat->roughness;
                if (sh_r > mix_val) {
                    mat = &sc.materials[mat->textures[RH_MIX_MAT1]];
                    sh_r = (sh_r - mix_val) / (1.0f - mix_val);
                } else {
                    mat = &sc.materials[mat->textures[RH_MIX_MAT2]];
                    sh_r = sh_r / mix_val;
                }
            }
            if (mat->type != RH_MAT_TRANSPARENT) {
                visibility = 0;
                break;
            }
        }
        const float t = sh.t + RH_HIT_BIAS;
        ro.x += dir.x * t;
        ro.y += dir.y * t;
        ro.z += dir.z * t;
        dist -= t;
    }
    return visibility;
}

// AcumulateLightContribution (CoreRef.cpp:1898-1933): one sample on a sphere light
template <int Tree = kTreeAny>
__device__ __forceinline__ void accumulate_light(const Scene &sc, const rh_light &l, f3 I, f3 P, f3 N, f3 B,
                                                 f3 plane_N, float sigma, const float *halton, int hi, int rand_hash2,
                                                 float rand_offset, float rand_offset2, f3 &col, Stack &st,
                                                 ShadeCounters &cnt) {
    const f3 lpos = ld3(l.pos);
    f3 L = P - lpos;
    const float distance = length(L);
    const float d = smax(distance - l.radius, 0.0f);
    L = L / distance;

    const float z = fract_modf(halton[hi + 0] + rand_offset);
    const float dir = __builtin_sqrtf(z);
    const float phi = 2 * RH_PI * fract_modf(halton[hi + 1] + rand_offset2);
    const float cos_phi = rcos(phi), sin_phi = rsin(phi);

    const f3 TT = cross(L, B);
    const f3 BB = cross(L, TT);
    const f3 V = (dir * sin_phi) * BB + __builtin_sqrtf(1.0f - dir) * L + (dir * cos_phi) * TT;

    L = normalize(lpos + V * l.radius - P);

    const float denom = d / l.radius + 1.0f;
    float atten = 1.0f / (denom * denom);
    atten = (atten - RH_LIGHT_ATTEN_CUTOFF / l.brightness) / (1.0f - RH_LIGHT_ATTEN_CUTOFF);
    atten = smax(atten, 0.0f);

    const float _dot1 = smax(dot(L, N), 0.0f);
    const float _dot2 = dot(L, ld3(l.dir));

    if (_dot1 > RH_FLT_EPS && _dot2 > l.spot && (l.brightness * atten) > RH_FLT_EPS) {
        const float visibility =
            compute_visibility<Tree>(sc, P + RH_HIT_BIAS * plane_N, lpos, halton, hi, rand_hash2, st, cnt);
        col = col + ((((ld3(l.col) * _dot1) * visibility) * atten) * brdf_oren_nayar(L, I, N, B, sigma));
    }
}

// ComputeDirectLighting (CoreRef.cpp:1935-1991): point-in-box walk of the light tree
template <int Tree = kTreeAny>
__device__ __forceinline__ f3 compute_direct_lighting(const Scene &sc, f3 I, f3 P, f3 N, f3 B, f3 plane_N,
                                                      float sigma, const float *halton, int hi, int rand_hash2,
                                                      float rand_offset, float rand_offset2, Stack &st,
                                                      ShadeCounters &cnt) {
    f3 col = f3{0.0f, 0.0f, 0.0f};
    // The reference walks the light tree depth first (right child popped
    // first) and enters a node only when P lies strictly inside its box.  A
    // parent box contains its children's boxes, so P strictly inside a leaf box
    // implies it is inside every ancestor: visiting the leaves in that
    // depth-first order and testing only the leaf boxes reaches the same lights
    // in the same order, without a per-lane stack.  The leaf list is built
    // once at scene upload.
    for (uint32_t k = 0; k < sc.light_leaf_count; k++) {
        const NodeBox n = load_node(sc.light_leaves, k);
        if (!sc.light_untested && !(P.x > n.bmin[0] && P.x < n.bmax[0] && P.y > n.bmin[1] && P.y < n.bmax[1] &&
                                    P.z > n.bmin[2] && P.z < n.bmax[2]))
            continue;
        const uint32_t first = n.a & RH_PRIM_INDEX_BITS;
        for (uint32_t i = first; i < first + n.b; i++) {
            const rh_light l = sc.lights[sc.li_indices[i]];
            accumulate_light<Tree>(sc, l, I, P, N, B, plane_N, sigma, halton, hi, rand_hash2, rand_offset, rand_offset2, col,
                             st, cnt);
        }
    }
    return col;
}

// Deferred next-event estimation.  ComputeVisibility only ever returns 0 or 1,
// and the reference adds each light's term ((l.col*dot1)*vis*atten)*brdf to
// the running colour in light order.  So shading can record, per light, the
// shadow-ray end points and the term with vis = 1 (x*1 == x exactly); a
// separate any-hit stage resolves each visibility, and a combine stage adds
// the visible terms in the same order: bit-identical to the inline form.
// Slots are per shaded ray: [ray * max_lights, ray * max_lights + count).
struct NeeSink {
    float4 *ray;  // p1.xyz, light index (bits)
    int2 *aux;    // halton index hi, rand_hash2
    float4 *term; // the light's term with visibility 1
    uint32_t base;
    uint32_t count;
};

__device__ __forceinline__ void defer_light(const Scene &sc, uint32_t light_index, const rh_light &l, f3 I, f3 P, f3 N,
                                            f3 B, f3 plane_N, float sigma, const float *halton, int hi, int rand_hash2,
                                            float rand_offset, float rand_offset2, NeeSink &sink) {
    const f3 lpos = ld3(l.pos);
    f3 L = P - lpos;
    const float distance = length(L);
    const float d = smax(distance - l.radius, 0.0f);
    L = L / distance;
    const float z = fract_modf(halton[hi + 0] + rand_offset);
    const float dir = __builtin_sqrtf(z);
    const float phi = 2 * RH_PI * fract_modf(halton[hi + 1] + rand_offset2);
    const float cos_phi = rcos(phi), sin_phi = rsin(phi);
    const f3 TT = cross(L, B);
    const f3 BB = cross(L, TT);
    const f3 V = (dir * sin_phi) * BB + __builtin_sqrtf(1.0f - dir) * L + (dir * cos_phi) * TT;
    L = normalize(lpos + V * l.radius - P);
    const float denom = d / l.radius + 1.0f;
    float atten = 1.0f / (denom * denom);
    atten = (atten - RH_LIGHT_ATTEN_CUTOFF / l.brightness) / (1.0f - RH_LIGHT_ATTEN_CUTOFF);
    atten = smax(atten, 0.0f);
    const float _dot1 = smax(dot(L, N), 0.0f);
    const float _dot2 = dot(L, ld3(l.dir));
    if (_dot1 > RH_FLT_EPS && _dot2 > l.spot && (l.brightness * atten) > RH_FLT_EPS) {
        const f3 p1 = P + RH_HIT_BIAS * plane_N;
        const f3 term = ((ld3(l.col) * _dot1) * atten) * brdf_oren_nayar(L, I, N, B, sigma);
        const uint32_t k = sink.base + sink.count++;
        sink.ray[k] = make_float4(p1.x, p1.y, p1.z, __uint_as_float(light_index));
        sink.aux[k] = make_int2(hi, rand_hash2);
        sink.term[k] = make_float4(term.x, term.y, term.z, 0.0f);
    }
}

__device__ __forceinline__ void defer_direct_lighting(const Scene &sc, f3 I, f3 P, f3 N, f3 B, f3 plane_N, float sigma,
                                                      const float *halton, int hi, int rand_hash2, float rand_offset,
                                                      float rand_offset2, NeeSink &sink) {
    for (uint32_t k = 0; k < sc.light_leaf_count; k++) {
        const NodeBox n = load_node(sc.light_leaves, k);
        if (!sc.light_untested && !(P.x > n.bmin[0] && P.x < n.bmax[0] && P.y > n.bmin[1] && P.y < n.bmax[1] &&
                                    P.z > n.bmin[2] && P.z < n.bmax[2]))
            continue;
        const uint32_t first = n.a & RH_PRIM_INDEX_BITS;
        for (uint32_t i = first; i < first + n.b; i++) {
            const uint32_t li = sc.li_indices[i];
            defer_light(sc, li, sc.lights[li], I, P, N, B, plane_N, sigma, halton, hi, rand_hash2, rand_offset,
                        rand_offset2, sink);
        }
    }
}

__device__ __forceinline__ float pow5(float v) { return (v * v) * (v * v) * v; }

struct ShadeOut {
    f4 col;
    bool emit;
    bool nee_pending; // diffuse hit whose direct light waits for the shadow stage
    bool consider_albedo;
    f3 albedo;
    RayFull next;
    // material whose main_color scales this vertex's colour and every later
    // contribution of the path exactly once (emissive; diffuse and glossy
    // when the albedo is considered), or -1
    int alb_mat;
    // the albedo texture sample of this vertex: texels and weights, the
    // sampled value before the sRGB conversion, and whether it was applied
    TexFoot alb_foot;
    f3 alb_raw;
    bool alb_srgb;
    uint32_t alb_tex; // the albedo texture
};

// ShadeSurface (CoreRef.cpp:2055-2391) for one (ray, hit) pair.
// With a sink, a diffuse hit's direct light is deferred (see NeeSink) and
// out.nee_pending is set; without one, visibility is traced inline.
template <int Tree = kTreeAny>
__device__ __forceinline__ void shade_surface(const Scene &sc, const rh_pass_info &pi, const Hit &inter,
                                              const RayFull &ray, const float *halton, Stack &st, ShadeCounters &cnt,
                                              ShadeOut &out, NeeSink *sink = nullptr) {
    out.emit = false;
    out.nee_pending = false;
    out.alb_mat = -1;
    const uint32_t flags = pi.settings.flags;
    const bool add_env = !(flags & RH_NO_BACKGROUND) || pi.bounce > 2;
    const bool add_direct = !(flags & RH_SKIP_DIRECT_LIGHT) || pi.bounce > 2;
    const bool consider_albedo = !(flags & RH_LIGHTING_ONLY) || pi.bounce > 2;
    const bool uniform_sampling = (flags & RH_OUTPUT_SH) && pi.bounce <= 2;

    if (!inter.mask) {
        f4 env_col = f4{0.0f, 0.0f, 0.0f, 0.0f};
        if (add_env) {
            env_col = sample_latlong_rgbe(sc, sc.env.env_map, ray.d);
            if (sc.env.env_clamp > RH_FLT_EPS) {
                const float c = sc.env.env_clamp;
                env_col = f4{env_col.x < c ? env_col.x : c, env_col.y < c ? env_col.y : c,
                             env_col.z < c ? env_col.z : c, env_col.w < c ? env_col.w : c};
            }
            env_col.w = 1.0f;
        }
        out.col = f4{ray.c.x * env_col.x * sc.env.env_col[0], ray.c.y * env_col.y * sc.env.env_col[1],
                     ray.c.z * env_col.z * sc.env.env_col[2], env_col.w};
        return;
    }

    const f3 I = ray.d;
    const f3 P = ray.o + inter.t * I;

    const rh_tri_accel tri = load_tri(sc.tris, (uint32_t)inter.prim);
    const rh_material *mat = &sc.materials[tri.mi];
    const float *inv = sc.transforms[sc.mesh_instances[inter.obj].tr_index].inv_xform;
    const rh_vertex &v1 = sc.vertices[sc.vtx_indices[inter.prim * 3 + 0]];
    const rh_vertex &v2 = sc.vertices[sc.vtx_indices[inter.prim * 3 + 1]];
    const rh_vertex &v3 = sc.vertices[sc.vtx_indices[inter.prim * 3 + 2]];

    const float w = 1.0f - inter.u - inter.v;
    f3 N = ld3(v1.n) * w + ld3(v2.n) * inter.u + ld3(v3.n) * inter.v;
    const f2 uvs = f2{v1.t[0][0] * w + v2.t[0][0] * inter.u + v3.t[0][0] * inter.v,
                      v1.t[0][1] * w + v2.t[0][1] * inter.u + v3.t[0][1] * inter.v};

    f3 plane_N = transform_normal(extract_plane_normal(tri), inv);

    const bool is_backfacing = dot(plane_N, I) > 0.0f;
    if (is_backfacing) {
        if (tri.back_mi == RH_INVALID) {
            out.col = f4{0.0f, 0.0f, 0.0f, 0.0f};
            return;
        }
        mat = &sc.materials[tri.back_mi];
        plane_N = -plane_N;
        N = -N;
    }

    Derivatives der;
    compute_derivatives(I, inter.t, ray.do_dx, ray.do_dy, ray.dd_dx, ray.dd_dy, v1, v2, v3, plane_N, der);

    const int rand_hash = hash(pi.rand_index), rand_hash2 = hash(rand_hash), rand_hash3 = hash(rand_hash2);
    const float rand_offset = construct_float((uint32_t)rand_hash), rand_offset2 = construct_float((uint32_t)rand_hash2),
                rand_offset3 = construct_float((uint32_t)rand_hash3);

    const int hi = (pi.iteration & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT + pi.bounce * 2;

    float mix_rand = fract_modf(halton[hi] + rand_offset);

    // mix materials: Schlick-weighted choice (CoreRef.cpp:2122-2158)
    while (mat->type == RH_MAT_MIX) {
        float mix_val = sample_bilinear(sc, mat->textures[RH_MAIN_TEXTURE], uvs, 0).x * mat->roughness;
        const float mix_ior = is_backfacing ? mat->ext_ior : mat->int_ior;
        float R0 = (ray.ior - mix_ior) / (ray.ior + mix_ior);
        R0 *= R0;
        float RR;
        if (ray.ior > mix_ior) {
            const float eta = ray.ior / mix_ior;
            const float cosi = -dot(I, N);
            const float cost2 = 1.0f - eta * eta * (1.0f - cosi * cosi);
            if (cost2 >= 0.0f) {
                const float m = eta * cosi - __builtin_sqrtf(cost2);
                const f3 V = eta * I + m * N;
                RR = R0 + (1.0f - R0) * pow5(1.0f + dot(V, N));
            } else {
                RR = 1.0f;
            }
        } else {
            RR = R0 + (1.0f - R0) * pow5(1.0f + dot(I, N));
        }
        mix_val *= clampf(RR, 0.0f, 1.0f);
        if (mix_rand > mix_val) {
            mat = &sc.materials[mat->textures[RH_MIX_MAT1]];
            mix_rand = (mix_rand - mix_val) / (1.0f - mix_val);
        } else {
            mat = &sc.materials[mat->textures[RH_MIX_MAT2]];
            mix_rand = mix_rand / mix_val;
        }
    }

    // normal map (CoreRef.cpp:2161-2170)
    f3 B = ld3(v1.b) * w + ld3(v2.b) * inter.u + ld3(v3.b) * inter.v;
    f3 T = cross(B, N);
    f4 nm = sample_bilinear(sc, mat->textures[RH_NORMALS_TEXTURE], uvs, 0);
    nm = f4{nm.x * 2.0f - 1.0f, nm.y * 2.0f - 1.0f, nm.z * 2.0f - 1.0f, nm.w * 2.0f - 1.0f};
    N = nm.x * B + nm.z * N + nm.y * T;
    N = transform_normal(N, inv);
    B = transform_normal(B, inv);
    T = transform_normal(T, inv);

    // albedo (CoreRef.cpp:2172-2181)
    const uint32_t main_tex = mat->textures[RH_MAIN_TEXTURE];
    const rh_texture &mt = sc.textures[main_tex];
    const float albedo_lod = get_texture_lod(mt, der.duv_dx, der.duv_dy);
    f4 albedo = sample_bilinear(sc, main_tex, uvs, (int)albedo_lod, &out.alb_foot);
    out.alb_raw = f3{albedo.x, albedo.y, albedo.z};
    out.alb_srgb = (mt.width & RH_TEXTURE_SRGB_BIT) != 0;
    out.alb_tex = main_tex;
    if (mt.width & RH_TEXTURE_SRGB_BIT) albedo = srgb_to_rgb(albedo);
    albedo.x *= mat->main_color[0];
    albedo.y *= mat->main_color[1];
    albedo.z *= mat->main_color[2];
    const f3 alb = f3{albedo.x, albedo.y, albedo.z};

    f3 col = f3{0.0f, 0.0f, 0.0f};
    if (mat->type == RH_MAT_EMISSIVE || (consider_albedo && (mat->type == RH_MAT_DIFFUSE || mat->type == RH_MAT_GLOSSY)))
        out.alb_mat = (int)(mat - sc.materials);

    const int diff_depth = ray.depth & 0xff;
    const int gloss_depth = (ray.depth >> 8) & 0xff;
    const int refr_depth = (ray.depth >> 16) & 0xff;
    const int transp_depth = (ray.depth >> 24) & 0xff;
    const int total_depth = diff_depth + gloss_depth + refr_depth + transp_depth;
    const bool cant_terminate = total_depth < pi.settings.termination_start_depth;

    RayFull r;
    bool emit = false;

    if (mat->type == RH_MAT_DIFFUSE) {
        if (add_direct) {
            if (sink) {
                defer_direct_lighting(sc, I, P, N, B, plane_N, mat->roughness, halton, hi, rand_hash2, rand_offset,
                                      rand_offset2, *sink);
                out.nee_pending = true;
                out.consider_albedo = consider_albedo;
                out.albedo = alb;
            } else {
                col = compute_direct_lighting<Tree>(sc, I, P, N, B, plane_N, mat->roughness, halton, hi, rand_hash2,
                                              rand_offset, rand_offset2, st, cnt);
                if (consider_albedo) col = col * alb;
            }
        }
        if (diff_depth < pi.settings.max_diff_depth && total_depth < pi.settings.max_total_depth) {
            const float u1 = fract_modf(halton[hi + 0] + rand_offset);
            const float u2 = fract_modf(halton[hi + 1] + rand_offset2);
            const float phi = 2 * RH_PI * u2;
            const float cos_phi = rcos(phi), sin_phi = rsin(phi);
            f3 V;
            float weight = 1.0f;
            if (uniform_sampling) {
                const float dir = __builtin_sqrtf(1.0f - u1 * u1);
                V = (dir * sin_phi) * B + u1 * N + (dir * cos_phi) * T;
                weight = 2 * u1;
            } else {
                const float dir = __builtin_sqrtf(u1);
                V = (dir * sin_phi) * B + __builtin_sqrtf(1.0f - u1) * N + (dir * cos_phi) * T;
            }
            weight *= brdf_oren_nayar(V, I, N, B, mat->roughness);

            r.xy = ray.xy;
            r.ior = ray.ior;
            r.depth = ray.depth + 0x00000001;
            r.o = P + RH_HIT_BIAS * plane_N;
            r.d = V;
            r.c = f3{ray.c.x * weight, ray.c.y * weight, ray.c.z * weight};
            if (consider_albedo) r.c = r.c * alb;
            r.do_dx = der.do_dx;
            r.do_dy = der.do_dy;
            const float dIN = dot(I, plane_N);
            r.dd_dx = der.dd_dx - 2 * (dIN * der.dndx + der.ddn_dx * plane_N);
            r.dd_dy = der.dd_dy - 2 * (dIN * der.dndy + der.ddn_dy * plane_N);
            emit = true;
        }
    } else if (mat->type == RH_MAT_GLOSSY) {
        if (gloss_depth < pi.settings.max_glossy_depth && total_depth < pi.settings.max_total_depth) {
            f3 V = I - (2 * dot(N, I)) * N;
            const float h = 1.0f - rcos(0.5f * RH_PI * mat->roughness * mat->roughness);
            const float z = h * fract_modf(halton[hi + 0] + rand_offset);
            const float dir = __builtin_sqrtf(z);
            const float phi = 2 * RH_PI * fract_modf(halton[hi + 1] + rand_offset2);
            const float cos_phi = rcos(phi), sin_phi = rsin(phi);
            const f3 TT = cross(V, B);
            const f3 BB = cross(V, TT);
            if (dot(V, plane_N) > 0) {
                V = (dir * sin_phi) * BB + __builtin_sqrtf(1.0f - dir) * V + (dir * cos_phi) * TT;
            } else {
                V = ((-dir) * sin_phi) * BB + __builtin_sqrtf(1.0f - dir) * V - (dir * cos_phi) * TT;
            }
            r.xy = ray.xy;
            r.ior = ray.ior;
            r.depth = ray.depth + 0x00000100;
            r.o = P + RH_HIT_BIAS * plane_N;
            r.d = V;
            r.c = ray.c;
            if (consider_albedo) r.c = r.c * alb;
            r.do_dx = der.do_dx;
            r.do_dy = der.do_dy;
            const float dIN = dot(I, plane_N);
            r.dd_dx = der.dd_dx - 2 * (dIN * der.dndx + der.ddn_dx * plane_N);
            r.dd_dy = der.dd_dy - 2 * (dIN * der.dndy + der.ddn_dy * plane_N);
            emit = true;
        }
    } else if (mat->type == RH_MAT_REFRACTIVE) {
        if (refr_depth < pi.settings.max_refr_depth && total_depth < pi.settings.max_total_depth) {
            const float ior = is_backfacing ? mat->ext_ior : mat->int_ior;
            const float eta = ray.ior / ior;
            const float cosi = -dot(I, N);
            const float cost2 = 1.0f - eta * eta * (1.0f - cosi * cosi);
            if (cost2 < 0) {
                out.col = f4{0.0f, 0.0f, 0.0f, 1.0f};
                return;
            }
            const float m = eta * cosi - __builtin_sqrtf(cost2);
            f3 V = eta * I + m * N;
            const float z = 1.0f - halton[hi + 0] * mat->roughness;
            const float temp = __builtin_sqrtf(1.0f - z * z);
            const float phi =
                halton[(((hash(hi) + pi.iteration) & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT + pi.bounce * 2) + 0] *
                2 * RH_PI;
            const float cos_phi = rcos(phi), sin_phi = rsin(phi);
            const f3 TT = normalize(cross(V, B));
            const f3 BB = normalize(cross(V, TT));
            V = (temp * sin_phi) * BB + z * V + (temp * cos_phi) * TT;

            const float k = (eta - eta * eta * dot(I, plane_N) / dot(V, plane_N));
            const float dmdx = k * der.ddn_dx, dmdy = k * der.ddn_dy;
            r.xy = ray.xy;
            r.ior = ior;
            r.depth = ray.depth + 0x00010000;
            r.o = P + RH_HIT_BIAS * I;
            r.d = V;
            r.c = ray.c * z;
            r.do_dx = der.do_dx;
            r.do_dy = der.do_dy;
            r.dd_dx = eta * der.dd_dx - (m * der.dndx + dmdx * plane_N);
            r.dd_dy = eta * der.dd_dy - (m * der.dndy + dmdy * plane_N);
            emit = true;
        }
    } else if (mat->type == RH_MAT_EMISSIVE) {
        col = mat->roughness * alb;
    } else if (mat->type == RH_MAT_TRANSPARENT) {
        if (transp_depth < pi.settings.max_transp_depth && total_depth < pi.settings.max_total_depth) {
            r.xy = ray.xy;
            r.ior = ray.ior;
            r.depth = ray.depth + 0x01000000;
            r.o = P + RH_HIT_BIAS * I;
            r.d = ray.d;
            r.c = ray.c;
            r.do_dx = ray.do_dx;
            r.do_dy = ray.do_dy;
            r.dd_dx = ray.dd_dx;
            r.dd_dy = ray.dd_dy;
            emit = true;
        }
    }

    if (emit) {
        // Russian roulette (CoreRef.cpp:2248-2255)
        const float lum = smax(r.c.x, smax(r.c.y, r.c.z));
        const float p = fract_modf(halton[hi + 0] + rand_offset3);
        const float q = cant_terminate ? 0.0f : smax(0.05f, 1.0f - lum);
        if (p >= q) {
            r.c.x /= 1.0f - q;
            r.c.y /= 1.0f - q;
            r.c.z /= 1.0f - q;
            out.emit = true;
            out.next = r;
        }
    }
    out.col = f4{ray.c.x * col.x, ray.c.y * col.y, ray.c.z * col.z, 1.0f};
}

// The direct light a deferred sink recorded for one shaded ray, resolved in
// the same lane: visibility of each sample (ComputeVisibility), the visible
// terms summed in light order, times the albedo when considered, times the
// ray's throughput -- the arithmetic of the separate shadow and combine
// stages, so adding the result after out.col gives the same pixel value.
template <int Tree = kTreeAny, bool Solid = false>
__device__ __forceinline__ f3 resolve_light_samples(const Scene &sc, const float *halton, const NeeSink &sink,
                                                    const ShadeOut &out, f3 throughput, Stack &st,
                                                    ShadeCounters &cnt, TraceCounters *tc = nullptr) {
    f3 col = f3{0.0f, 0.0f, 0.0f};
    for (uint32_t j = 0; j < sink.count; j++) {
        const uint32_t k = sink.base + j;
        const float4 r = sink.ray[k];
        const int2 x = sink.aux[k];
        const rh_light &l = sc.lights[__float_as_uint(r.w)];
        st.sp = 0;
        if (compute_visibility<Tree, Solid>(sc, f3{r.x, r.y, r.z}, ld3(l.pos), halton, x.x, x.y, st, cnt, tc) != 0.0f) {
            const float4 t = sink.term[k];
            col = col + f3{t.x, t.y, t.z};
        }
    }
    if (out.consider_albedo) col = col * out.albedo;
    return f3{throughput.x * col.x, throughput.y * col.y, throughput.z * col.z};
}

} // namespace rh
