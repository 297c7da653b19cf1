/*
 * rh_types.h — plain-old-data layouts shared by the C++ host layer, the HIP
 * kernels and the CPU oracle.  Every scene record keeps the byte layout of the
 * Ray library's own records so that a scene built by the reference and one
 * built here can be compared array by array, and so that the GPU walks the
 * same BVH node layout the reference's CPU backend walks.
 *
 * Reference layouts restated (src/Ray/internal/Core.h):
 *   rh_tri_accel     <- tri_accel_t      Core.h:27-36   (48 B)
 *   rh_bvh_node      <- bvh_node_t       Core.h:65-85   (32 B)
 *   rh_texture       <- texture_t        Core.h:103-109 (74 B)
 *   rh_material      <- material_t       Core.h:121-131 (48 B)
 *   rh_light         <- light_t          Core.h:133-138 (48 B)
 *   rh_vertex        <- vertex_t         Core.h:219-222 (52 B)
 *   rh_mesh          <- mesh_t           Core.h:224-228 (16 B)
 *   rh_transform     <- transform_t      Core.h:230-233 (128 B)
 *   rh_mesh_instance <- mesh_instance_t  Core.h:235-241 (32 B)
 *   rh_environment   <- environment_t    Core.h:243-247
 *   rh_pass_settings <- pass_settings_t  Types.h:48-57
 *   rh_camera        <- camera_t         Types.h:59-71
 *
 * The wavefront records (rh_ray, rh_hit) are this backend's own: the ray is
 * split into a 32-byte traversal half and a 64-byte shading half so that the
 * trace kernels read only what traversal needs.
 */
#ifndef RH_TYPES_H
#define RH_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (Core.h:40-63, 94-119, 207-217) ---- */
#define RH_TRI_W_BITS 0x03u
#define RH_TRI_AXIS_ALIGNED_BIT 0x04u
#define RH_TRI_INV_NORMAL_BIT 0x08u
#define RH_TRI_SOLID_BIT 0x10u

#define RH_HIT_BIAS 0.001f
#define RH_HIT_EPS 0.000001f
#define RH_FLT_EPS 0.0000001f
#define RH_PI 3.141592653589793238463f
#define RH_MAX_DIST 3.402823466e+38F
#define RH_MAX_BOUNCES 10     /* MAX_BOUNCES (Core.h): the reference clamps max_total_depth to it */
#define RH_MAX_BOUNCES_EXT 16 /* the largest depth the "max_bounces" option allows (extension) */
#define RH_LIGHT_ATTEN_CUTOFF 0.001f

#define RH_LEAF_NODE_BIT 0x80000000u
#define RH_PRIM_INDEX_BITS 0x7fffffffu
#define RH_LEFT_CHILD_BITS 0x7fffffffu
#define RH_SEP_AXIS_BITS 0xc0000000u
#define RH_PRIM_COUNT_BITS 0x3fffffffu
#define RH_RIGHT_CHILD_BITS 0x3fffffffu

#define RH_NUM_MIP_LEVELS 14
#define RH_MAX_MIP_LEVEL (RH_NUM_MIP_LEVELS - 1)
#define RH_TEXTURE_ATLAS_SIZE 8192
#define RH_TEXTURE_SRGB_BIT 0x8000u
#define RH_TEXTURE_WIDTH_BITS 0x7fffu

#define RH_NORMALS_TEXTURE 0
#define RH_MAIN_TEXTURE 1
#define RH_MIX_MAT1 2
#define RH_MIX_MAT2 3

#define RH_MAX_STACK_SIZE 32

#define RH_HALTON_COUNT 24
#define RH_HALTON_2D_COUNT 12
#define RH_HALTON_SEQ_LEN 256
/* Device (and oracle) Halton tables hold one more row, a copy of row 0.  A
   pass reads dimensions 2*bounce and 2*bounce+1 of its row: bounce index 12
   (the last bounce at depth 10, CoreRef.cpp ShadeSurface) and the bounces of
   the max_bounces extension run past the row's 24 entries, into the next row,
   and on the last row of an epoch past the 256-row table; the extra row keeps
   those reads in bounds and defined (the wrap to row 0). */
#define RH_HALTON_TABLE_FLOATS (RH_HALTON_COUNT * (RH_HALTON_SEQ_LEN + 1))

#define RH_INVALID 0xffffffffu

/* material types (SceneBase.h:45-52) */
enum { RH_MAT_DIFFUSE = 0, RH_MAT_GLOSSY, RH_MAT_REFRACTIVE, RH_MAT_EMISSIVE, RH_MAT_MIX, RH_MAT_TRANSPARENT };
/* light types (SceneBase.h:106-110) */
enum { RH_LIGHT_POINT = 0, RH_LIGHT_SPOT, RH_LIGHT_DIRECTIONAL };
/* camera (Types.h:34-38) */
enum { RH_CAM_PERSP = 0, RH_CAM_ORTHO, RH_CAM_GEO };
enum { RH_FILTER_BOX = 0, RH_FILTER_TENT };
enum { RH_DEVICE_NONE = 0, RH_DEVICE_SRGB };
/* pass flags (Types.h:40-46) */
enum {
    RH_SKIP_DIRECT_LIGHT = 1 << 0,
    RH_SKIP_INDIRECT_LIGHT = 1 << 1,
    RH_LIGHTING_ONLY = 1 << 2,
    RH_NO_BACKGROUND = 1 << 3,
    RH_CLAMP = 1 << 4,
    RH_OUTPUT_SH = 1 << 5,
    RH_USE_COHERENT_SAMPLING = 1 << 6
};
/* vertex layouts (SceneBase.h:27-42) */
enum { RH_PxyzNxyzTuv = 0, RH_PxyzNxyzTuvTuv, RH_PxyzNxyzBxyzTuv, RH_PxyzNxyzBxyzTuvTuv };

/* ---- scene records ---- */
typedef struct rh_tri_accel {
    float nu, nv;
    float np;
    float pu, pv;
    int32_t ci;
    float e0u, e0v;
    float e1u, e1v;
    uint32_t mi, back_mi;
} rh_tri_accel;

typedef struct rh_bvh_node {
    float bbox_min[3];
    uint32_t prim_index; /* == left_child for interior nodes; bit 31 marks a leaf */
    float bbox_max[3];
    uint32_t prim_count; /* == right_child for interior nodes; bits 30-31 hold the split axis */
} rh_bvh_node;

/* mbvh_node_t (Core.h:87-92): the 8-wide node of the reference's default
 * tree (settings use_wide_bvh).  Interior: child[i] is a node index, or
 * RH_MBVH_EMPTY with NaN bounds for an unused slot; leaf: child[0] = prim
 * index | RH_LEAF_NODE_BIT, child[1] = prim count, bounds in slot 0. */
#define RH_MBVH_EMPTY 0x7fffffffu
typedef struct rh_mbvh_node {
    float bbox_min[3][8];
    float bbox_max[3][8];
    uint32_t child[8];
} rh_mbvh_node;

typedef struct rh_texture {
    uint16_t width; /* bit 15 = sRGB flag */
    uint16_t height;
    uint8_t page[RH_NUM_MIP_LEVELS];
    uint16_t pos[RH_NUM_MIP_LEVELS][2];
} rh_texture;

typedef struct rh_material {
    uint32_t textures[4];
    float main_color[3], pad;
    uint32_t type;
    float roughness; /* == strength for emissive and mix materials */
    float int_ior, ext_ior;
} rh_material;

typedef struct rh_light {
    float pos[3], radius;
    float col[3], brightness;
    float dir[3], spot;
} rh_light;

typedef struct rh_vertex {
    float p[3], n[3], b[3], t[2][2];
} rh_vertex;

typedef struct rh_mesh {
    uint32_t node_index, node_count;
    uint32_t tris_index, tris_count;
} rh_mesh;

typedef struct rh_transform {
    float xform[16], inv_xform[16];
} rh_transform;

typedef struct rh_mesh_instance {
    float bbox_min[3];
    uint32_t tr_index;
    float bbox_max[3];
    uint32_t mesh_index;
} rh_mesh_instance;

typedef struct rh_environment {
    float env_col[3];
    float env_clamp;
    uint32_t env_map;
} rh_environment;

typedef struct rh_pass_settings {
    uint8_t max_diff_depth, max_glossy_depth, max_refr_depth, max_transp_depth, max_total_depth;
    uint8_t termination_start_depth;
    uint8_t pad[2];
    uint32_t flags;
} rh_pass_settings;

typedef struct rh_camera {
    int32_t type;
    int32_t filter;
    int32_t dtype;
    float fov, gamma;
    float focus_distance, focus_factor;
    float origin[3], fwd[3], side[3], up[3];
    uint32_t mi_index, uv_index;
    rh_pass_settings pass_settings;
} rh_camera;

/*
 * Texel storage.  The reference keeps every mip level in an 8192x8192 tiled
 * RGBA8 atlas page with a one-texel wrap border (TextureAtlasRef.cpp:192-231).
 * Here each mip level is stored compactly as its own (w+2) x (h+2) bordered
 * block inside one texel pool; rh_texture keeps the reference's atlas page and
 * position so that the sampling arithmetic (and its rounding) is unchanged, and
 * rh_tex_mip maps (texture, mip) to the compact block.
 */
typedef struct rh_tex_mip {
    uint32_t offset; /* first texel of the bordered block in the texel pool */
    uint16_t bw, bh; /* bordered width/height = mip size + 2 */
} rh_tex_mip;

/* A read-only view of a whole scene: what the C ABI uploads and what the CPU
 * oracle walks.  Pointers are host pointers. */
typedef struct rh_scene_view {
    const rh_environment *env;
    const rh_mesh_instance *mesh_instances;
    uint32_t mesh_instances_count;
    const uint32_t *mi_indices;
    uint32_t mi_indices_count;
    const rh_mesh *meshes;
    uint32_t meshes_count;
    const rh_transform *transforms;
    uint32_t transforms_count;
    const uint32_t *vtx_indices;
    uint32_t vtx_indices_count;
    const rh_vertex *vertices;
    uint32_t vertices_count;
    const rh_bvh_node *nodes;
    uint32_t nodes_count;
    const rh_tri_accel *tris;
    uint32_t tris_count;
    const uint32_t *tri_indices;
    uint32_t tri_indices_count;
    const rh_material *materials;
    uint32_t materials_count;
    const rh_texture *textures;
    uint32_t textures_count;
    const rh_tex_mip *tex_mips; /* textures_count * RH_NUM_MIP_LEVELS */
    const uint32_t *texels;     /* RGBA8 packed, little-endian r in bits 0-7 */
    uint32_t texels_count;
    const rh_light *lights;
    uint32_t lights_count;
    const uint32_t *li_indices;
    uint32_t li_indices_count;
    uint32_t macro_nodes_root;
    uint32_t light_nodes_root;
    /* 8-wide trees (use_wide_bvh): when mnodes is set, nodes is empty and
     * meshes[].node_index, macro_nodes_root and light_nodes_root index mnodes */
    const rh_mbvh_node *mnodes;
    uint32_t mnodes_count;
} rh_scene_view;

/* ---- wavefront records (this backend's own) ---- */

/* Traversal half of a ray: origin + pixel xy, direction + packed depth
 * counters.  xy = (x << 16) | y as in the reference (CoreRef.h:41-44). */
typedef struct rh_ray {
    float o[3];
    int32_t xy;
    float d[3];
    int32_t depth; /* four 8-bit counters: diffuse | glossy<<8 | refr<<16 | transp<<24 */
} rh_ray;

/* Shading half of a ray: throughput, ior and the ray differentials. */
typedef struct rh_ray_aux {
    float c[3], ior;
    float do_dx[3], dd_dx[3], do_dy[3], dd_dy[3];
} rh_ray_aux;

/* Closest/any hit.  mask != 0 marks a hit (CoreRef.h:51-60). */
typedef struct rh_hit {
    int32_t mask;
    int32_t obj;  /* mesh instance index, -1 when no hit */
    int32_t prim; /* triangle index, -1 when no hit */
    float t, u, v;
    int32_t xy;
    int32_t pad;
} rh_hit;

/* A full reference-layout ray (CoreRef.h:34-45), used at the test boundary. */
typedef struct rh_ray_packet {
    float o[3], d[3];
    float c[3], ior;
    float do_dx[3], dd_dx[3], do_dy[3], dd_dy[3];
    int32_t xy;
    int32_t ray_depth;
} rh_ray_packet;

/* per-pass shading state (Core.h:269-293) */
typedef struct rh_pass_info {
    int32_t index, rand_index;
    int32_t iteration, bounce;
    rh_pass_settings settings;
} rh_pass_info;

/* stage timers (RendererBase.h:176-183) */
typedef struct rh_stats {
    unsigned long long time_primary_ray_gen_us;
    unsigned long long time_primary_trace_us;
    unsigned long long time_primary_shade_us;
    unsigned long long time_secondary_sort_us;
    unsigned long long time_secondary_trace_us;
    unsigned long long time_secondary_shade_us;
} rh_stats;

/* rays traced, counted on the device */
typedef struct rh_ray_counts {
    unsigned long long primary;
    unsigned long long secondary;
    unsigned long long shadow;
    unsigned long long sorted;
    unsigned long long adjoint;        /* path segments traced by the adjoint replay */
    unsigned long long adjoint_shadow; /* visibility segments of the adjoint replay */
    unsigned long long overflow;       /* traversal stack pushes dropped past the per-lane stack's capacity
                                          (LDS part + global spill); non-zero means hits may have been missed */
} rh_ray_counts;

/* per-kernel profile: kinds 0 trace, 1 shade, 2 sort, 3 shadow (deferred NEE visibility),
 * 4 path tail (the last bounces traced and shaded in one launch) */
#define RH_KERNEL_KINDS 5
/* stages of a pass's closest-hit trace launches: 0 the primary rays, b + 1 bounce b */
#define RH_TRACE_STAGES (RH_MAX_BOUNCES_EXT + 1)
typedef struct rh_kernel_profile {
    double ms[RH_KERNEL_KINDS];
    unsigned long long launches[RH_KERNEL_KINDS];
    unsigned long long node_visits, tri_tests, instance_tests;
    unsigned long long counted_rays; /* rays traced by the counting trace kernels */
    /* the counting shadow kernels: node visits, triangle tests, instance tests, queue entries, segments */
    unsigned long long shadow_visits[5];
    /* the counting path-tail kernels: node visits, triangle tests, instance tests (path and shadow
     * traversals together), path segments, shadow segments */
    unsigned long long tail_visits[5];
    /* the closest-hit trace launches of rh_render_pass by stage: time and launches (with kernel_timing),
     * node visits, triangle tests, instance tests and rays (with count_visits; also in the totals above) */
    double trace_stage_ms[RH_TRACE_STAGES];
    unsigned long long trace_stage_launches[RH_TRACE_STAGES];
    unsigned long long trace_stage_visits[RH_TRACE_STAGES][4];
} rh_kernel_profile;

#ifdef __cplusplus
}

static_assert(sizeof(rh_tri_accel) == 48, "tri_accel_t is 48 bytes");
static_assert(sizeof(rh_bvh_node) == 32, "bvh_node_t is 32 bytes");
static_assert(sizeof(rh_mbvh_node) == 224, "mbvh_node_t is 224 bytes");
static_assert(sizeof(rh_texture) == 74, "texture_t is 74 bytes");
static_assert(sizeof(rh_material) == 48, "material_t is 48 bytes");
static_assert(sizeof(rh_light) == 48, "light_t is 48 bytes");
static_assert(sizeof(rh_vertex) == 52, "vertex_t is 52 bytes");
static_assert(sizeof(rh_mesh) == 16, "mesh_t is 16 bytes");
static_assert(sizeof(rh_transform) == 128, "transform_t is 128 bytes");
static_assert(sizeof(rh_mesh_instance) == 32, "mesh_instance_t is 32 bytes");
static_assert(sizeof(rh_pass_settings) == 12, "pass_settings_t is 12 bytes");
static_assert(sizeof(rh_pass_info) == 28, "pass_info_t is 28 bytes");
static_assert(sizeof(rh_ray) == 32, "traversal half of a ray");
static_assert(sizeof(rh_ray_aux) == 64, "shading half of a ray");
static_assert(sizeof(rh_hit) == 32, "hit record");
static_assert(sizeof(rh_ray_packet) == 96, "reference ray_packet_t is 96 bytes");
#endif

#endif /* RH_TYPES_H */
