// Host-side scene preprocessing for the HIP backend: the parts of the Ray
// library's core that run on the CPU at scene-build time.  Every function
// restates a reference routine (cited beside it) so that a scene built here
// has the same triangle records, BVH node arrays and index orders as one
// built by the reference's CPU backends.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include "rh_types.h"

namespace RayHip {

struct vec3 {
    float v[3];
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

// component-wise min/max with the argument order of std::min/std::max
inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
inline float fmax_std(float a, float b) { return (a < b) ? b : a; }
inline vec3 vmin(const vec3 &a, const vec3 &b) { return {{fmin_std(a[0], b[0]), fmin_std(a[1], b[1]), fmin_std(a[2], b[2])}}; }
inline vec3 vmax(const vec3 &a, const vec3 &b) { return {{fmax_std(a[0], b[0]), fmax_std(a[1], b[1]), fmax_std(a[2], b[2])}}; }

// prim_t (BVHSplit.h:9-12)
struct prim_t {
    uint32_t i0, i1, i2;
    vec3 bbox_min, bbox_max;
};

// bvh_settings_t (Core.h:142-147)
struct bvh_settings_t {
    float oversplit_threshold = 0.95f;
    float node_traversal_cost = 0.025f;
    bool allow_spatial_splits = false;
    bool use_fast_bvh_build = false;
};

// PreprocessTri (Core.cpp:162-214)
bool PreprocessTri(const float *p, int stride, rh_tri_accel *acc);
// ExtractPlaneNormal (Core.cpp:216-228)
void ExtractPlaneNormal(const rh_tri_accel &tri, float *out_normal);

struct HlbvhTreelet;
// The bottom half of the fast build done elsewhere (on the device): given the
// primitives and the index offset of their first sorted position, fills
// order (the primitives in Morton order), bottom (the treelets' nodes) and
// treelets exactly as the host's own steps would; throws on failure.
using LbvhBottom = std::function<void(const prim_t *prims, size_t n, uint32_t index_offset,
                                      std::vector<uint32_t> &order, std::vector<rh_bvh_node> &bottom,
                                      std::vector<HlbvhTreelet> &treelets)>;

// An SAH build of boxes (6 floats each: bbox_min, bbox_max) done elsewhere
// (on the device): fills nodes (root first, child links relative to
// nodes[0], leaves indexing `indices` from 0) and indices exactly as
// PreprocessPrims_SAH without spatial splits would, or returns false when it
// cannot (the caller then builds on the host); throws on failure.
using SahBoxes = std::function<bool(const float *boxes, size_t n, const bvh_settings_t &s,
                                    std::vector<rh_bvh_node> &nodes, std::vector<uint32_t> &indices)>;

// PreprocessMesh (Core.cpp:230-275); with s.use_fast_bvh_build the tree is
// built by PreprocessPrims_HLBVH, its bottom half by `device` and its top
// tree by `sah` when given.
uint32_t PreprocessMesh(const float *attrs, const uint32_t *vtx_indices, size_t vtx_indices_count, int layout,
                        int base_vertex, const bvh_settings_t &s, std::vector<rh_bvh_node> &out_nodes,
                        std::vector<rh_tri_accel> &out_tris, std::vector<uint32_t> &out_tri_indices,
                        const LbvhBottom *device = nullptr, const SahBoxes *sah = nullptr);

// PreprocessPrims_SAH (Core.cpp:431-530) + SplitPrimitives_SAH (BVHSplit.cpp:126-390)
uint32_t PreprocessPrims_SAH(const prim_t *prims, size_t prims_count, const float *positions, size_t stride,
                             const bvh_settings_t &s, std::vector<rh_bvh_node> &out_nodes,
                             std::vector<uint32_t> &out_indices);
// The same without spatial splits, built by `sah` where it can: its nodes are
// appended as the host's would be (child links and leaf offsets rebased).
uint32_t PreprocessPrims_SAH_Boxes(const prim_t *prims, size_t prims_count, const bvh_settings_t &s,
                                   std::vector<rh_bvh_node> &out_nodes, std::vector<uint32_t> &out_indices,
                                   const SahBoxes *sah);

// The fast build (PreprocessPrims_HLBVH, Core.cpp:532-680): primitives sorted
// (stably) by the 30-bit Morton code of their centres in the mesh box
// (10 bits per axis, Core.cpp:21-40, 549-562); runs sharing the top 12 bits
// are treelets, each built by EmitLBVH (Core.cpp:342-429) from bit 17 down;
// a SAH tree over the treelets' boxes (every treelet its own leaf) joins them.
// The node array is the reference's: the top tree's interior nodes, then the
// treelets' nodes in treelet order.
constexpr int kMortonBitsPerAxis = 10;
constexpr uint32_t kTreeletMask = 0x3ffc0000u; // Morton bits 18..29
constexpr int kTreeletBit = 17;
struct HlbvhTreelet {
    uint32_t first, count; // range of the code-sorted primitives
    uint32_t node_index;   // root within the treelets' node array
};
uint32_t F2U32_x86(float f);
uint32_t MortonCode3(uint32_t x, uint32_t y, uint32_t z);
void MortonCodes(const prim_t *prims, size_t n, uint32_t *codes);
uint32_t EmitLBVH(const prim_t *prims, const uint32_t *indices, const uint32_t *codes, uint32_t first, uint32_t count,
                  uint32_t index_offset, int bit, std::vector<rh_bvh_node> &out);
void TopLevelSAH(const std::vector<HlbvhTreelet> &treelets, const std::vector<rh_bvh_node> &bottom,
                 std::vector<rh_bvh_node> &top, std::vector<uint32_t> &top_indices, const SahBoxes *sah = nullptr);
uint32_t AssembleHLBVH(const std::vector<HlbvhTreelet> &treelets, const std::vector<rh_bvh_node> &bottom,
                       const std::vector<rh_bvh_node> &top, const std::vector<uint32_t> &top_indices,
                       std::vector<rh_bvh_node> &out_nodes);
uint32_t PreprocessPrims_HLBVH(const prim_t *prims, size_t prims_count, std::vector<rh_bvh_node> &out_nodes,
                               std::vector<uint32_t> &out_indices, const LbvhBottom *device = nullptr,
                               const SahBoxes *sah = nullptr);

// FlattenBVH_Recursive (Core.cpp:683-819): appends the 8-wide form of the
// binary tree rooted at nodes[root] to out (depth-first, a node before its
// children) and returns the index of its root.  An interior node takes the
// binary nodes two levels below it (a leaf child stands for itself), placed
// in the slot of their centre's Morton code within the children's centre box,
// the next free slot on a collision.
uint32_t FlattenBVH(const rh_bvh_node *nodes, uint32_t root, std::vector<rh_mbvh_node> &out);
// Bounds of the wide node FlattenBVH makes for nodes[root]: slot 0 of a leaf,
// else the union of its children's boxes (SceneRef.cpp:383-406, 52-66)
void WideRootBounds(const rh_bvh_node *nodes, uint32_t root, float bmin[3], float bmax[3]);

// ConstructCamera (Core.cpp:856-886)
void ConstructCamera(int type, int filter, int dtype, const float origin[3], const float fwd[3], const float up[3],
                     float fov, float gamma, float focus_distance, float focus_factor, rh_camera *cam);

// TransformBoundingBox (Core.cpp:888-907)
void TransformBoundingBox(const float bbox_min[3], const float bbox_max[3], const float *xform, float out_min[3],
                          float out_max[3]);

// InverseMatrix (Core.cpp:909-950)
void InverseMatrix(const float mat[16], float out_mat[16]);

// ComputeTangentBasis (TextureUtilsRef.cpp:40-178)
void ComputeTangentBasis(size_t vtx_offset, size_t vtx_start, std::vector<rh_vertex> &vertices,
                         std::vector<uint32_t> &new_vtx_indices, const uint32_t *indices, size_t indices_count);

// DownsampleTexture (TextureUtilsRef.cpp:9-38); RGBA8 packed as uint32 (r in bits 0-7)
std::vector<uint32_t> DownsampleTexture(const std::vector<uint32_t> &tex, const int res[2]);

// Radical-inverse permutations (Halton.h:6-43) seeded as RendererRef.cpp:11-12
std::vector<uint16_t> ComputeRadicalInversePermutations();
// One 256 x 24 Halton table (RendererRef.cpp:327-358)
void UpdateHaltonSequence(const std::vector<uint16_t> &perms, int iteration, float *seq);

// Guillotine rectangle packer restating TextureSplitter::Allocate
// (TextureSplitter.cpp:3-10, 53-101): only placement is needed here, so the
// atlas position of every mip level equals the reference's.
class TextureSplitter {
    struct node_t {
        int parent = -1;
        int child[2] = {-1, -1};
        int pos[2] = {0, 0}, size[2] = {0, 0};
        bool is_free = true;
    };
    std::vector<node_t> nodes_;
    int Insert_Recursive(int i, const int res[2]);

public:
    explicit TextureSplitter(const int res[2]);
    int Allocate(const int res[2], int pos[2]);
};

} // namespace RayHip
