// Internal declarations shared by the host-side scene builder and the HIP
// tracer of libaa_tracer.so.  Not part of the public C ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/aa_tracer.h"

struct aa_scene_s;

namespace aa {

// Canonical binary BVH in the reference layout (acoustic_agent/bvh.py:8-77):
// preorder nodes, leaves hold `count` primitives starting at `start` in
// `prims`, escape[n] = index of the node after n's subtree.
struct CanonicalBVH {
    std::vector<double> lo, hi;  // [nodes,3]
    std::vector<int64_t> start, count, escape, prims;
    int64_t leaf_count = 0;
    int64_t max_depth = 0;
    int64_t node_count() const { return (int64_t)start.size(); }
};

// Returns an empty string on success, else the reason (mirrors the
// ValueErrors of bvh.py:18-26).
std::string build_canonical_bvh(const double* lower, const double* upper, int64_t n, int64_t leaf_size,
                                CanonicalBVH& out);

// Binned-SAH binary tree in the same layout (used for the device tree: the
// traversal result does not depend on the tree, only its cost does).
std::string build_sah_bvh(const double* lower, const double* upper, int64_t n, int64_t leaf_size,
                          CanonicalBVH& out);

// 4-wide BVH node as stored on the device: 8 x float4 = 128 bytes.
//   v[0] lo_x[4]  v[1] hi_x[4]  v[2] lo_y[4]  v[3] hi_y[4]
//   v[4] lo_z[4]  v[5] hi_z[4]  v[6] child[4] (int32 bits)  v[7] unused
// child >= 0: internal node index; child == -1: empty slot (its box is a
// point at 1e30, which no ray reaches); child <= -2: leaf with
// v = -(child + 2), first triangle v >> 3, count (v & 7) + 1.
struct WideNode {
    float v[8][4];
};

constexpr int kEmptyChild = -1;
constexpr int kMaxLeafSize = 8;

inline int32_t encode_leaf(int64_t start, int64_t count) {
    return (int32_t)(-((int64_t)((start << 3) | (count - 1)) + 2));
}

struct WideBVH {
    std::vector<WideNode> nodes;  // breadth-first order: node 0 is the root
    int depth = 0;                // levels of 4-wide nodes
    std::vector<int32_t> level_end;  // nodes [0, level_end[l]) cover levels 0..l
};

// Collapse the canonical tree to 4-wide nodes (children chosen by expanding
// the largest-area internal child first), laid out breadth first.
void collapse_to_wide(const CanonicalBVH& bin, WideBVH& out);


// Record a message for aa_last_error() and return `code`.
int set_error(int code, const char* fmt, ...);

struct DeviceProps {
    int cu = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
    bool valid = false;
};
const DeviceProps& device_props(int device);

// Validate one trace request (shared by aa_trace and the batch path).
int validate_trace(const ::aa_scene_s* scene, const aa_trace_params* p, const float* d_sources,
                   const float* d_directions, const float* d_diffuse_bank, const float* d_diffuse_random,
                   const int32_t* d_diffuse_indices, const aa_trace_outputs* o);

}  // namespace aa

// Opaque scene handle of the C ABI.
struct aa_scene_s {
    int32_t device = 0;
    int32_t num_triangles = 0, num_surfaces = 0;
    int32_t leaf_size = 4;
    int32_t canonical_nodes = 0, canonical_depth = 0;
    aa::WideBVH wide;
    float* d_nodes = nullptr;      // [nodes, 8, 4]
    float* d_tris = nullptr;       // [T, 3, 4] leaf order
    // canonical binary tree for the plain one-thread-per-ray kernel:
    // per node (lo xyz | escape) (hi xyz | start << 4 | count); triangles in
    // its primitive order, same 3 x float4 format as d_tris
    float* d_plain_nodes = nullptr;
    float* d_plain_tris = nullptr;
    int32_t plain_node_count = 0;
    float* d_normals = nullptr;    // [T, 4]
    float* d_reflection = nullptr; // [S, 6]
    float* d_scattering = nullptr; // [S]
    uint64_t device_bytes = 0;
    uint64_t content_hash = 0;
};

