// lbvh.h — batched LBVH builder (Morton codes, LDS radix sort, Karras hierarchy,
// bottom-up box fit, collapse to 4-wide 128-byte nodes).
//
// It builds the structures the Vulkan driver builds for
// vkCmdBuildAccelerationStructuresKHR in the reference
// (crates/libs/vulkan/src/command.rs:354-372): one BVH per geometry (BLAS) and one
// over the instances (TLAS).  Several BLASes are built in one pass: the segment id
// sits above the Morton bits of each key, so one sort and one radix tree cover the
// whole batch and every segment is its own subtree.
#pragma once

#include <vector>

#include "common.h"

namespace rt {

// One segment (= one BLAS) of a batched triangle build.
struct TriSegment {
    const uint8_t* vertex_data;
    const uint32_t* index_data;  // already offset by the primitive offset
    uint32_t stride;
    uint32_t max_vertex;
    uint32_t first_vertex;
    uint32_t prim_count;
    uint32_t prim_start;         // first global primitive of the segment
    uint32_t _pad;
};

// Scratch and results of one build; device pointers are owned by the caller
// (api.hip) so that they can be pooled.
struct LbvhResult {
    QNode* nodes = nullptr;         // final (quantised) node array, exact size
    uint32_t num_nodes = 0;
    uint32_t* perm = nullptr;       // sorted position -> input primitive index
    float* seg_bounds = nullptr;    // 6 floats per segment (padded root box)
    std::vector<uint32_t> level_start;  // nodes of tree level L: [level_start[L], level_start[L + 1])
};

// How the binary radix tree becomes 4-wide nodes.
enum : uint32_t {
    kCollapseArea = 0,  // greedy: open the largest-surface-area child until four are gathered
    kCollapseSah = 1    // least surface-area-heuristic cost (dynamic program over the binary tree)
};

// Builds the wide BVH of `n` primitives with boxes (lo[i], hi[i]) split into
// `nseg` segments (d_seg_start has nseg + 1 entries, on device and host).
// Segment s's root is node s.  Leaves hold at most leaf_max primitives.
// treelet_passes: bottom-up treelet restructurings of the binary tree before the
// collapse (0: none).  Returns a hipError_t-compatible code.
int lbvh_build(hipStream_t stream, const float4* d_lo, const float4* d_hi, uint32_t n,
               const uint32_t* d_seg_start, const std::vector<uint32_t>& h_seg_start,
               uint32_t leaf_max, uint32_t collapse, uint32_t treelet_passes, LbvhResult* out, std::string* err);

// Triangle preparation for a BLAS batch: reads the caller's index and vertex
// buffers, writes triangle records (input order) and boxes.  Sets *d_bad to 1 if an
// index exceeds its segment's max_vertex (that triangle is then degenerate).
int prepare_triangles(hipStream_t stream, const TriSegment* d_segs, uint32_t nseg, uint32_t n,
                      float4* d_tri, float4* d_lo, float4* d_hi, uint32_t* d_bad);

// Gathers triangle records into BVH leaf order.
int reorder_triangles(hipStream_t stream, const float4* d_tri_in, const uint32_t* d_perm, uint32_t n,
                      float4* d_tri_out);

// One flattened instance: `count` triangle records of a BLAS pool starting at
// record `first_rec`, placed from flattened triangle `start` on and tagged with
// `inst_bits` (instance index | kFlatOpaque).
struct FlatSource {
    const float4* tris;
    uint32_t first_rec;
    uint32_t start;
    uint32_t inst_bits;
    uint32_t _pad;
};

// Copies the triangle records of identity instances into one array (input order of
// a flattened BVH build) with their boxes; records get v1.w = inst_bits and
// v2.w = source record index.
int gather_flat(hipStream_t stream, const FlatSource* d_src, uint32_t nsrc, uint32_t n, float4* d_tri, float4* d_lo,
                float4* d_hi);

// Refit of a built BLAS pool to new vertex data with the same topology: rewrites the
// triangle records' positions from `d_segs` (record r keeps its primitive index and
// segment), then recomputes every node's child boxes bottom-up level by level, padded
// and quantised as the build does, and the BlasDescs' root bounds.  Sets *d_bad to 1
// if an index exceeds its segment's max_vertex or a primitive index its segment's
// count.
int lbvh_refit(hipStream_t stream, const TriSegment* d_segs, uint32_t nseg, uint32_t n, float4* d_tri, QNode* d_nodes,
               uint32_t num_nodes, const std::vector<uint32_t>& level_start, BlasDesc* d_descs, uint32_t* d_bad,
               std::string* err);

// Refit of a flattened BVH (one segment, root node 0) after the BLASes its records
// come from were refit: every record takes its positions again from its BLAS record,
// then the node boxes are recomputed bottom-up as in lbvh_refit.  `bounds` gets the
// padded root box.
int lbvh_refit_flat(hipStream_t stream, const InstanceRecord* d_inst, uint32_t n, float4* d_tri, QNode* d_nodes,
                    uint32_t num_nodes, const std::vector<uint32_t>& level_start, float bounds[6], std::string* err);

// Refit of the TLAS over n instance records in leaf order (their BLAS bounds and
// transforms as they are now), node boxes recomputed bottom-up.
int lbvh_refit_instances(hipStream_t stream, const InstanceRecord* d_inst, uint32_t n, QNode* d_nodes,
                         uint32_t num_nodes, const std::vector<uint32_t>& level_start, float bounds[6],
                         std::string* err);

// Instance boxes (world space) for the TLAS build.
int prepare_instances(hipStream_t stream, const InstanceRecord* d_inst, uint32_t n, float4* d_lo,
                      float4* d_hi);
int reorder_instances(hipStream_t stream, const InstanceRecord* d_in, const uint32_t* d_perm, uint32_t n,
                      InstanceRecord* d_out);

}  // namespace rt
