// Host-side KD-tree (RenderX's median-split object-partition tree) and its
// conversion to the device layout used by the traversal kernels.
#pragma once
#include <cstdint>
#include <vector>

namespace rx {

// One node in pre-order (left child = index + 1), the order of
// KDTree::serializeForGPU (src/KDTree.cpp:346-383).
struct PNode {
  float lo[3], hi[3];
  int32_t parent;
  int32_t right;        // -1 for a leaf
  uint32_t leaf_begin;  // leaf: range in leaf_tris
  uint32_t leaf_end;
};

struct HostTree {
  std::vector<PNode> nodes;
  std::vector<uint32_t> leaf_tris;  // triangle ids in leaf (DFS) order
  uint32_t max_depth = 0;
};

// KDTree::build(BEST) (src/KDTree.cpp:230-246).  abs_int selects the int
// abs(int) binding of the split criterion (src/KDTree.cpp:191).
void build_tree(const float* verts, uint32_t n_verts, const uint32_t* tris, uint32_t n_tris, bool abs_int,
                HostTree& out);

// Parse KDTree::serializeForGPU arrays back into a HostTree; fills the leaf
// triangle vertices (9 floats per leaf triangle) from v1/v2/v3.
bool parse_serialized(uint32_t n_nodes, const float* nodes4, const float* min4, const float* max4,
                      uint32_t n_leaf, const float* v1, const float* v2, const float* v3, HostTree& out,
                      std::vector<float>& leaf_verts);

// Device layout (see rx_dev.h DevTree): interior nodes carry both children's
// boxes (64 B), leaves are ranges of a leaf-ordered triangle array.
struct DeviceTreeHost {
  std::vector<float> nodes;       // 16 floats per interior node
  std::vector<int32_t> leaf_start;
  std::vector<float> tv0, tv1, tv2, tn;  // 4 floats per leaf-ordered triangle
  float root_lo[3], root_hi[3];
  int32_t root_ref = 0;
  uint32_t n_interior = 0, n_leaves = 0;
};

// leaf_verts: 9 floats per leaf position; tri_obj/tri_normal/tri_local may be
// null (object -1, normal computed from the vertices, local = tid).
void make_device_tree(const HostTree& t, const std::vector<float>& leaf_verts, const int32_t* tri_obj,
                      const uint32_t* tri_local, DeviceTreeHost& out);

// Face normal normalize((v1-v0) x (v2-v0)) with the reference's float
// operation order (SimpleShape::getWorldNormal flat branch, src/SimpleShape.cpp:404-421).
void face_normal(const float* v0, const float* v1, const float* v2, float* n);

}  // namespace rx
