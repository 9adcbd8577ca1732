// RT::RadixTree (src/RadixTree.hpp:16-61 of the reference) on the MI355X
// pipeline.
//
// The constructor copies the cloud to HBM and runs the whole index build on
// the GPU (bounds, Morton encode, radix sort, duplicate removal, Karras tree,
// boxes, octree) through sx_index_build.  d_tree's arrays are device arrays
// (hipMalloc, not managed memory) sized n_pts; prefixN/leftChild/... hold the
// reference's values for the n_nodes = n_pts - 1 internal nodes.
// Differences from the reference: errors throw shapes::Error; the object is
// move-only (the reference's implicit copy would double-free); the root
// covers all leaves unless root_mode is SX_ROOT_REFERENCE.
#pragma once

#include <cstdint>
#include <memory>

#include "MortonUtils.hpp"
#include "PointCloud.hpp"
#include "ShapesError.hpp"

namespace shapes {
// Shared owner of one sx_index (a radix tree and the octree built from it
// share it, so the octree outlives the radix tree as in the reference's main).
struct IndexHandle {
    sx_index* index = nullptr;
    sx_index_view view{};
    explicit IndexHandle(int64_t capacity, int code_bits, int root_mode);
    ~IndexHandle();
    IndexHandle(const IndexHandle&) = delete;
    IndexHandle& operator=(const IndexHandle&) = delete;
};
}  // namespace shapes

namespace RT {
struct Nodes {
    Code_t* mortonCode;   // 63-bit codes, sorted and unique
    bool* hasLeafLeft;
    bool* hasLeafRight;
    uint8_t* prefixN;     // delta_node of Karras
    int* leftChild;       // right child is leftChild + 1
    int* parent;
};

class RadixTree {
public:
    explicit RadixTree(const PointCloud<float>& cloud, int root_mode = SX_ROOT_TEXTBOOK);
    ~RadixTree();
    RadixTree(const RadixTree&) = delete;
    RadixTree& operator=(const RadixTree&) = delete;
    RadixTree(RadixTree&&) = default;

    Nodes d_tree;
    int n_pts;    // unique points
    int n_nodes;  // n_pts - 1
    float min_coord, max_coord;

    // The index the octree and kNN use.
    std::shared_ptr<shapes::IndexHandle> index() const { return index_; }

private:
    std::shared_ptr<shapes::IndexHandle> index_;
};
}  // namespace RT
