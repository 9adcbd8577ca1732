// OT::Octree (src/Octree.hpp:16-126 of the reference) on the MI355X pipeline.
//
// u_nodes / u_points are the device arrays of the index (OTNode layout and node
// numbering identical to the reference; points are the decoded unique points
// in Morton order); h_points is their host copy.  deviceKnnSearch / knnSearch
// run the exact GPU kNN (rows ascending by squared distance, ties by index,
// the query itself included when it is an index point); eps is accepted for
// source compatibility and ignored: the reference never uses it either
// (src/Octree.hpp:152-155).
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "PointCloud.hpp"
#include "RadixTree.hpp"
#include "ShapesError.hpp"

namespace OT {
constexpr static int SEARCH_Q_SIZE = 32;  // the reference's queue bound; the exact search has none

struct OTNode {
    int children[8];
    Point corner;
    float cell_size;
    int child_node_mask;  // bit i: children[i] is an octree node
    int child_leaf_mask;  // bit i: children[i] is a point index
};
static_assert(sizeof(OTNode) == sizeof(sx_otnode), "OTNode must match sx_otnode (56 bytes)");

class Octree {
public:
    Octree() {}
    explicit Octree(const RT::RadixTree& radix_tree);
    // From host data (the reference uses it after an MPI broadcast): uploads
    // nodes and points and indexes the points again for the GPU search.
    Octree(std::shared_ptr<std::vector<OTNode>> nodes, const int n_nodes, std::shared_ptr<std::vector<Point>> points,
           const int n_pts);
    ~Octree();
    Octree(const Octree&) = delete;
    Octree& operator=(const Octree&) = delete;
    Octree& operator=(Octree&& other);

    template <int k>
    std::vector<std::array<int, k>> knnSearch(const std::vector<Point>& points, const float eps = 0.01f) const;

    template <int k>
    void deviceKnnSearch(const Point* d_query, int* d_nn_indices, const int N, const float eps) const;

    int n_pts = 0;
    Point* u_points = nullptr;
    std::shared_ptr<std::vector<Point>> h_points;
    int n_nodes = 0;
    OTNode* u_nodes = nullptr;

    std::shared_ptr<shapes::IndexHandle> index() const { return index_; }

private:
    void knn(const Point* d_query, int* d_nn, int N, int k) const;
    std::vector<int> knn_host(const std::vector<Point>& points, int k) const;
    void release();
    std::shared_ptr<shapes::IndexHandle> index_;
    void* own_nodes_ = nullptr;  // device copies made by the host-data constructor
};

template <int k>
void Octree::deviceKnnSearch(const Point* d_query, int* d_nn_indices, const int N, const float) const {
    static_assert(k == 8 || k == 16 || k == 24 || k == 32, "k must be 8, 16, 24 or 32");
    knn(d_query, d_nn_indices, N, k);
}

template <int k>
std::vector<std::array<int, k>> Octree::knnSearch(const std::vector<Point>& points, const float) const {
    static_assert(k == 8 || k == 16 || k == 24 || k == 32, "k must be 8, 16, 24 or 32");
    const std::vector<int> flat = knn_host(points, k);
    std::vector<std::array<int, k>> out(points.size());
    for (size_t i = 0; i < points.size(); ++i)
        for (int j = 0; j < k; ++j) out[i][j] = flat[i * k + j];
    return out;
}
}  // namespace OT
