// NormalEstimation (src/NormalEstimation.hpp:14-20 of the reference).
//
// estimateNormals<k>(octree[, start_idx, n]) returns, for the octree points
// [start_idx, start_idx + n) (Morton order), the unit eigenvector of the
// smallest eigenvalue of their k-neighbourhood covariance, flipped toward the
// viewpoint (0,0,0) - the reference's result, computed by one kNN kernel and
// one fused covariance/eigensolve kernel instead of a kNN kernel, a covariance
// kernel and cuSOLVER syevjBatched.  The extended overload takes options
// (precision, learned search radius).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "Octree.hpp"
#include "PointCloud.hpp"

namespace shapes {

// Learned kNN search radius (weights file written by shapeextraction_amd/learned.py).
class RadiusModel {
public:
    explicit RadiusModel(const std::string& path);
    ~RadiusModel();
    RadiusModel(const RadiusModel&) = delete;
    RadiusModel& operator=(const RadiusModel&) = delete;
    sx_mlp* handle() const { return mlp_; }
    int k() const { return k_; }

private:
    sx_mlp* mlp_ = nullptr;
    int k_ = 0;
};

struct NormalOptions {
    int precision = SX_PREC_FP32;           // or SX_PREC_FP64
    const RadiusModel* radius_model = nullptr;  // seeds the exact kNN when set
    Point viewpoint = Point(0.0f, 0.0f, 0.0f);
};

std::vector<Point> estimate_normals(const OT::Octree& octree, int start_idx, int n, int k, const NormalOptions& opt);
// Normals of every input point of the cloud the octree was built from, in
// that cloud's order (duplicates share their code's normal).
std::vector<Point> estimate_normals_input_order(const OT::Octree& octree, int k, const NormalOptions& opt);

}  // namespace shapes

class NormalEstimation {
public:
    template <int k>
    static std::vector<Point> estimateNormals(const OT::Octree& octree) {
        return estimateNormals<k>(octree, 0, octree.n_pts);
    }
    template <int k>
    static std::vector<Point> estimateNormals(const OT::Octree& octree, const int start_idx, const int n) {
        static_assert(k == 8 || k == 16 || k == 24 || k == 32, "k must be 8, 16, 24 or 32");
        return shapes::estimate_normals(octree, start_idx, n, k, shapes::NormalOptions());
    }
    template <int k>
    static std::vector<Point> estimateNormals(const OT::Octree& octree, const int start_idx, const int n,
                                              const shapes::NormalOptions& opt) {
        static_assert(k == 8 || k == 16 || k == 24 || k == 32, "k must be 8, 16, 24 or 32");
        return shapes::estimate_normals(octree, start_idx, n, k, opt);
    }
    // Not in the reference, which returns Morton order (src/main.cu:119-120):
    // one normal per point of the input cloud, in its order, ready for
    // PointCloud::setNormals.
    template <int k>
    static std::vector<Point> estimateNormalsInputOrder(const OT::Octree& octree,
                                                        const shapes::NormalOptions& opt = shapes::NormalOptions()) {
        static_assert(k == 8 || k == 16 || k == 24 || k == 32, "k must be 8, 16, 24 or 32");
        return shapes::estimate_normals_input_order(octree, k, opt);
    }
};
