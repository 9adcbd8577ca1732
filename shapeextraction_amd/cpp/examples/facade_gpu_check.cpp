// GPU checks of the drop-in API that need device results, driven by
// tests/test_facade.py:
//   * Octree(host nodes, n, host points, n) - the constructor the reference's
//     MPI ranks use after a broadcast (src/main.cu:30-47, src/Octree.cu:
//     179-192) - rebuilt from a built octree's host copy gives the same nodes,
//     points, kNN rows and normals;
//   * Octree::knnSearch<k> on arbitrary query points (src/Octree.hpp:214-242)
//     for k = 8, 16, 24, 32: rows written to <dir>/knn<k>.i32 with the
//     queries (<dir>/queries.f32) and the octree points (<dir>/points.f32) for
//     the test to compare with brute force;
//   * NormalEstimation::estimateNormalsInputOrder<16>: one normal per input
//     point (<dir>/normals_input.f32), with the Morton-order normals
//     (<dir>/normals_unique.f32).
// Usage: facade_gpu_check <dir> [n_points]; prints "facade gpu check ok".
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "NormalEstimation.hpp"
#include "Octree.hpp"
#include "RadixTree.hpp"

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);   \
            return 1;                                                             \
        }                                                                         \
    } while (0)

namespace {
template <typename T>
void dump(const std::string& path, const T* data, size_t count) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(count * sizeof(T)));
}

template <int k>
bool same_rows(const OT::Octree& a, const OT::Octree& b, const std::vector<Point>& q) {
    return a.knnSearch<k>(q) == b.knnSearch<k>(q);
}

template <int k>
void knn_dump(const OT::Octree& octree, const std::vector<Point>& q, const std::string& dir) {
    const std::vector<std::array<int, k>> rows = octree.knnSearch<k>(q);
    dump(dir + "/knn" + std::to_string(k) + ".i32", rows.empty() ? nullptr : rows[0].data(), rows.size() * k);
}

// Deterministic queries: inside the cube, around it, far away, and on points.
std::vector<Point> make_queries(const std::vector<Point>& pts, float lo, float hi, size_t n) {
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto u01 = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<float>((s >> 40) & 0xFFFFFF) / 16777216.0f;
    };
    const float ext = hi - lo;
    std::vector<Point> q;
    for (size_t i = 0; i < n; ++i) {
        switch (i % 4) {
            case 0: q.emplace_back(lo + u01() * ext, lo + u01() * ext, lo + u01() * ext); break;
            case 1: q.emplace_back(lo - ext + 3 * u01() * ext, lo - ext + 3 * u01() * ext, lo - ext + 3 * u01() * ext); break;
            case 2: q.emplace_back(1e4f * (u01() - 0.5f), 1e4f * (u01() - 0.5f), 1e4f * (u01() - 0.5f)); break;
            default: q.push_back(pts[static_cast<size_t>(u01() * static_cast<float>(pts.size() - 1))]); break;
        }
    }
    return q;
}
}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const long n = argc > 2 ? std::atol(argv[2]) : 200000;
    try {
        PointCloud<float> cloud;
        cloud.x_vals.resize(n);
        cloud.y_vals.resize(n);
        cloud.z_vals.resize(n);
        shapes::check(sx_gen_scene(7, n, 200, 100.0f, 0.01f, 0.01f, cloud.x_vals.data(), cloud.y_vals.data(),
                                   cloud.z_vals.data()),
                      "sx_gen_scene");
        // duplicates: every 9th input point again, so input order != unique order
        for (long i = 0; i < n; i += 9) {
            cloud.x_vals.push_back(cloud.x_vals[i]);
            cloud.y_vals.push_back(cloud.y_vals[i]);
            cloud.z_vals.push_back(cloud.z_vals[i]);
        }
        RT::RadixTree radix_tree(cloud);
        OT::Octree a(radix_tree);
        std::printf("octree: %d points, %d nodes\n", a.n_pts, a.n_nodes);

        // ---- host-data constructor round trip
        auto h_nodes = std::make_shared<std::vector<OT::OTNode>>(static_cast<size_t>(a.n_nodes));
        shapes::check(sx_memcpy_dtoh(h_nodes->data(), a.u_nodes, sizeof(OT::OTNode) * h_nodes->size(), nullptr),
                      "sx_memcpy_dtoh");
        auto h_points = std::make_shared<std::vector<Point>>(*a.h_points);
        OT::Octree b(h_nodes, a.n_nodes, h_points, a.n_pts);
        CHECK(b.n_pts == a.n_pts && b.n_nodes == a.n_nodes);
        std::vector<Point> pb(static_cast<size_t>(b.n_pts));
        shapes::check(sx_memcpy_dtoh(pb.data(), b.u_points, sizeof(Point) * pb.size(), nullptr), "sx_memcpy_dtoh");
        CHECK(std::memcmp(pb.data(), a.h_points->data(), sizeof(Point) * pb.size()) == 0);
        std::vector<OT::OTNode> nb(h_nodes->size());
        shapes::check(sx_memcpy_dtoh(nb.data(), b.u_nodes, sizeof(OT::OTNode) * nb.size(), nullptr), "sx_memcpy_dtoh");
        CHECK(std::memcmp(nb.data(), h_nodes->data(), sizeof(OT::OTNode) * nb.size()) == 0);
        CHECK(same_rows<16>(a, b, *a.h_points));
        const std::vector<Point> na = NormalEstimation::estimateNormals<16>(a);
        const std::vector<Point> nbn = NormalEstimation::estimateNormals<16>(b);
        CHECK(std::memcmp(na.data(), nbn.data(), sizeof(Point) * na.size()) == 0);
        // a moved-to octree keeps working
        OT::Octree c;
        c = std::move(b);
        CHECK(b.n_pts == 0 && c.n_pts == a.n_pts);
        CHECK(same_rows<8>(a, c, *a.h_points));
        // points that are not the octree's grid points are rejected
        auto bad = std::make_shared<std::vector<Point>>(*a.h_points);
        (*bad)[a.n_pts / 2].x = std::nextafter((*bad)[a.n_pts / 2].x, 1e9f);
        bool threw = false;
        try {
            OT::Octree d(h_nodes, a.n_nodes, bad, a.n_pts);
        } catch (const shapes::Error& e) {
            threw = e.status() == SX_EINVAL;
        }
        CHECK(threw);
        std::printf("host-data octree: nodes, points, kNN rows and normals identical\n");

        // ---- knnSearch<k> on arbitrary queries
        const std::vector<Point> q = make_queries(*a.h_points, radix_tree.min_coord, radix_tree.max_coord, 4000);
        CHECK(same_rows<16>(a, c, q));
        dump(dir + "/points.f32", &(*a.h_points)[0].x, 3 * a.h_points->size());
        dump(dir + "/queries.f32", &q[0].x, 3 * q.size());
        knn_dump<8>(a, q, dir);
        knn_dump<16>(a, q, dir);
        knn_dump<24>(a, q, dir);
        knn_dump<32>(a, q, dir);

        // ---- normals in input order
        const std::vector<Point> nin = NormalEstimation::estimateNormalsInputOrder<16>(a);
        CHECK(nin.size() == cloud.x_vals.size());
        dump(dir + "/normals_input.f32", &nin[0].x, 3 * nin.size());
        dump(dir + "/normals_unique.f32", &na[0].x, 3 * na.size());
        dump(dir + "/cloud_x.f32", cloud.x_vals.data(), cloud.x_vals.size());
        dump(dir + "/cloud_y.f32", cloud.y_vals.data(), cloud.y_vals.size());
        dump(dir + "/cloud_z.f32", cloud.z_vals.data(), cloud.z_vals.size());
    } catch (const shapes::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("facade gpu check ok\n");
    return 0;
}
