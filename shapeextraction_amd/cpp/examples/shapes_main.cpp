// The reference's driver (src/main.cu:50-124) for one process, written
// against the drop-in API: load a cloud, build the radix tree and octree,
// estimate normals, write a PLY.  Usage:
//   shapes [out.ply] [input.{bin,obj,txt} | synthetic:N] [k] [radius_model.bin]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "DataIO.hpp"
#include "NormalEstimation.hpp"
#include "Octree.hpp"
#include "RadixTree.hpp"

namespace {
double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

template <int k>
std::vector<Point> normals_k(const OT::Octree& octree, const shapes::NormalOptions& opt) {
    return NormalEstimation::estimateNormals<k>(octree, 0, octree.n_pts, opt);
}
}  // namespace

int main(int argc, char* argv[]) {
    std::string out_file_name = argc > 1 ? argv[1] : "cloud.ply";
    std::string in_file_name = argc > 2 ? argv[2] : "synthetic:100000";
    const int k = argc > 3 ? std::atoi(argv[3]) : 8;
    try {
        PointCloud<float> input_cloud;
        if (in_file_name.rfind("synthetic:", 0) == 0) {
            const long n = std::atol(in_file_name.c_str() + 10);
            input_cloud.x_vals.resize(n);
            input_cloud.y_vals.resize(n);
            input_cloud.z_vals.resize(n);
            shapes::check(sx_gen_scene(2, n, 200, 100.0f, 0.01f, 0.01f, input_cloud.x_vals.data(),
                                       input_cloud.y_vals.data(), input_cloud.z_vals.data()),
                          "sx_gen_scene");
        } else {
            input_cloud = DataIO::loadFile(in_file_name);
        }
        std::cout << "Input data has " << input_cloud.x_vals.size() << " points" << std::endl;

        auto t = std::chrono::steady_clock::now();
        RT::RadixTree radix_tree(input_cloud);
        std::cout << "Radix tree construction took " << ms_since(t) << "ms" << std::endl;
        t = std::chrono::steady_clock::now();
        OT::Octree octree(radix_tree);
        std::cout << "Octree construction took " << ms_since(t) << "ms (" << octree.n_nodes << " nodes)" << std::endl;

        shapes::NormalOptions opt;
        std::unique_ptr<shapes::RadiusModel> model;
        if (argc > 4) {
            model.reset(new shapes::RadiusModel(argv[4]));
            opt.radius_model = model.get();
        }
        t = std::chrono::steady_clock::now();
        std::vector<Point> normals;
        switch (k) {
            case 8: normals = normals_k<8>(octree, opt); break;
            case 16: normals = normals_k<16>(octree, opt); break;
            case 24: normals = normals_k<24>(octree, opt); break;
            case 32: normals = normals_k<32>(octree, opt); break;
            default: std::cerr << "k must be 8, 16, 24 or 32" << std::endl; return 2;
        }
        std::cout << "Normal estimation took " << ms_since(t) << "ms" << std::endl;

        PointCloud<float> output_cloud(octree.h_points, normals);
        const bool binary = out_file_name.size() > 8 && out_file_name.rfind(".bin.ply") == out_file_name.size() - 8;
        if (!(binary ? output_cloud.saveAsBinaryPly(out_file_name) : output_cloud.saveAsPly(out_file_name))) {
            std::cerr << "cannot write " << out_file_name << std::endl;
            return 1;
        }
    } catch (const shapes::Error& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
