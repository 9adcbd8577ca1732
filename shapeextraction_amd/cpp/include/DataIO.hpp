// DataIO (src/DataIO.hpp:5-19 of the reference): point-cloud file loaders.
// loadFile dispatches on the extension (.obj, .bin KITTI float4, .txt
// Semantic3D) and throws shapes::Error for any other extension (the
// reference falls off the end of the function).
#pragma once

#include <string>

#include "PointCloud.hpp"

class DataIO {
public:
    static PointCloud<float> loadFile(const std::string& filepath);
    // KITTI .bin: float4 (x, y, z, reflectance) records; num = maximum floats to read.
    static PointCloud<float> loadKitti(const std::string& filepath, size_t num);
    static void saveKitti(const PointCloud<float> cloud, const std::string& filepath);
    static PointCloud<float> loadObj(const std::string& filepath);
    static PointCloud<float> loadSemantic3D(const std::string& filepath);
};
