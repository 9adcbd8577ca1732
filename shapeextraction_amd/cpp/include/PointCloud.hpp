// Point and PointCloud (src/PointCloud.hpp:21-149 of the reference), host side.
//
// Point keeps its 12-byte layout (x, y, z) so that arrays of Point are the
// same bytes as the library's 12-byte AoS point arrays.  distance2 fixes the
// operation order ((dx*dx + dy*dy) + dz*dz), the order the GPU kernels use.
// getMpiDatatype is not provided: this build has no MPI transport.
#pragma once

#include <cassert>
#include <cstddef>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

struct Point {
    // Be warned, x, y, z are not initialised by the default constructor.
    Point() {}
    Point(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    float x, y, z;

    static float distance2(const Point& p1, const Point& p2) {
        const float dx = p1.x - p2.x, dy = p1.y - p2.y, dz = p1.z - p2.z;
        const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
        const float s = xx + yy;
        return s + zz;
    }
    float dot(const Point& o) const {
        const float a = x * o.x, b = y * o.y, c = z * o.z;
        const float s = a + b;
        return s + c;
    }
    Point& operator+=(const Point& r) { x += r.x; y += r.y; z += r.z; return *this; }
    Point& operator-=(const Point& r) { x -= r.x; y -= r.y; z -= r.z; return *this; }
    Point& operator/=(float r) { x /= r; y /= r; z /= r; return *this; }
    friend Point operator+(Point l, const Point& r) { l += r; return l; }
    friend Point operator-(Point l, const Point& r) { l -= r; return l; }
};
static_assert(sizeof(Point) == 12 && std::is_standard_layout<Point>::value, "Point must stay 12 bytes");

template <typename T>
class PointCloud {
public:
    PointCloud() {}
    PointCloud(const std::shared_ptr<const std::vector<Point>>& points, const std::vector<Point>& normals);
    std::vector<T> x_vals, y_vals, z_vals;
    std::vector<T> x_normals, y_normals, z_normals;

    // ASCII PLY as the reference writes it (src/PointCloud.hpp:152-190).
    bool saveAsPly(std::string filename);
    // Binary little-endian PLY with the same header fields (much faster).
    bool saveAsBinaryPly(std::string filename);
    void setNormals(std::vector<Point> normals);
};

template <typename T>
PointCloud<T>::PointCloud(const std::shared_ptr<const std::vector<Point>>& points_ptr,
                          const std::vector<Point>& normals) {
    const auto& points = *points_ptr;
    assert(points.size() == normals.size());
    const size_t n = points.size();
    x_vals.resize(n); y_vals.resize(n); z_vals.resize(n);
    x_normals.resize(n); y_normals.resize(n); z_normals.resize(n);
    for (size_t i = 0; i < n; ++i) {
        x_vals[i] = points[i].x; y_vals[i] = points[i].y; z_vals[i] = points[i].z;
        x_normals[i] = normals[i].x; y_normals[i] = normals[i].y; z_normals[i] = normals[i].z;
    }
}

template <>
bool PointCloud<float>::saveAsPly(std::string filename);
template <>
bool PointCloud<float>::saveAsBinaryPly(std::string filename);
template <>
void PointCloud<float>::setNormals(std::vector<Point> normals);
