// Host implementation of the drop-in C++ API (RadixTree, Octree,
// NormalEstimation, PointCloud, MortonUtils, DataIO) over the C ABI of
// libshapes_hip.so.  No HIP headers: every device operation goes through
// include/shapes_c.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "DataIO.hpp"
#include "MortonUtils.hpp"
#include "NormalEstimation.hpp"
#include "Octree.hpp"
#include "RadixTree.hpp"
#include "ShapesError.hpp"

namespace {

struct DeviceBuffer {
    void* ptr = nullptr;
    explicit DeviceBuffer(size_t bytes) { shapes::check(sx_malloc(&ptr, bytes), "sx_malloc"); }
    ~DeviceBuffer() { sx_free(ptr); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

inline uint32_t axis_scale() { return 0xFFFFFFFFu >> (32 - CODE_LEN / 3); }

}  // namespace

// ---------------------------------------------------------------- Morton
// Same operations, in the same order, as the GPU encode/decode kernels.
Code_t pointToCode(const float x, const float y, const float z, const float min_coord, const float range) {
    const float fs = static_cast<float>(axis_scale());
    auto q = [&](float v) -> uint64_t {
        if (!(range > 0.0f)) return 0;
        float t = (v - min_coord) / range;
        float s = fs * t;
        if (s < 0.0f) s = 0.0f;
        if (s > fs) s = fs;
        return static_cast<uint32_t>(s);
    };
    const uint64_t c[3] = {q(x), q(y), q(z)};
    Code_t code = 0;
    for (int i = 0; i < CODE_LEN / 3; ++i)
        for (int a = 0; a < 3; ++a) code |= ((c[a] >> i) & 1u) << (3 * i + a);
    return code;
}

Point codeToPoint(const Code_t code, const float min_coord, const float range) {
    const float fs = static_cast<float>(axis_scale());
    uint32_t r[3] = {0, 0, 0};
    for (int i = 0; i < CODE_LEN / 3; ++i)
        for (int a = 0; a < 3; ++a) r[a] |= static_cast<uint32_t>((code >> (3 * i + a)) & 1u) << i;
    float p[3];
    for (int a = 0; a < 3; ++a) {
        const float u = static_cast<float>(r[a]) / fs;
        const float v = u * range;
        p[a] = v + min_coord;
    }
    return Point(p[0], p[1], p[2]);
}

// ---------------------------------------------------------------- PointCloud
template <>
void PointCloud<float>::setNormals(std::vector<Point> normals) {
    if (normals.size() != x_vals.size()) throw shapes::Error(SX_EINVAL, "setNormals: size mismatch");
    x_normals.resize(normals.size());
    y_normals.resize(normals.size());
    z_normals.resize(normals.size());
    for (size_t i = 0; i < normals.size(); ++i) {
        x_normals[i] = normals[i].x;
        y_normals[i] = normals[i].y;
        z_normals[i] = normals[i].z;
    }
}

namespace {
void ply_header(std::ostream& of, const char* format, size_t n, bool has_normals) {
    of << "ply\nformat " << format << " 1.0\nelement vertex " << n
       << "\nproperty float x\nproperty float y\nproperty float z\n";
    if (has_normals) of << "property float nx\nproperty float ny\nproperty float nz\n";
    of << "end_header\n";
}
}  // namespace

template <>
bool PointCloud<float>::saveAsPly(std::string filename) {
    const bool has_normals = !x_normals.empty();
    std::ofstream of(filename);
    if (!of) return false;
    ply_header(of, "ascii", x_vals.size(), has_normals);
    for (size_t i = 0; i < x_vals.size(); ++i) {
        of << x_vals[i] << " " << y_vals[i] << " " << z_vals[i];
        if (has_normals) of << " " << x_normals[i] << " " << y_normals[i] << " " << z_normals[i];
        of << "\n";
    }
    return static_cast<bool>(of);
}

template <>
bool PointCloud<float>::saveAsBinaryPly(std::string filename) {
    const bool has_normals = !x_normals.empty();
    std::ofstream of(filename, std::ios::binary);
    if (!of) return false;
    ply_header(of, "binary_little_endian", x_vals.size(), has_normals);
    const size_t n = x_vals.size();
    const size_t w = has_normals ? 6 : 3;
    std::vector<float> rows(w * n);
    for (size_t i = 0; i < n; ++i) {
        float* r = &rows[w * i];
        r[0] = x_vals[i]; r[1] = y_vals[i]; r[2] = z_vals[i];
        if (has_normals) { r[3] = x_normals[i]; r[4] = y_normals[i]; r[5] = z_normals[i]; }
    }
    of.write(reinterpret_cast<const char*>(rows.data()), static_cast<std::streamsize>(rows.size() * sizeof(float)));
    return static_cast<bool>(of);
}

// ---------------------------------------------------------------- index handle
namespace shapes {

IndexHandle::IndexHandle(int64_t capacity, int code_bits, int root_mode) {
    check(sx_index_create(&index, capacity, code_bits, root_mode, 0.0f), "sx_index_create");
}

IndexHandle::~IndexHandle() { sx_index_destroy(index); }

namespace {
std::shared_ptr<IndexHandle> build_index(const float* x, const float* y, const float* z, size_t n, int root_mode) {
    if (n == 0) throw Error(SX_EINVAL, "cannot index an empty cloud");
    auto h = std::make_shared<IndexHandle>(static_cast<int64_t>(n), CODE_LEN, root_mode);
    const size_t bytes = n * sizeof(float);
    DeviceBuffer dx(bytes), dy(bytes), dz(bytes);
    check(sx_memcpy_htod(dx.ptr, x, bytes, nullptr), "sx_memcpy_htod");
    check(sx_memcpy_htod(dy.ptr, y, bytes, nullptr), "sx_memcpy_htod");
    check(sx_memcpy_htod(dz.ptr, z, bytes, nullptr), "sx_memcpy_htod");
    check(sx_index_build(h->index, static_cast<const float*>(dx.ptr), static_cast<const float*>(dy.ptr),
                         static_cast<const float*>(dz.ptr), static_cast<int64_t>(n), nullptr),
          "sx_index_build");
    check(sx_index_get_view(h->index, &h->view, nullptr), "sx_index_get_view");  // synchronises
    return h;
}
}  // namespace

}  // namespace shapes

// ---------------------------------------------------------------- RadixTree
RT::RadixTree::RadixTree(const PointCloud<float>& cloud, int root_mode) {
    const size_t n = cloud.x_vals.size();
    if (cloud.y_vals.size() != n || cloud.z_vals.size() != n) throw shapes::Error(SX_EINVAL, "ragged point cloud");
    index_ = shapes::build_index(cloud.x_vals.data(), cloud.y_vals.data(), cloud.z_vals.data(), n, root_mode);
    const sx_index_view& v = index_->view;
    d_tree.mortonCode = const_cast<Code_t*>(v.codes);
    d_tree.hasLeafLeft = reinterpret_cast<bool*>(const_cast<uint8_t*>(v.has_leaf_left));
    d_tree.hasLeafRight = reinterpret_cast<bool*>(const_cast<uint8_t*>(v.has_leaf_right));
    d_tree.prefixN = const_cast<uint8_t*>(v.prefix_n);
    d_tree.leftChild = const_cast<int*>(v.left_child);
    d_tree.parent = const_cast<int*>(v.parent);
    n_pts = static_cast<int>(v.n_unique);
    n_nodes = n_pts - 1;
    min_coord = v.min_coord;
    max_coord = v.max_coord;
    if (static_cast<int64_t>(n) != v.n_unique) std::printf("%lld duplicates removed\n", static_cast<long long>(n - v.n_unique));
}

RT::RadixTree::~RadixTree() = default;

// ---------------------------------------------------------------- Octree
OT::Octree::Octree(const RT::RadixTree& radix_tree) {
    index_ = radix_tree.index();
    const sx_index_view& v = index_->view;
    n_nodes = static_cast<int>(v.n_oct);
    u_nodes = reinterpret_cast<OTNode*>(const_cast<sx_otnode*>(v.oct_nodes));
    n_pts = static_cast<int>(v.n_unique);
    u_points = reinterpret_cast<Point*>(const_cast<float*>(v.points));
    h_points = std::make_shared<std::vector<Point>>(static_cast<size_t>(n_pts));
    shapes::check(sx_memcpy_dtoh(h_points->data(), u_points, sizeof(Point) * static_cast<size_t>(n_pts), nullptr),
                  "sx_memcpy_dtoh");
}

namespace {
// The raw grid coordinate whose decoded value (codeToPoint's arithmetic) is
// exactly v under the cube (mn, range), or -1 when v is not a decoded grid
// value.  The encode formula gives the candidate; decoding is monotone, so
// the search only looks a few steps either side of it.
int64_t snap_to_grid(float v, float mn, float range) {
    const float fs = static_cast<float>(axis_scale());
    float t = (v - mn) / range;
    float sc = fs * t;
    if (!(sc >= 0.0f)) sc = 0.0f;
    if (sc > fs) sc = fs;
    const int64_t r0 = static_cast<int64_t>(sc);
    for (int64_t d = 0; d <= 3; ++d) {
        for (int64_t r : {r0 - d, r0 + d}) {
            if (r < 0 || r > static_cast<int64_t>(axis_scale())) continue;
            const float u = static_cast<float>(static_cast<uint32_t>(r)) / fs;
            const float w = u * range;
            if (w + mn == v) return r;
        }
    }
    return -1;
}

Code_t interleave_raw(uint64_t rx, uint64_t ry, uint64_t rz) {
    Code_t code = 0;
    for (int i = 0; i < CODE_LEN / 3; ++i)
        code |= (((rx >> i) & 1u) << (3 * i)) | (((ry >> i) & 1u) << (3 * i + 1)) | (((rz >> i) & 1u) << (3 * i + 2));
    return code;
}
}  // namespace

OT::Octree::Octree(std::shared_ptr<std::vector<OTNode>> nodes, const int nn, std::shared_ptr<std::vector<Point>> points,
                   const int np) {
    if (!nodes || !points || nn < 0 || np <= 0 || static_cast<size_t>(nn) > nodes->size() ||
        static_cast<size_t>(np) > points->size())
        throw shapes::Error(SX_EINVAL, "Octree: bad host data");
    if (nn == 0 && np > 1) throw shapes::Error(SX_EINVAL, "Octree: host points without octree nodes");
    h_points = points;
    if (nn == 0) {
        // one point: no node carries a cube; index the point itself
        const Point& p = (*points)[0];
        index_ = shapes::build_index(&p.x, &p.y, &p.z, 1, SX_ROOT_TEXTBOOK);
    } else {
        // The root's cell is the whole cube: cell_size = range
        // (src/Octree.cu:257-261), and its corner is the cube minimum on the
        // axis where some point sits at raw coordinate 0 (every other corner
        // coordinate decodes a raw prefix >= 0, so it is not smaller).
        const OTNode& root = (*nodes)[0];
        const float mn = std::min(root.corner.x, std::min(root.corner.y, root.corner.z));
        const float range = root.cell_size;
        if (!(range > 0.0f) || !std::isfinite(range) || !std::isfinite(mn))
            throw shapes::Error(SX_EINVAL, "Octree: the root node does not describe a cube");
        // The points are the decoded unique points of that cube in Morton
        // order: recover their codes exactly and index those.
        std::vector<Code_t> codes(static_cast<size_t>(np));
        for (int i = 0; i < np; ++i) {
            const Point& p = (*points)[i];
            const int64_t rx = snap_to_grid(p.x, mn, range), ry = snap_to_grid(p.y, mn, range),
                          rz = snap_to_grid(p.z, mn, range);
            if (rx < 0 || ry < 0 || rz < 0)
                throw shapes::Error(SX_EINVAL, "Octree: host point " + std::to_string(i) +
                                                   " is not a decoded grid point of the octree's cube");
            codes[i] = interleave_raw(static_cast<uint64_t>(rx), static_cast<uint64_t>(ry), static_cast<uint64_t>(rz));
            if (i > 0 && codes[i] <= codes[i - 1])
                throw shapes::Error(SX_EINVAL, "Octree: host points are not distinct grid points in Morton order (point " +
                                                   std::to_string(i) + ")");
        }
        index_ = std::make_shared<shapes::IndexHandle>(static_cast<int64_t>(np), CODE_LEN, SX_ROOT_TEXTBOOK);
        DeviceBuffer dc(codes.size() * sizeof(Code_t));
        shapes::check(sx_memcpy_htod(dc.ptr, codes.data(), codes.size() * sizeof(Code_t), nullptr), "sx_memcpy_htod");
        shapes::check(sx_index_build_codes(index_->index, static_cast<const uint64_t*>(dc.ptr), np, mn, range, nullptr),
                      "sx_index_build_codes");
        shapes::check(sx_index_get_view(index_->index, &index_->view, nullptr), "sx_index_get_view");
    }
    const sx_index_view& v = index_->view;
    // The search runs on the index points; they must be the caller's points,
    // bit for bit and in the same order.
    std::vector<Point> decoded(static_cast<size_t>(v.n_unique));
    shapes::check(sx_memcpy_dtoh(decoded.data(), v.points, decoded.size() * sizeof(Point), nullptr), "sx_memcpy_dtoh");
    bool same = v.n_unique == np;
    for (int i = 0; same && i < np; ++i)
        same = std::memcmp(&decoded[i], &(*points)[i], sizeof(Point)) == 0;
    if (!same)
        throw shapes::Error(SX_EINVAL, "Octree: the indexed points differ from the host points");
    n_pts = np;
    u_points = reinterpret_cast<Point*>(const_cast<float*>(v.points));
    n_nodes = nn;
    shapes::check(sx_malloc(&own_nodes_, sizeof(OTNode) * static_cast<size_t>(std::max(nn, 1))), "sx_malloc");
    shapes::check(sx_memcpy_htod(own_nodes_, nodes->data(), sizeof(OTNode) * static_cast<size_t>(nn), nullptr),
                  "sx_memcpy_htod");
    shapes::check(sx_stream_sync(nullptr), "sx_stream_sync");
    u_nodes = static_cast<OTNode*>(own_nodes_);
}

void OT::Octree::release() {
    if (own_nodes_) sx_free(own_nodes_);
    own_nodes_ = nullptr;
    index_.reset();
    u_nodes = nullptr;
    u_points = nullptr;
    h_points.reset();
    n_nodes = n_pts = 0;
}

OT::Octree::~Octree() { release(); }

OT::Octree& OT::Octree::operator=(Octree&& other) {
    if (this != &other) {
        release();
        index_ = std::move(other.index_);
        own_nodes_ = other.own_nodes_;
        u_nodes = other.u_nodes;
        u_points = other.u_points;
        h_points = std::move(other.h_points);
        n_nodes = other.n_nodes;
        n_pts = other.n_pts;
        other.own_nodes_ = nullptr;
        other.u_nodes = nullptr;
        other.u_points = nullptr;
        other.n_nodes = other.n_pts = 0;
    }
    return *this;
}

void OT::Octree::knn(const Point* d_query, int* d_nn, int N, int k) const {
    if (!index_) throw shapes::Error(SX_ESTATE, "Octree: empty octree");
    if (N <= 0) return;
    sx_index* ix = index_->index;
    // Queries that are the octree's own points use the index path (same
    // results; the query is its own first neighbour); others the generic one.
    const Point* base = u_points;
    if (d_query >= base && d_query + N <= base + n_pts) {
        shapes::check(sx_knn(ix, d_query - base, N, k, SX_KNN_PLAIN, nullptr, d_nn, nullptr), "sx_knn");
    } else {
        shapes::check(sx_knn_points(ix, reinterpret_cast<const float*>(d_query), N, k, d_nn, nullptr),
                      "sx_knn_points");
    }
    shapes::check(sx_stream_sync(nullptr), "sx_stream_sync");
}

std::vector<int> OT::Octree::knn_host(const std::vector<Point>& points, int k) const {
    std::vector<int> out(points.size() * static_cast<size_t>(k));
    if (points.empty()) return out;
    DeviceBuffer dq(points.size() * sizeof(Point)), dn(out.size() * sizeof(int));
    shapes::check(sx_memcpy_htod(dq.ptr, points.data(), points.size() * sizeof(Point), nullptr), "sx_memcpy_htod");
    knn(static_cast<const Point*>(dq.ptr), static_cast<int*>(dn.ptr), static_cast<int>(points.size()), k);
    shapes::check(sx_memcpy_dtoh(out.data(), dn.ptr, out.size() * sizeof(int), nullptr), "sx_memcpy_dtoh");
    return out;
}

// ---------------------------------------------------------------- normals
namespace shapes {

RadiusModel::RadiusModel(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(SX_EINVAL, "cannot open radius model " + path);
    char magic[8];
    int32_t hdr[3];
    f.read(magic, 8);
    f.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!f || std::memcmp(magic, "SXRADIUS", 8) != 0 || hdr[0] != 1 || hdr[2] != sx_mlp_num_params())
        throw Error(SX_EINVAL, path + ": not a version-1 radius model");
    k_ = hdr[1];
    std::vector<float> p(static_cast<size_t>(hdr[2]));
    float mu[8], sigma[8], margin;
    f.read(reinterpret_cast<char*>(p.data()), static_cast<std::streamsize>(p.size() * sizeof(float)));
    f.read(reinterpret_cast<char*>(mu), sizeof(mu));
    f.read(reinterpret_cast<char*>(sigma), sizeof(sigma));
    f.read(reinterpret_cast<char*>(&margin), sizeof(margin));
    if (!f) throw Error(SX_EINVAL, path + ": truncated radius model");
    check(sx_mlp_create(&mlp_, 1024), "sx_mlp_create");
    check(sx_mlp_set_params(mlp_, p.data(), nullptr), "sx_mlp_set_params");
    check(sx_mlp_set_normalizer(mlp_, mu, sigma, margin, nullptr), "sx_mlp_set_normalizer");
}

RadiusModel::~RadiusModel() { sx_mlp_destroy(mlp_); }

namespace {
// kNN + normals of the octree points [start_idx, start_idx + n) into the
// device buffer `normals` (n rows of 12 bytes).
void normals_on_device(const OT::Octree& octree, int start_idx, int n, int k, const NormalOptions& opt,
                       DeviceBuffer& normals) {
    sx_index* ix = octree.index()->index;
    DeviceBuffer nn(sizeof(int32_t) * static_cast<size_t>(n) * k);
    int mode = SX_KNN_PLAIN;
    std::unique_ptr<DeviceBuffer> feats, r2;
    if (opt.radius_model) {
        feats.reset(new DeviceBuffer(sizeof(float) * 8 * static_cast<size_t>(n)));
        r2.reset(new DeviceBuffer(sizeof(float) * static_cast<size_t>(n)));
        check(sx_radius_features(ix, start_idx, n, static_cast<float*>(feats->ptr), nullptr), "sx_radius_features");
        check(sx_mlp_predict_radius2(opt.radius_model->handle(), static_cast<const float*>(feats->ptr), n,
                                     static_cast<float*>(r2->ptr), nullptr),
              "sx_mlp_predict_radius2");
        mode = SX_KNN_RADIUS;
    }
    check(sx_knn(ix, start_idx, n, k, mode, r2 ? static_cast<const float*>(r2->ptr) : nullptr,
                 static_cast<int32_t*>(nn.ptr), nullptr),
          "sx_knn");
    check(sx_normals(ix, static_cast<const int32_t*>(nn.ptr), start_idx, n, k, opt.precision, opt.viewpoint.x,
                     opt.viewpoint.y, opt.viewpoint.z, static_cast<float*>(normals.ptr), nullptr, nullptr),
          "sx_normals");
}
}  // namespace

std::vector<Point> estimate_normals(const OT::Octree& octree, int start_idx, int n, int k, const NormalOptions& opt) {
    if (!octree.index()) throw Error(SX_ESTATE, "estimateNormals: empty octree");
    if (start_idx < 0 || n < 0 || start_idx + n > octree.n_pts)
        throw Error(SX_EINVAL, "estimateNormals: range outside the octree points");
    std::vector<Point> out(static_cast<size_t>(n));
    if (n == 0) return out;
    DeviceBuffer normals(sizeof(Point) * static_cast<size_t>(n));
    normals_on_device(octree, start_idx, n, k, opt, normals);
    check(sx_memcpy_dtoh(out.data(), normals.ptr, sizeof(Point) * out.size(), nullptr), "sx_memcpy_dtoh");
    return out;
}

std::vector<Point> estimate_normals_input_order(const OT::Octree& octree, int k, const NormalOptions& opt) {
    if (!octree.index()) throw Error(SX_ESTATE, "estimateNormals: empty octree");
    const sx_index_view& v = octree.index()->view;
    std::vector<Point> out(static_cast<size_t>(v.n_input));
    DeviceBuffer normals(sizeof(Point) * static_cast<size_t>(octree.n_pts));
    normals_on_device(octree, 0, octree.n_pts, k, opt, normals);
    DeviceBuffer in_order(sizeof(Point) * out.size());
    check(sx_index_to_input_order(octree.index()->index, normals.ptr, sizeof(Point), in_order.ptr, nullptr),
          "sx_index_to_input_order");
    check(sx_memcpy_dtoh(out.data(), in_order.ptr, sizeof(Point) * out.size(), nullptr), "sx_memcpy_dtoh");
    return out;
}

}  // namespace shapes

// ---------------------------------------------------------------- DataIO
PointCloud<float> DataIO::loadFile(const std::string& filepath) {
    auto ends_with = [&](const std::string& ext) {
        return filepath.size() >= ext.size() && filepath.compare(filepath.size() - ext.size(), ext.size(), ext) == 0;
    };
    if (ends_with(".obj")) return loadObj(filepath);
    if (ends_with(".bin")) {
        std::ifstream f(filepath, std::ios::binary | std::ios::ate);
        if (!f) throw shapes::Error(SX_EINVAL, "cannot open " + filepath);
        return loadKitti(filepath, static_cast<size_t>(f.tellg()) / sizeof(float));
    }
    if (ends_with(".txt")) return loadSemantic3D(filepath);
    throw shapes::Error(SX_EINVAL, "unsupported point-cloud file extension: " + filepath);
}

PointCloud<float> DataIO::loadKitti(const std::string& filepath, size_t num) {
    PointCloud<float> cloud;
    std::ifstream f(filepath, std::ios::binary);
    if (!f) throw shapes::Error(SX_EINVAL, "cannot open " + filepath);
    std::vector<float> data(num);
    f.read(reinterpret_cast<char*>(data.data()), static_cast<std::streamsize>(num * sizeof(float)));
    const size_t got = static_cast<size_t>(f.gcount()) / sizeof(float) / 4;
    cloud.x_vals.resize(got);
    cloud.y_vals.resize(got);
    cloud.z_vals.resize(got);
    for (size_t i = 0; i < got; ++i) {
        cloud.x_vals[i] = data[4 * i];
        cloud.y_vals[i] = data[4 * i + 1];
        cloud.z_vals[i] = data[4 * i + 2];
    }
    return cloud;
}

void DataIO::saveKitti(const PointCloud<float> cloud, const std::string& filepath) {
    std::ofstream f(filepath, std::ios::binary);
    if (!f) throw shapes::Error(SX_EINVAL, "cannot write " + filepath);
    const float reflectance = 255.0f;
    std::vector<float> rec(4 * cloud.x_vals.size());
    for (size_t i = 0; i < cloud.x_vals.size(); ++i) {
        rec[4 * i] = cloud.x_vals[i];
        rec[4 * i + 1] = cloud.y_vals[i];
        rec[4 * i + 2] = cloud.z_vals[i];
        rec[4 * i + 3] = reflectance;
    }
    f.write(reinterpret_cast<const char*>(rec.data()), static_cast<std::streamsize>(rec.size() * sizeof(float)));
}

PointCloud<float> DataIO::loadObj(const std::string& filepath) {
    PointCloud<float> cloud;
    std::ifstream file(filepath);
    if (!file) throw shapes::Error(SX_EINVAL, "cannot open " + filepath);
    std::string line;
    while (std::getline(file, line)) {
        if (line.size() > 1 && line[0] == 'v' && line[1] == ' ') {
            std::istringstream iss(line.substr(2));
            float x, y, z;
            if (iss >> x >> y >> z) {
                cloud.x_vals.push_back(x);
                cloud.y_vals.push_back(y);
                cloud.z_vals.push_back(z);
            } else {
                std::fprintf(stderr, "Failed to parse vertex line %s\n", line.c_str());
            }
        }
    }
    return cloud;
}

PointCloud<float> DataIO::loadSemantic3D(const std::string& filepath) {
    PointCloud<float> cloud;
    std::ifstream file(filepath);
    if (!file) throw shapes::Error(SX_EINVAL, "cannot open " + filepath);
    std::string line;
    while (std::getline(file, line)) {
        std::istringstream iss(line);
        float x, y, z, r, g, b;
        if (iss >> x >> y >> z >> r >> g >> b) {
            cloud.x_vals.push_back(x);
            cloud.y_vals.push_back(y);
            cloud.z_vals.push_back(z);
        } else {
            std::fprintf(stderr, "Failed to parse line %s\n", line.c_str());
        }
    }
    return cloud;
}
